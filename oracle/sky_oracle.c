/*
 * CPU oracle for the sky-model visibility predictor.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path. Only tests,
 * the smoke check and the benchmark's CPU-baseline leg load this library.
 *
 * Every function restates the reference's algorithm in the reference's operation
 * order (Rust never contracts a*b+c into an FMA, so this file is compiled with
 * -ffp-contract=off), citing the file:line it follows. Parity status of each part
 * is recorded next to it:
 *   - geometry (LMN, UVW, HA/Dec): pinned bit-exactly by the reference's golden
 *     vectors (src/model/tests/mod.rs:370-444,1530-1568; components/tests.rs:80-82).
 *   - flux estimation: pinned by src/srclist/types/flux_density/tests.rs KATs and
 *     src/model/tests/gpu.rs:598-600.
 *   - point / Gaussian / shapelet folds and the f32 demotion: pinned bit-exactly by
 *     the no-beam golden arrays (src/model/tests/mod.rs:350-1600).
 *   - Gaussian envelope with nonzero size: pinned to f32 precision by the ratio of the
 *     FEE golden arrays (mod.rs:1177-1390 / 524-597, 810-885, 1097-1171).
 *   - shapelet envelope with nonzero size or n > 0: pinned by a brute-force Hermite
 *     evaluation and interpolation bounds in the tests (no reference vector exists).
 *   - az/el (marlu HADec::to_azel): parity unpinned (only feeds the beam, whose reference
 *     implementation, mwa_hyperbeam 0.11.0, is absent); checked against closed-form cases.
 *   - analytic beam: this project's own beam model, no reference counterpart; checked
 *     against closed-form special cases (zenith, dead dipoles, steering).
 */
#include "sky_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define VEL_C 299792458.0
#define TAU 6.283185307179586
#define FRAC_PI_2 1.5707963267948966
#define LN_2 0.6931471805599453
#define SQRT_FRAC_PI_SQ_2_LN_2 2.6682231283184983 /* src/constants.rs:23 */
#define DEFAULT_SPEC_INDEX (-0.8)                 /* src/constants.rs:8 */
#define SBF_C 5000.0                              /* src/model/shapelets.rs:11-14 */
#define SBF_L 10001
#define SBF_N 101
#define SBF_DX 0.01
#define POWER_LAW_FD_REF_FREQ 150e6 /* src/gpu/model.h:9 */

/* ------------------------------------------------------------------ complex / Jones */
typedef struct {
    double re, im;
} c64;
typedef struct {
    c64 j[4]; /* [XX, XY, YX, YY] = [(0,0), (0,1), (1,0), (1,1)] */
} jones;

/* num-complex 0.4.6 Mul/Add (no FMA). */
static inline c64 cmul(c64 a, c64 b) {
    c64 r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}
static inline c64 cadd(c64 a, c64 b) {
    c64 r = {a.re + b.re, a.im + b.im};
    return r;
}
static inline c64 cscale(c64 a, double s) {
    c64 r = {a.re * s, a.im * s};
    return r;
}
static inline c64 cconj(c64 a) {
    c64 r = {a.re, -a.im};
    return r;
}
/* marlu Jones * Jones: standard 2x2 product, element order [a0 b0 + a1 b2, ...]. */
static inline jones jmul(jones a, jones b) {
    jones r;
    r.j[0] = cadd(cmul(a.j[0], b.j[0]), cmul(a.j[1], b.j[2]));
    r.j[1] = cadd(cmul(a.j[0], b.j[1]), cmul(a.j[1], b.j[3]));
    r.j[2] = cadd(cmul(a.j[2], b.j[0]), cmul(a.j[3], b.j[2]));
    r.j[3] = cadd(cmul(a.j[2], b.j[1]), cmul(a.j[3], b.j[3]));
    return r;
}
/* marlu Jones::h(): conjugate transpose. */
static inline jones jh(jones a) {
    jones r;
    r.j[0] = cconj(a.j[0]);
    r.j[1] = cconj(a.j[2]);
    r.j[2] = cconj(a.j[1]);
    r.j[3] = cconj(a.j[3]);
    return r;
}
static inline jones jmulc(jones a, c64 s) {
    jones r;
    for (int k = 0; k < 4; k++) r.j[k] = cmul(a.j[k], s);
    return r;
}
static inline jones jmuls(jones a, double s) {
    jones r;
    for (int k = 0; k < 4; k++) r.j[k] = cscale(a.j[k], s);
    return r;
}
static inline void jadd_inplace(jones *a, jones b) {
    for (int k = 0; k < 4; k++) a->j[k] = cadd(a->j[k], b.j[k]);
}
static const jones JONES_IDENTITY = {{{1.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {1.0, 0.0}}};
static const jones JONES_ZERO = {{{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}}};

/* ------------------------------------------------------------------ geometry */
/* marlu RADec::to_lmn + LMN::prepare_for_rime (called at src/srclist/types/components/mod.rs:214). */
void orc_radec_to_lmn_rime(double ra, double dec, double pc_ra, double pc_dec, double out[3]) {
    double d_ra = ra - pc_ra;
    double s_d_ra = sin(d_ra), c_d_ra = cos(d_ra);
    double s_dec = sin(dec), c_dec = cos(dec);
    double pc_s_dec = sin(pc_dec), pc_c_dec = cos(pc_dec);
    double l = c_dec * s_d_ra;
    double m = s_dec * pc_c_dec - c_dec * pc_s_dec * c_d_ra;
    double n = s_dec * pc_s_dec + c_dec * pc_c_dec * c_d_ra;
    out[0] = TAU * l;
    out[1] = TAU * m;
    out[2] = TAU * (n - 1.0);
}

/* marlu xyzs_to_cross_uvws (src/model/cpu.rs:767; formula mdbook/src/defs/coords.md "UVWs"):
 * per-tile UVW, then baseline = tile1 - tile2 over the upper triangle. */
void orc_xyzs_to_cross_uvws(int32_t n, const double *xyzs, double ha, double dec, double *out) {
    double s_ha = sin(ha), c_ha = cos(ha);
    double s_dec = sin(dec), c_dec = cos(dec);
    double *tile = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    for (int32_t i = 0; i < n; i++) {
        double x = xyzs[3 * i], y = xyzs[3 * i + 1], z = xyzs[3 * i + 2];
        tile[3 * i + 0] = s_ha * x + c_ha * y;
        tile[3 * i + 1] = -s_dec * c_ha * x + s_dec * s_ha * y + c_dec * z;
        tile[3 * i + 2] = c_dec * c_ha * x - c_dec * s_ha * y + s_dec * z;
    }
    size_t bl = 0;
    for (int32_t i = 0; i < n; i++) {
        for (int32_t j = i + 1; j < n; j++) {
            out[3 * bl + 0] = tile[3 * i + 0] - tile[3 * j + 0];
            out[3 * bl + 1] = tile[3 * i + 1] - tile[3 * j + 1];
            out[3 * bl + 2] = tile[3 * i + 2] - tile[3 * j + 2];
            bl++;
        }
    }
    free(tile);
}

/* marlu HADec::to_azel (ERFA hd2ae formulation), called at components/mod.rs:394. */
void orc_hadec_to_azel(double ha, double dec, double lat, double *az, double *el) {
    double s_ha = sin(ha), c_ha = cos(ha);
    double s_dec = sin(dec), c_dec = cos(dec);
    double s_lat = sin(lat), c_lat = cos(lat);
    double x = -c_ha * c_dec * s_lat + s_dec * c_lat;
    double y = -s_ha * c_dec;
    double z = c_ha * c_dec * c_lat + s_dec * s_lat;
    double r = hypot(x, y);
    double a = (r != 0.0) ? atan2(y, x) : 0.0;
    *az = (a < 0.0) ? a + TAU : a;
    *el = atan2(z, r);
}

/* TileBaselineFlags ordering (src/math/mod.rs:96-109): row-major upper triangle. */
void orc_baseline_to_tiles(int32_t num_tiles, int32_t bl, int32_t *t1, int32_t *t2) {
    int32_t b = bl;
    for (int32_t i = 0; i < num_tiles; i++) {
        int32_t row = num_tiles - 1 - i;
        if (b < row) {
            *t1 = i;
            *t2 = i + 1 + b;
            return;
        }
        b -= row;
    }
    *t1 = -1;
    *t2 = -1;
}

/* ------------------------------------------------------------------ flux densities */
static double calc_flux_ratio(double desired, double cat, double si) { /* flux_density/mod.rs:273-275 */
    return pow(desired / cat, si);
}
static void fd_scale(OrcFluxDensity *fd, double s) { /* FluxDensity * f64 (mod.rs:103-115) */
    fd->i *= s;
    fd->q *= s;
    fd->u *= s;
    fd->v *= s;
}

/* FluxDensityType::estimate_at_freq (src/srclist/types/flux_density/mod.rs:158-270). */
void orc_estimate_at_freq(const OrcComponent *c, double freq_hz, OrcFluxDensity *out) {
    if (c->flux_type == ORC_POWER_LAW) {
        double ratio = calc_flux_ratio(freq_hz, c->fds[0].freq, c->si);
        *out = c->fds[0];
        fd_scale(out, ratio);
        out->freq = freq_hz;
        return;
    }
    if (c->flux_type == ORC_CURVED_POWER_LAW) {
        OrcFluxDensity pl = c->fds[0];
        fd_scale(&pl, calc_flux_ratio(freq_hz, c->fds[0].freq, c->si));
        pl.freq = freq_hz;
        double lr = log(freq_hz / c->fds[0].freq);
        double curved = exp(c->q * (lr * lr));
        fd_scale(&pl, curved);
        *out = pl;
        return;
    }
    /* List. */
    const OrcFluxDensity *fds = c->fds;
    int32_t n = c->num_fds;
    double spec_index;
    const OrcFluxDensity *smaller;
    if (n == 1) {
        spec_index = DEFAULT_SPEC_INDEX;
        smaller = &fds[0];
    } else {
        const OrcFluxDensity *p0 = &fds[0], *p1 = &fds[1];
        for (int32_t w = 0; w + 1 < n; w++) {
            p0 = &fds[w];
            p1 = &fds[w + 1];
            if (fabs(p0->freq - freq_hz) < 1e-3) {
                *out = *p0;
                return;
            }
            if (fabs(p1->freq - freq_hz) < 1e-3) {
                *out = *p1;
                return;
            }
            if (freq_hz < p1->freq) break;
        }
        /* f64::signum: +1 for +0.0 and positives, -1 for -0.0 and negatives. */
        double s0 = signbit(p0->i) ? -1.0 : 1.0;
        double s1 = signbit(p1->i) ? -1.0 : 1.0;
        if (s0 != s1) {
            double run = p1->freq - p0->freq;
            double i_slope = (p1->i - p0->i) / run;
            double q_slope = (p1->q - p0->q) / run;
            double u_slope = (p1->u - p0->u) / run;
            double v_slope = (p1->v - p0->v) / run;
            out->freq = freq_hz;
            out->i = p0->i + i_slope * (freq_hz - p0->freq);
            out->q = p0->q + q_slope * (freq_hz - p0->freq);
            out->u = p0->u + u_slope * (freq_hz - p0->freq);
            out->v = p0->v + v_slope * (freq_hz - p0->freq);
            return;
        }
        spec_index = log(p1->i / p0->i) / log(p1->freq / p0->freq); /* calc_spec_index, mod.rs:46-48 */
        smaller = (p1->freq < freq_hz) ? p1 : p0;
    }
    double ratio = calc_flux_ratio(freq_hz, smaller->freq, spec_index);
    *out = *smaller;
    out->freq = freq_hz;
    fd_scale(out, ratio);
}

/* FluxDensity::to_inst_stokes (flux_density/mod.rs:58-65). */
static jones to_inst_stokes(OrcFluxDensity fd) {
    jones j;
    j.j[0].re = fd.i - fd.q;
    j.j[0].im = 0.0;
    j.j[1].re = fd.u;
    j.j[1].im = -fd.v;
    j.j[2].re = fd.u;
    j.j[2].im = fd.v;
    j.j[3].re = fd.i + fd.q;
    j.j[3].im = 0.0;
    return j;
}

/* CPL spectral index re-derived for the 150 MHz reference (src/model/gpu.rs:664-680). */
double orc_cpl_si_at_150mhz(const OrcComponent *c) {
    if (c->fds[0].freq == POWER_LAW_FD_REF_FREQ) return c->si;
    OrcFluxDensity at150;
    orc_estimate_at_freq(c, POWER_LAW_FD_REF_FREQ, &at150);
    double logratio = log(c->fds[0].freq / POWER_LAW_FD_REF_FREQ);
    return (log(c->fds[0].i / at150.i) - c->q * (logratio * logratio)) / logratio;
}

/* ------------------------------------------------------------------ shapelet basis table */
/* Hermite functions B_n(x) = exp(-x^2/2) H_n(x) / sqrt(2^n n!) on x_j = j*0.01 - 50
 * (src/model/make_shapelet_basis_values.py:12-34, numpy.linspace(-50, 50, 10001)),
 * evaluated with the normalised three-term recurrence
 * B_{n+1} = sqrt(2/(n+1)) x B_n - sqrt(n/(n+1)) B_{n-1}. */
void orc_shapelet_basis_table(double *out) {
    const double step = 100.0 / 10000.0;
    for (int j = 0; j < SBF_L; j++) {
        double x = (j == SBF_L - 1) ? 50.0 : (double)j * step + -50.0;
        double b0 = exp(-0.5 * (x * x));
        double bm1 = 0.0, b = b0;
        out[j] = b0;
        for (int n = 0; n + 1 < SBF_N; n++) {
            double nb = sqrt(2.0 / (double)(n + 1)) * x * b - sqrt((double)n / (double)(n + 1)) * bm1;
            bm1 = b;
            b = nb;
            out[(size_t)(n + 1) * SBF_L + j] = b;
        }
    }
}

static double *g_table = NULL;
const double *orc_shapelet_basis_table_cached(void) {
    if (!g_table) {
        double *t = (double *)malloc(sizeof(double) * SBF_L * SBF_N);
        orc_shapelet_basis_table(t);
        g_table = t;
    }
    return g_table;
}

/* ------------------------------------------------------------------ analytic MWA-like beam */
/* The product's analytic beam, restated on the CPU. A 4x4 grid of crossed short dipoles
 * (1.1 m pitch) over a ground plane (dipole height 0.278 m), steered by delays in 435 ps
 * steps, per-dipole gains (X then Y), delay 32 = dead dipole (src/model/cpu.rs:944-964).
 * Rows are the X (east-west) and Y (north-south) dipoles; columns are the components of
 * the sky field along a = normalised projection of East onto the plane of the sky and
 * b = r x a, so the Jones matrix is the identity at zenith. Normalised by the zenith
 * response of an ideal tile. */
#define BEAM_DIPOLE_SEP 1.1
#define BEAM_DIPOLE_HEIGHT 0.278
#define BEAM_DELAY_STEP 435e-12
void orc_analytic_beam_jones(double az, double za, double freq_hz, const uint32_t delays[16],
                             const double gains[32], double out[8]) {
    double s_za = sin(za), c_za = cos(za), s_az = sin(az), c_az = cos(az);
    double r_e = s_za * s_az, r_n = s_za * c_az, r_u = c_za;
    double k = TAU * freq_hz / VEL_C;
    double afx_re = 0, afx_im = 0, afy_re = 0, afy_im = 0;
    for (int d = 0; d < 16; d++) {
        uint32_t delay = delays ? delays[d] : 0;
        double gx = gains ? gains[d] : 1.0;
        double gy = gains ? gains[16 + d] : 1.0;
        if (delay == 32) {
            gx = 0.0;
            gy = 0.0;
            delay = 0;
        }
        double x = ((double)(d % 4) - 1.5) * BEAM_DIPOLE_SEP;
        double y = (1.5 - (double)(d / 4)) * BEAM_DIPOLE_SEP;
        double ph = k * (x * r_e + y * r_n) - TAU * freq_hz * ((double)delay * BEAM_DELAY_STEP);
        double s = sin(ph), c = cos(ph);
        afx_re += gx * c;
        afx_im += gx * s;
        afy_re += gy * c;
        afy_im += gy * s;
    }
    double g = sin(k * BEAM_DIPOLE_HEIGHT * r_u) / (16.0 * sin(k * BEAM_DIPOLE_HEIGHT));
    double sx = sqrt(1.0 - r_e * r_e);
    if (sx < 1e-12) sx = 1e-12;
    double px = g * sx;
    double pyx = g * (-r_e * r_n / sx);
    double pyy = g * (r_u / sx);
    out[0] = afx_re * px;
    out[1] = afx_im * px;
    out[2] = 0.0;
    out[3] = 0.0;
    out[4] = afy_re * pyx;
    out[5] = afy_im * pyx;
    out[6] = afy_re * pyy;
    out[7] = afy_im * pyy;
}

/* ------------------------------------------------------------------ modeller set-up */
/* Unique tiles by (delays, gains with delay 32 -> gain 0) (src/model/cpu.rs:105-143,944-964). */
int32_t orc_unique_tiles(int32_t total_tiles, const int32_t *flagged, int32_t num_flagged,
                         const uint32_t *delays, const double *gains, int32_t *tile_map) {
    int32_t *uniq = (int32_t *)malloc(sizeof(int32_t) * (size_t)(total_tiles > 0 ? total_tiles : 1));
    int32_t n_uniq = 0;
    for (int32_t t = 0; t < total_tiles; t++) {
        int is_flagged = 0;
        for (int32_t f = 0; f < num_flagged; f++)
            if (flagged[f] == t) is_flagged = 1;
        if (is_flagged) {
            tile_map[t] = 0;
            continue;
        }
        double amps_t[32];
        uint32_t del_t[16];
        for (int k = 0; k < 16; k++) del_t[k] = delays ? delays[16 * t + k] : 0;
        for (int k = 0; k < 32; k++) {
            double a = gains ? gains[32 * t + k] : 1.0;
            amps_t[k] = (del_t[k % 16] == 32) ? 0.0 : a;
        }
        int32_t found = -1;
        for (int32_t u = 0; u < n_uniq && found < 0; u++) {
            int32_t o = uniq[u];
            double amps_o[32];
            uint32_t del_o[16];
            for (int k = 0; k < 16; k++) del_o[k] = delays ? delays[16 * o + k] : 0;
            for (int k = 0; k < 32; k++) {
                double a = gains ? gains[32 * o + k] : 1.0;
                amps_o[k] = (del_o[k % 16] == 32) ? 0.0 : a;
            }
            if (memcmp(del_t, del_o, sizeof del_t) == 0 && memcmp(amps_t, amps_o, sizeof amps_t) == 0) found = u;
        }
        if (found < 0) {
            uniq[n_uniq] = t;
            found = n_uniq++;
        }
        tile_map[t] = found;
    }
    free(uniq);
    return n_uniq;
}

typedef struct {
    int32_t n;           /* components of this type */
    const OrcComponent **c; /* in modelling order */
    double *lmn;         /* [n][3] */
    jones *fds;          /* [num_freqs][n] instrumental FDs */
    double *shapelet_uvw; /* [num_bl][n][3], shapelets only */
} type_list;

typedef struct {
    const OrcObs *obs;
    int32_t total_tiles;
    int32_t num_bl;
    int32_t *tile_map;     /* total-array tile index -> unique tile */
    int32_t n_unique_tiles;
    int32_t *unique_tiles; /* unique tile -> total-array tile index */
    int32_t *unflagged_to_total; /* unflagged tile index -> total-array tile index */
    int32_t *freq_map;     /* freq index -> unique freq index */
    double *unique_freqs;
    int32_t n_unique_freqs;
    double *uvws;          /* [num_bl][3] */
    int32_t *bl_t1, *bl_t2; /* unflagged-tile indices per baseline */
    type_list types[3];
    int32_t f_start, f_count; /* the channels being modelled */
} modeller;

static int is_flagged(const OrcObs *obs, int32_t t) {
    for (int32_t f = 0; f < obs->num_flagged_tiles; f++)
        if (obs->flagged_tiles[f] == t) return 1;
    return 0;
}

/* ComponentList::new over the reversed source list (src/model/cpu.rs:86-93,
 * src/srclist/types/components/mod.rs:181-291). */
static void build_modeller(modeller *m, const OrcObs *obs, const OrcComponent *comps, const int32_t *src_off,
                           int32_t num_sources, int32_t f_start, int32_t f_count) {
    memset(m, 0, sizeof *m);
    m->obs = obs;
    m->f_start = f_start;
    m->f_count = f_count;
    int32_t nt = obs->num_unflagged_tiles;
    m->total_tiles = nt + obs->num_flagged_tiles;
    m->num_bl = nt * (nt - 1) / 2;
    m->tile_map = (int32_t *)calloc((size_t)m->total_tiles + 1, sizeof(int32_t));
    m->n_unique_tiles = orc_unique_tiles(m->total_tiles, obs->flagged_tiles, obs->num_flagged_tiles,
                                         obs->dipole_delays, obs->dipole_gains, m->tile_map);
    m->unique_tiles = (int32_t *)calloc((size_t)m->total_tiles + 1, sizeof(int32_t));
    {
        int32_t next = 0;
        for (int32_t t = 0; t < m->total_tiles; t++) {
            if (is_flagged(obs, t)) continue;
            if (m->tile_map[t] == next) m->unique_tiles[next++] = t;
        }
    }
    m->unflagged_to_total = (int32_t *)calloc((size_t)nt + 1, sizeof(int32_t));
    {
        int32_t u = 0;
        for (int32_t t = 0; t < m->total_tiles; t++)
            if (!is_flagged(obs, t)) m->unflagged_to_total[u++] = t;
    }
    /* Unique beam freqs: the analytic beam and NoBeam are defined at every frequency,
     * so find_closest_freq is the identity (src/beam/mod.rs:319-321). */
    m->freq_map = (int32_t *)calloc((size_t)obs->num_freqs + 1, sizeof(int32_t));
    m->unique_freqs = (double *)calloc((size_t)obs->num_freqs + 1, sizeof(double));
    for (int32_t f = 0; f < obs->num_freqs; f++) {
        int32_t idx = -1;
        for (int32_t u = 0; u < m->n_unique_freqs; u++)
            if (m->unique_freqs[u] == obs->freqs[f]) idx = u;
        if (idx < 0) {
            idx = m->n_unique_freqs;
            m->unique_freqs[m->n_unique_freqs++] = obs->freqs[f];
        }
        m->freq_map[f] = idx;
    }
    m->uvws = (double *)malloc(sizeof(double) * 3 * (size_t)(m->num_bl > 0 ? m->num_bl : 1));
    orc_xyzs_to_cross_uvws(nt, obs->uvw_xyzs != NULL ? obs->uvw_xyzs : obs->unflagged_xyzs,
                           obs->lst - obs->phase_centre_ra, obs->phase_centre_dec, m->uvws);
    m->bl_t1 = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m->num_bl > 0 ? m->num_bl : 1));
    m->bl_t2 = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m->num_bl > 0 ? m->num_bl : 1));
    {
        size_t b = 0;
        for (int32_t i = 0; i < nt; i++)
            for (int32_t j = i + 1; j < nt; j++) {
                m->bl_t1[b] = i;
                m->bl_t2[b] = j;
                b++;
            }
    }
    int32_t total_comps = src_off[num_sources];
    for (int ty = 0; ty < 3; ty++) {
        type_list *tl = &m->types[ty];
        tl->c = (const OrcComponent **)malloc(sizeof(void *) * (size_t)(total_comps + 1));
        tl->n = 0;
    }
    for (int32_t s = num_sources - 1; s >= 0; s--) {
        for (int32_t i = src_off[s]; i < src_off[s + 1]; i++) {
            type_list *tl = &m->types[comps[i].comp_type];
            tl->c[tl->n++] = &comps[i];
        }
    }
    for (int ty = 0; ty < 3; ty++) {
        type_list *tl = &m->types[ty];
        int32_t n = tl->n;
        tl->lmn = (double *)malloc(sizeof(double) * 3 * (size_t)(n + 1));
        tl->fds = (jones *)malloc(sizeof(jones) * (size_t)(n + 1) * (size_t)(f_count + 1));
        for (int32_t i = 0; i < n; i++) {
            orc_radec_to_lmn_rime(tl->c[i]->ra, tl->c[i]->dec, obs->phase_centre_ra, obs->phase_centre_dec,
                                  &tl->lmn[3 * i]);
            for (int32_t f = 0; f < f_count; f++) {
                OrcFluxDensity fd;
                orc_estimate_at_freq(tl->c[i], obs->freqs[f_start + f], &fd);
                tl->fds[(size_t)f * n + i] = to_inst_stokes(fd);
            }
        }
        if (ty == ORC_SHAPELET && n > 0) {
            /* ShapeletComponentParams::get_shapelet_uvws (components/mod.rs:299-315):
             * each component's own HA/Dec with the unprecessed tile XYZs. */
            tl->shapelet_uvw = (double *)malloc(sizeof(double) * 3 * (size_t)n * (size_t)(m->num_bl + 1));
            double *tmp = (double *)malloc(sizeof(double) * 3 * (size_t)(m->num_bl + 1));
            for (int32_t i = 0; i < n; i++) {
                orc_xyzs_to_cross_uvws(nt, obs->unflagged_xyzs, obs->lst - tl->c[i]->ra, tl->c[i]->dec, tmp);
                for (int32_t b = 0; b < m->num_bl; b++) {
                    for (int k = 0; k < 3; k++) tl->shapelet_uvw[((size_t)b * n + i) * 3 + k] = tmp[3 * b + k];
                }
            }
            free(tmp);
        }
    }
}

static void free_modeller(modeller *m) {
    free(m->tile_map);
    free(m->unique_tiles);
    free(m->unflagged_to_total);
    free(m->freq_map);
    free(m->unique_freqs);
    free(m->uvws);
    free(m->bl_t1);
    free(m->bl_t2);
    for (int ty = 0; ty < 3; ty++) {
        free(m->types[ty].c);
        free(m->types[ty].lmn);
        free(m->types[ty].fds);
        free(m->types[ty].shapelet_uvw);
    }
}

/* get_beam_responses (src/model/cpu.rs:187-230): [unique tile][unique freq][direction]. */
static jones *beam_responses(const modeller *m, const type_list *tl) {
    const OrcObs *obs = m->obs;
    size_t n = (size_t)tl->n;
    jones *out = (jones *)malloc(sizeof(jones) * (size_t)m->n_unique_tiles * (size_t)m->n_unique_freqs * (n + 1));
    if (obs->beam_type == ORC_BEAM_NONE) {
        for (size_t k = 0; k < (size_t)m->n_unique_tiles * (size_t)m->n_unique_freqs * n; k++) out[k] = JONES_IDENTITY;
        return out;
    }
    double *az = (double *)malloc(sizeof(double) * (n + 1));
    double *za = (double *)malloc(sizeof(double) * (n + 1));
    for (size_t i = 0; i < n; i++) {
        double a, e;
        orc_hadec_to_azel(obs->lst - tl->c[i]->ra, tl->c[i]->dec, obs->latitude, &a, &e);
        az[i] = a;
        za[i] = FRAC_PI_2 - e;
    }
    /* Only the unique frequencies of the channels being modelled are evaluated (the others
     * are never read). */
    char *used = (char *)calloc((size_t)m->n_unique_freqs + 1, 1);
    for (int32_t f = m->f_start; f < m->f_start + m->f_count; f++) used[m->freq_map[f]] = 1;
    for (int32_t ut = 0; ut < m->n_unique_tiles; ut++) {
        int32_t t = m->unique_tiles[ut];
        const uint32_t *dl = obs->dipole_delays ? &obs->dipole_delays[16 * t] : NULL;
        const double *gn = obs->dipole_gains ? &obs->dipole_gains[32 * t] : NULL;
        for (int32_t uf = 0; uf < m->n_unique_freqs; uf++) {
            if (!used[uf]) continue;
            for (size_t i = 0; i < n; i++) {
                double jj[8];
                orc_analytic_beam_jones(az[i], za[i], m->unique_freqs[uf], dl, gn, jj);
                jones *o = &out[((size_t)ut * m->n_unique_freqs + uf) * n + i];
                for (int k = 0; k < 4; k++) {
                    o->j[k].re = jj[2 * k];
                    o->j[k].im = jj[2 * k + 1];
                }
            }
        }
    }
    free(used);
    free(az);
    free(za);
    return out;
}

static const double GAUSSIAN_EXP_CONST = -(FRAC_PI_2 * FRAC_PI_2) / LN_2; /* src/model/cpu.rs:35 */

/* The component folds of model_points / model_gaussians / model_shapelets
 * (src/model/cpu.rs:303-350, 426-492, 592-723) for channels [f_start, f_start+f_count). */
static void model_type(const modeller *m, int ty, int32_t f_start, int32_t f_count, int32_t vis_nf,
                       int32_t vis_f0, float *vis_f32, double *vis_f64, int32_t num_threads,
                       const int32_t *bl_list, int32_t n_bl_list) {
    const type_list *tl = &m->types[ty];
    if (tl->n == 0) return;
    const OrcObs *obs = m->obs;
    jones *beam = beam_responses(m, tl);
    const double *table = (ty == ORC_SHAPELET) ? orc_shapelet_basis_table_cached() : NULL;
    const double shapelet_const = SQRT_FRAC_PI_SQ_2_LN_2 / SBF_DX;
    int32_t n = tl->n;
    /* Output rows: every baseline, or (bl_list non-null) the listed baselines in list order. */
    int32_t nbl = bl_list ? n_bl_list : m->num_bl;
    (void)vis_nf;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) num_threads(num_threads > 0 ? num_threads : 1)
#else
    (void)num_threads;
#endif
    for (int32_t row = 0; row < nbl; row++) {
        int32_t b = bl_list ? bl_list[row] : row;
        int32_t t1 = m->tile_map[m->unflagged_to_total[m->bl_t1[b]]];
        int32_t t2 = m->tile_map[m->unflagged_to_total[m->bl_t2[b]]];
        c64 *env = (ty == ORC_SHAPELET) ? (c64 *)malloc(sizeof(c64) * (size_t)n) : NULL;
        for (int32_t ff = 0; ff < f_count; ff++) {
            int32_t f = f_start + ff;
            double freq = obs->freqs[f];
            int32_t uf = m->freq_map[f];
            const jones *b1 = &beam[((size_t)t1 * m->n_unique_freqs + uf) * n];
            const jones *b2 = &beam[((size_t)t2 * m->n_unique_freqs + uf) * n];
            const jones *fds = &tl->fds[(size_t)ff * n];
            double u, v, w;
            double one_on_lambda = freq / VEL_C;
            if (ty == ORC_SHAPELET) {
                u = m->uvws[3 * b] * one_on_lambda;
                v = m->uvws[3 * b + 1] * one_on_lambda;
                w = m->uvws[3 * b + 2] * one_on_lambda;
                for (int32_t i = 0; i < n; i++) {
                    const OrcComponent *c = tl->c[i];
                    const double *suvw = &tl->shapelet_uvw[((size_t)b * n + i) * 3];
                    double su = suvw[0] * one_on_lambda;
                    double sv = suvw[1] * one_on_lambda;
                    double s_pa = sin(c->pa), c_pa = cos(c->pa);
                    double x = su * s_pa + sv * c_pa;
                    double y = su * c_pa - sv * s_pa;
                    double const_x = c->maj * shapelet_const;
                    double const_y = -c->min * shapelet_const;
                    double x_pos = x * const_x + SBF_C;
                    double y_pos = y * const_y + SBF_C;
                    /* `as usize` saturates negatives to 0; the reference has no upper check, the
                     * table rows end in exact zeros so the last index is used beyond the end. */
                    double fx = floor(x_pos), fy = floor(y_pos);
                    long xi = fx < 0 ? 0 : (fx > (double)(SBF_L - 2) ? SBF_L - 2 : (long)fx);
                    long yi = fy < 0 ? 0 : (fy > (double)(SBF_L - 2) ? SBF_L - 2 : (long)fy);
                    c64 acc = {0.0, 0.0};
                    for (int32_t k = 0; k < c->num_coeffs; k++) {
                        const OrcShapeletCoeff *cf = &c->coeffs[k];
                        double f_hat = cf->value;
                        double x_low = table[(size_t)SBF_L * cf->n1 + xi];
                        double x_high = table[(size_t)SBF_L * cf->n1 + xi + 1];
                        double u_value = x_low + (x_high - x_low) * (x_pos - fx);
                        double y_low = table[(size_t)SBF_L * cf->n2 + yi];
                        double y_high = table[(size_t)SBF_L * cf->n2 + yi + 1];
                        double v_value = y_low + (y_high - y_low) * (y_pos - fy);
                        static const c64 I_POWER[4] = {{1.0, 0.0}, {0.0, 1.0}, {-1.0, 0.0}, {0.0, -1.0}};
                        c64 term = cscale(cscale(cscale(I_POWER[(cf->n1 + cf->n2) % 4], f_hat), u_value), v_value);
                        acc = cadd(acc, term);
                    }
                    env[i] = acc;
                }
            } else {
                u = m->uvws[3 * b] * freq / VEL_C;
                v = m->uvws[3 * b + 1] * freq / VEL_C;
                w = m->uvws[3 * b + 2] * freq / VEL_C;
            }
            jones accum = JONES_ZERO;
            for (int32_t i = 0; i < n; i++) {
                const double *lmn = &tl->lmn[3 * i];
                jones t = jmul(jmul(b1[i], fds[i]), jh(b2[i]));
                double ph = u * lmn[0] + v * lmn[1] + w * lmn[2];
                c64 cis = {cos(ph), sin(ph)};
                t = jmulc(t, cis);
                if (ty == ORC_GAUSSIAN) {
                    const OrcComponent *c = tl->c[i];
                    double s_pa = sin(c->pa), c_pa = cos(c->pa);
                    double k_x = u * s_pa + v * c_pa;
                    double k_y = u * c_pa - v * s_pa;
                    double envelope = exp(GAUSSIAN_EXP_CONST * ((c->maj * c->maj) * (k_x * k_x) +
                                                                (c->min * c->min) * (k_y * k_y)));
                    t = jmuls(t, envelope);
                } else if (ty == ORC_SHAPELET) {
                    t = jmulc(t, env[i]);
                }
                jadd_inplace(&accum, t);
            }
            size_t o = ((size_t)(f - vis_f0) * nbl + row) * 8;
            for (int k = 0; k < 4; k++) {
                vis_f32[o + 2 * k] += (float)accum.j[k].re;
                vis_f32[o + 2 * k + 1] += (float)accum.j[k].im;
                if (vis_f64) {
                    vis_f64[o + 2 * k] += accum.j[k].re;
                    vis_f64[o + 2 * k + 1] += accum.j[k].im;
                }
            }
        }
        free(env);
    }
    free(beam);
}

/* mask_pols (src/model/mod.rs:264-285). */
static void mask_pols(float *vis, double *vis64, size_t n_jones, int32_t pols) {
    if (pols == ORC_POLS_ALL) return;
    int keep[4] = {0, 0, 0, 0};
    switch (pols) {
    case ORC_POLS_XX: keep[0] = 1; break;
    case ORC_POLS_YY: keep[3] = 1; break;
    case ORC_POLS_XX_YY: keep[0] = keep[3] = 1; break;
    case ORC_POLS_XX_YY_XY: keep[0] = keep[1] = keep[3] = 1; break;
    default: return;
    }
    for (size_t i = 0; i < n_jones; i++)
        for (int k = 0; k < 4; k++)
            if (!keep[k]) {
                vis[8 * i + 2 * k] = 0.0f;
                vis[8 * i + 2 * k + 1] = 0.0f;
                if (vis64) {
                    vis64[8 * i + 2 * k] = 0.0;
                    vis64[8 * i + 2 * k + 1] = 0.0;
                }
            }
}

/* SkyModellerCpu::model_timestep_with without precession (src/model/cpu.rs:787-812). */
int orc_model_timestep(const OrcObs *obs, const OrcComponent *comps, const int32_t *source_offsets,
                       int32_t num_sources, int32_t comp_type_mask, float *vis_f32, double *vis_f64,
                       int32_t num_threads) {
    if (obs->num_unflagged_tiles < 2 || obs->num_freqs < 1) return 1;
    modeller m;
    build_modeller(&m, obs, comps, source_offsets, num_sources, 0, obs->num_freqs);
    for (int ty = 0; ty < 3; ty++)
        if (comp_type_mask & (1 << ty)) model_type(&m, ty, 0, obs->num_freqs, obs->num_freqs, 0, vis_f32, vis_f64,
                                                   num_threads, NULL, 0);
    mask_pols(vis_f32, vis_f64, (size_t)obs->num_freqs * m.num_bl, obs->pols);
    free_modeller(&m);
    return 0;
}

int orc_model_timestep_slice(const OrcObs *obs, const OrcComponent *comps, const int32_t *source_offsets,
                             int32_t num_sources, int32_t chan_start, int32_t chan_count, float *vis_f32,
                             double *vis_f64, int32_t num_threads) {
    if (obs->num_unflagged_tiles < 2 || chan_start < 0 || chan_count < 1 || chan_start + chan_count > obs->num_freqs)
        return 1;
    modeller m;
    build_modeller(&m, obs, comps, source_offsets, num_sources, chan_start, chan_count);
    for (int ty = 0; ty < 3; ty++)
        model_type(&m, ty, chan_start, chan_count, chan_count, chan_start, vis_f32, vis_f64, num_threads, NULL, 0);
    free_modeller(&m);
    return 0;
}

int orc_model_timestep_baselines(const OrcObs *obs, const OrcComponent *comps, const int32_t *source_offsets,
                                 int32_t num_sources, int32_t chan_start, int32_t chan_count,
                                 const int32_t *baselines, int32_t num_listed, float *vis_f32, double *vis_f64,
                                 int32_t num_threads) {
    if (obs->num_unflagged_tiles < 2 || chan_start < 0 || chan_count < 1 || chan_start + chan_count > obs->num_freqs ||
        baselines == NULL || num_listed < 1)
        return 1;
    int32_t nt = obs->num_unflagged_tiles;
    int32_t nbl = nt * (nt - 1) / 2;
    for (int32_t k = 0; k < num_listed; k++)
        if (baselines[k] < 0 || baselines[k] >= nbl) return 1;
    modeller m;
    build_modeller(&m, obs, comps, source_offsets, num_sources, chan_start, chan_count);
    for (int ty = 0; ty < 3; ty++)
        model_type(&m, ty, chan_start, chan_count, chan_count, chan_start, vis_f32, vis_f64, num_threads, baselines,
                   num_listed);
    free_modeller(&m);
    return 0;
}

/* SkyModellerCpu::model_timestep_autos_with (src/model/cpu.rs:828-936), tile map indexed with
 * the unflagged tile index as the reference does (cpu.rs:877). */
int orc_model_timestep_autos(const OrcObs *obs, const OrcComponent *comps, const int32_t *source_offsets,
                             int32_t num_sources, float *vis_f32, double *vis_f64, int32_t num_threads) {
    (void)num_threads;
    if (obs->num_unflagged_tiles < 1 || obs->num_freqs < 1) return 1;
    modeller m;
    build_modeller(&m, obs, comps, source_offsets, num_sources, 0, obs->num_freqs);
    int32_t nt = obs->num_unflagged_tiles;
    for (int ty = 0; ty < 3; ty++) {
        const type_list *tl = &m.types[ty];
        if (tl->n == 0) continue;
        jones *beam = beam_responses(&m, tl);
        for (int32_t t = 0; t < nt; t++) {
            int32_t ut = m.tile_map[t];
            for (int32_t f = 0; f < obs->num_freqs; f++) {
                int32_t uf = m.freq_map[f];
                const jones *bm = &beam[((size_t)ut * m.n_unique_freqs + uf) * tl->n];
                jones accum = JONES_ZERO;
                for (int32_t i = 0; i < tl->n; i++) {
                    jones tj = jmul(jmul(bm[i], tl->fds[(size_t)f * tl->n + i]), jh(bm[i]));
                    jadd_inplace(&accum, tj);
                }
                size_t o = ((size_t)f * nt + t) * 8;
                for (int k = 0; k < 4; k++) {
                    vis_f32[o + 2 * k] += (float)accum.j[k].re;
                    vis_f32[o + 2 * k + 1] += (float)accum.j[k].im;
                    if (vis_f64) {
                        vis_f64[o + 2 * k] += accum.j[k].re;
                        vis_f64[o + 2 * k + 1] += accum.j[k].im;
                    }
                }
            }
        }
        free(beam);
    }
    mask_pols(vis_f32, vis_f64, (size_t)obs->num_freqs * nt, obs->pols);
    free_modeller(&m);
    return 0;
}
