// SkyModeller handle: the host side of the predictor (src/model/gpu.rs:150-1527 and
// src/model/cpu.rs:73-180 re-expressed in C++ over the HIP kernels).
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "sm_abi_common.h"
#include "sm_tiled.h"
#include "sm_factored.h"

namespace sm {
namespace {

constexpr double kTau = 6.283185307179586;
constexpr double kDefaultSpecIndex = -0.8;  // src/constants.rs:8

struct Cplx {
    double re, im;
};

// --------------------------------------------------------------------- flux estimation
// FluxDensityType::estimate_at_freq (src/srclist/types/flux_density/mod.rs:158-270).
SmFluxDensity estimate_at_freq(const SmComponent &c, double freq) {
    auto scaled = [](SmFluxDensity fd, double r) {
        fd.i *= r;
        fd.q *= r;
        fd.u *= r;
        fd.v *= r;
        return fd;
    };
    if (c.flux_type == SM_POWER_LAW) {
        SmFluxDensity out = scaled(c.fds[0], std::pow(freq / c.fds[0].freq, c.si));
        out.freq = freq;
        return out;
    }
    if (c.flux_type == SM_CURVED_POWER_LAW) {
        SmFluxDensity pl = scaled(c.fds[0], std::pow(freq / c.fds[0].freq, c.si));
        pl.freq = freq;
        const double lr = std::log(freq / c.fds[0].freq);
        return scaled(pl, std::exp(c.q * (lr * lr)));
    }
    const SmFluxDensity *fds = c.fds;
    const int n = c.num_fds;
    double si;
    const SmFluxDensity *base;
    if (n == 1) {
        si = kDefaultSpecIndex;
        base = &fds[0];
    } else {
        const SmFluxDensity *a = &fds[0], *b = &fds[1];
        for (int w = 0; w + 1 < n; ++w) {
            a = &fds[w];
            b = &fds[w + 1];
            if (std::fabs(a->freq - freq) < 1e-3) return *a;
            if (std::fabs(b->freq - freq) < 1e-3) return *b;
            if (freq < b->freq) break;
        }
        const bool sa = std::signbit(a->i), sb = std::signbit(b->i);
        if (sa != sb) {
            const double run = b->freq - a->freq;
            SmFluxDensity out;
            out.freq = freq;
            out.i = a->i + (b->i - a->i) / run * (freq - a->freq);
            out.q = a->q + (b->q - a->q) / run * (freq - a->freq);
            out.u = a->u + (b->u - a->u) / run * (freq - a->freq);
            out.v = a->v + (b->v - a->v) / run * (freq - a->freq);
            return out;
        }
        si = std::log(b->i / a->i) / std::log(b->freq / a->freq);
        base = b->freq < freq ? b : a;
    }
    SmFluxDensity out = scaled(*base, std::pow(freq / base->freq, si));
    out.freq = freq;
    return out;
}

// FluxDensity::to_inst_stokes (flux_density/mod.rs:58-65): XX=I-Q, XY=U-iV, YX=U+iV, YY=I+Q.
JonesF64 inst_stokes(const SmFluxDensity &fd) {
    JonesF64 j;
    j.j00_re = fd.i - fd.q;
    j.j00_im = 0.0;
    j.j01_re = fd.u;
    j.j01_im = -fd.v;
    j.j10_re = fd.u;
    j.j10_im = fd.v;
    j.j11_re = fd.i + fd.q;
    j.j11_im = 0.0;
    return j;
}

// ------------------------------------------------------------------------- geometry
// RADec::to_lmn(..).prepare_for_rime() (marlu; src/srclist/types/components/mod.rs:214).
double4 lmn_rime(double ra, double dec, double pc_ra, double pc_dec) {
    const double d_ra = ra - pc_ra;
    const double l = std::cos(dec) * std::sin(d_ra);
    const double m = std::sin(dec) * std::cos(pc_dec) - std::cos(dec) * std::sin(pc_dec) * std::cos(d_ra);
    const double n = std::sin(dec) * std::sin(pc_dec) + std::cos(dec) * std::cos(pc_dec) * std::cos(d_ra);
    return make_double4(kTau * l, kTau * m, kTau * (n - 1.0), 0.0);
}

// Baseline UVWs, antenna1 - antenna2 over the upper triangle (mdbook/src/defs/coords.md).
void cross_uvws(const std::vector<double> &xyz, int nt, double ha, double dec, std::vector<UVW> &out,
                std::vector<double> *tile_out = nullptr) {
    const double s_ha = std::sin(ha), c_ha = std::cos(ha), s_dec = std::sin(dec), c_dec = std::cos(dec);
    std::vector<UVW> tile(nt);
    for (int i = 0; i < nt; ++i) {
        const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        tile[i].u = s_ha * x + c_ha * y;
        tile[i].v = -s_dec * c_ha * x + s_dec * s_ha * y + c_dec * z;
        tile[i].w = c_dec * c_ha * x - c_dec * s_ha * y + s_dec * z;
    }
    if (tile_out != nullptr) {
        tile_out->resize(3 * (size_t)nt);
        for (int i = 0; i < nt; ++i) {
            (*tile_out)[3 * i] = tile[i].u;
            (*tile_out)[3 * i + 1] = tile[i].v;
            (*tile_out)[3 * i + 2] = tile[i].w;
        }
    }
    out.resize((size_t)nt * (nt - 1) / 2);
    size_t b = 0;
    for (int i = 0; i < nt; ++i)
        for (int j = i + 1; j < nt; ++j, ++b) {
            out[b].u = tile[i].u - tile[j].u;
            out[b].v = tile[i].v - tile[j].v;
            out[b].w = tile[i].w - tile[j].w;
        }
}

// HADec::to_azel (ERFA hd2ae formulation), returns (az, za).
void azza(double ha, double dec, double lat, double &az, double &za) {
    const double s_ha = std::sin(ha), c_ha = std::cos(ha), s_dec = std::sin(dec), c_dec = std::cos(dec);
    const double s_lat = std::sin(lat), c_lat = std::cos(lat);
    const double x = -c_ha * c_dec * s_lat + s_dec * c_lat;
    const double y = -s_ha * c_dec;
    const double z = c_ha * c_dec * c_lat + s_dec * s_lat;
    const double r = std::hypot(x, y);
    const double a = r != 0.0 ? std::atan2(y, x) : 0.0;
    az = a < 0.0 ? a + kTau : a;
    za = 1.5707963267948966 - std::atan2(z, r);
}

// Hermite-function table B_n(x_j), x_j = j/100 - 50 (src/model/make_shapelet_basis_values.py),
// by the normalised recurrence B_{n+1} = sqrt(2/(n+1)) x B_n - sqrt(n/(n+1)) B_{n-1}.
std::vector<double> shapelet_table() {
    std::vector<double> t((size_t)SM_SBF_N * SM_SBF_L);
    for (int j = 0; j < SM_SBF_L; ++j) {
        const double x = j == SM_SBF_L - 1 ? 50.0 : (double)j * (100.0 / 10000.0) + -50.0;
        double prev = 0.0, cur = std::exp(-0.5 * (x * x));
        t[j] = cur;
        for (int n = 0; n + 1 < SM_SBF_N; ++n) {
            const double next = std::sqrt(2.0 / (n + 1)) * x * cur - std::sqrt((double)n / (n + 1)) * prev;
            prev = cur;
            cur = next;
            t[(size_t)(n + 1) * SM_SBF_L + j] = cur;
        }
    }
    return t;
}

// --------------------------------------------------------------------- component lists
struct HostType {
    int n_pl = 0, n_cpl = 0, n_list = 0;
    std::vector<RADec> radec;  // concatenated PL, CPL, list
    std::vector<double4> lmn, shape;
    std::vector<JonesF64> pl_fds, cpl_fds, list_fds;  // list: [nchan][n_list]
    std::vector<double> pl_sis, cpl_sis, cpl_qs;
    std::vector<ShapeletCoeff> coeffs;
    std::vector<int2> ranges;
    int n() const { return n_pl + n_cpl + n_list; }
};

struct DevType {
    DevBuf lmn, shape, pl_fds, pl_sis, cpl_fds, cpl_sis, cpl_qs, list_fds, coeffs, ranges;
    DevBuf bl_uv, hadec, az, za, beam, coh;
    HostType h;
};

template <class T> hipError_t upload(DevBuf &b, const std::vector<T> &v) {
    if (v.empty()) return hipSuccess;
    hipError_t e = b.reserve(sizeof(T) * v.size());
    if (e != hipSuccess) return e;
    return hipMemcpy(b.ptr, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
}

}  // namespace
}  // namespace sm

using namespace sm;

struct SmModeller {
    // configuration
    int nt = 0, n_total = 0, nbl = 0, nchan = 0, nchan_pad = 0;
    std::vector<double> xyz, freqs, freqs_pad;
    std::vector<int> flagged;
    double pc_ra = 0, pc_dec = 0, lon = 0, lat = 0;
    double dut1 = 0.0;      // UT1 - UTC [s] for the timestamp entry points
    bool precess = false;   // timestamps precessed to J2000
    std::vector<double> xyz_j2000;  // precessed tile XYZs of the last timestamp call
    int beam_type = SM_BEAM_NONE, pols = SM_POLS_XX_XY_YX_YY;
    SmSurrogate *surrogate = nullptr;  // beam surrogate with exact fallback (not owned)
    double surrogate_threshold = 0.0;
    int surrogate_fallbacks = 0;       // tiles that took the exact beam, summed over calls
    std::vector<unsigned> delays;  // [n_total][16]
    std::vector<double> gains;     // [n_total][32]
    // unique tiles / freqs (src/model/cpu.rs:105-161)
    int n_utile = 1, n_ufreq = 1;
    std::vector<int> tile_rows;  // unflagged tile -> unique tile row
    std::vector<int> unique_tile_total;  // unique row -> total-array tile
    std::vector<int> freq_map;
    std::vector<double> ufreqs;
    bool even = false;
    // device
    DevBuf d_freqs, d_xyz, d_bl_tiles, d_bl_rows, d_tile_rows, d_freq_map, d_ufreqs, d_delays, d_gains, d_sbf;
    DevBuf d_udelays, d_ugains;  // the unique tiles' delays [n_utile][16] and gains [n_utile][32]
    DevBuf d_sbf32;  // the basis table in FP32 (precision path)
    DevBuf d_chunks4, d_chunks6, d_chunks8;  // evenly spaced channel runs of at most 4, 6 and 8 channels
    int nchunks4 = 0, nchunks6 = 0, nchunks8 = 0;
    DevBuf d_uvw, d_vis32, d_vis64, d_accum, d_tile_uvw, d_pairs;
    // split-precision point path
    DevBuf d_fblocks, d_fblock_of, d_fout, d_fscratch, d_fkbuf, d_fcohf, d_fmap;
    DevBuf d_fkhi, d_fklo;  // pre-split phasor planes (factored kernel 2)
    int nfblocks = 0;
    // Per-tile beams. Class 0 is the most common beam row; every other beam row is a class
    // of its own. Factored tiles: class-0 tiles first, padded to a 64-tile block boundary,
    // then the other tiles. A pair with at least one class-0 tile q takes the MFMA path
    // with row coherency J_r S J_0^H (r = the class of the other tile p): blocks class-0 x
    // class-0 and other x class-0, the latter giving V_pq, or its conjugate transpose when
    // the baseline's first tile is the class-0 one. Pairs of two other-class tiles take the
    // FP64 direct kernel through the baseline subset d_g0_rest.
    int g0_row = -1, g0_n = 0, g0_nblocks = 0, g0_nmap = 0, g0_nrest = 0, g0_nclass = 1;
    DevBuf d_g0_tiles, d_g0_blocks, d_g0_block_of, d_g0_map, d_g0_rest, d_g0_uvw, d_g0_coh;
    DevBuf d_g0_rclass, d_g0_class_row, d_class_of_row;
    // Class-0 tiles for the tile-pair kernel: their tile-group pairs, gathered UVWs and the
    // baselines with a tile outside class 0.
    int g0_n0 = 0, g0_npairs = 0;
    DevBuf d_other_tiles, d_other_cls;
    int n_other = 0;
    DevBuf d_g0_pairs, d_g0_only_uvw, d_notg0_bl;
    std::vector<int> h_seg_off, h_class_row;  // segments of d_notg0_bl; beam row per class
    int npairs = 0;
    std::vector<double> h_tile_uvw;
    DevType types[3];
    Scratch scratch;  // run_type's work buffers
    // timing: event pairs bracketing each direct-kernel launch, resolved lazily
    bool timing = false;
    struct Timed {
        hipEvent_t e0, e1;
        int type;
    };
    std::vector<Timed> pending;
    std::vector<hipEvent_t> pool;
    int launches = 0;
    double ms_total = 0, ms_type[4] = {0, 0, 0, 0};
    int launches_factored = 0;
    std::vector<UVW> h_uvw;
};

namespace {

bool tile_is_flagged(const SmModeller *m, int t) {
    for (int f : m->flagged)
        if (f == t) return true;
    return false;
}

// Unique tiles by (delays, gains with delay 32 -> gain 0) (src/model/cpu.rs:105-143,944-964).
void unique_tiles(SmModeller *m) {
    m->tile_rows.assign(m->nt, 0);
    m->unique_tile_total.clear();
    if (m->beam_type == SM_BEAM_NONE) {
        m->n_utile = 1;
        m->unique_tile_total.push_back(0);
        return;
    }
    std::vector<std::vector<double>> keys;
    int u = 0;
    for (int t = 0; t < m->n_total; ++t) {
        if (tile_is_flagged(m, t)) continue;
        std::vector<double> key(48);
        for (int k = 0; k < 16; ++k) key[k] = (double)m->delays[16 * t + k];
        for (int k = 0; k < 32; ++k) key[16 + k] = m->delays[16 * t + (k % 16)] == 32 ? 0.0 : m->gains[32 * t + k];
        int found = -1;
        for (size_t i = 0; i < keys.size() && found < 0; ++i)
            if (std::memcmp(keys[i].data(), key.data(), sizeof(double) * 48) == 0) found = (int)i;
        if (found < 0) {
            found = (int)keys.size();
            keys.push_back(key);
            m->unique_tile_total.push_back(t);
        }
        m->tile_rows[u++] = found;
    }
    m->n_utile = (int)keys.size();
}

const char *validate_component(const SmComponent &c) {
    if (c.comp_type < SM_POINT || c.comp_type > SM_SHAPELET) return "unknown component type";
    if (c.flux_type < SM_LIST || c.flux_type > SM_CURVED_POWER_LAW) return "unknown flux-density type";
    if (c.fds == nullptr || c.num_fds < 1) return "a component needs at least one flux density";
    if (c.flux_type == SM_LIST)
        for (int i = 1; i < c.num_fds; ++i)
            if (c.fds[i].freq < c.fds[i - 1].freq) return "list flux densities are not sorted by frequency";
    if (c.comp_type == SM_SHAPELET) {
        if (c.num_coeffs < 0 || (c.num_coeffs > 0 && c.coeffs == nullptr)) return "bad shapelet coefficients";
        for (int i = 0; i < c.num_coeffs; ++i)
            if (c.coeffs[i].n1 >= SM_SBF_N || c.coeffs[i].n2 >= SM_SBF_N)
                return "shapelet coefficient order exceeds the basis table (n1, n2 <= 100)";
    }
    return nullptr;
}

// ComponentList / SkyModellerGpu::update_source_list (src/model/gpu.rs:521-861): nine buckets
// concatenated per type as power law, curved power law, list.
const char *set_components(SmModeller *m, const SmComponent *const *comps, int ncomp) {
    for (int i = 0; i < ncomp; ++i)
        if (const char *e = validate_component(*comps[i])) return e;
    for (int ty = 0; ty < 3; ++ty) {
        HostType h;
        for (int pass = 0; pass < 3; ++pass) {
            const int ft = pass == 0 ? SM_POWER_LAW : (pass == 1 ? SM_CURVED_POWER_LAW : SM_LIST);
            for (int i = 0; i < ncomp; ++i) {
                const SmComponent &c = *comps[i];
                if (c.comp_type != ty || c.flux_type != ft) continue;
                h.radec.push_back(RADec{c.ra, c.dec});
                h.lmn.push_back(lmn_rime(c.ra, c.dec, m->pc_ra, m->pc_dec));
                GaussianParams gp{c.maj, c.min, c.pa};
                if (ty == SM_GAUSSIAN) h.shape.push_back(host_gaussian_shape(gp));
                if (ty == SM_SHAPELET) {
                    h.shape.push_back(host_shapelet_shape(gp));
                    h.ranges.push_back(make_int2((int)h.coeffs.size(), c.num_coeffs));
                    for (int k = 0; k < c.num_coeffs; ++k) h.coeffs.push_back(c.coeffs[k]);
                }
                if (ft == SM_POWER_LAW) {
                    h.pl_fds.push_back(inst_stokes(estimate_at_freq(c, SM_POWER_LAW_FD_REF_FREQ)));
                    h.pl_sis.push_back(c.si);
                    h.n_pl++;
                } else if (ft == SM_CURVED_POWER_LAW) {
                    const SmFluxDensity at150 = estimate_at_freq(c, SM_POWER_LAW_FD_REF_FREQ);
                    h.cpl_fds.push_back(inst_stokes(at150));
                    // Re-referenced spectral index (src/model/gpu.rs:664-680).
                    double si = c.si;
                    if (c.fds[0].freq != SM_POWER_LAW_FD_REF_FREQ) {
                        const double lr = std::log(c.fds[0].freq / SM_POWER_LAW_FD_REF_FREQ);
                        si = (std::log(c.fds[0].i / at150.i) - c.q * (lr * lr)) / lr;
                    }
                    h.cpl_sis.push_back(si);
                    h.cpl_qs.push_back(c.q);
                    h.n_cpl++;
                } else {
                    h.n_list++;
                }
            }
        }
        // List flux densities [freq][component] (src/model/gpu.rs:768-774).
        if (h.n_list > 0) {
            std::vector<const SmComponent *> lists;
            for (int i = 0; i < ncomp; ++i)
                if (comps[i]->comp_type == ty && comps[i]->flux_type == SM_LIST) lists.push_back(comps[i]);
            h.list_fds.resize((size_t)m->nchan * h.n_list);
            for (int f = 0; f < m->nchan; ++f)
                for (int i = 0; i < h.n_list; ++i)
                    h.list_fds[(size_t)f * h.n_list + i] = inst_stokes(estimate_at_freq(*lists[i], m->freqs[f]));
        }
        DevType &d = m->types[ty];
        SM_TRY(upload(d.lmn, h.lmn));
        SM_TRY(upload(d.shape, h.shape));
        SM_TRY(upload(d.pl_fds, h.pl_fds));
        SM_TRY(upload(d.pl_sis, h.pl_sis));
        SM_TRY(upload(d.cpl_fds, h.cpl_fds));
        SM_TRY(upload(d.cpl_sis, h.cpl_sis));
        SM_TRY(upload(d.cpl_qs, h.cpl_qs));
        SM_TRY(upload(d.list_fds, h.list_fds));
        SM_TRY(upload(d.coeffs, h.coeffs));
        SM_TRY(upload(d.ranges, h.ranges));
        d.h = std::move(h);
    }
    return nullptr;
}

TypeInputs type_inputs(DevType &d) {
    const HostType &h = d.h;
    TypeInputs in{};
    in.fb = FluxBuckets{h.n_pl,
                        h.n_cpl,
                        h.n_list,
                        d.pl_fds.as<JonesF64>(),
                        d.pl_sis.as<double>(),
                        d.cpl_fds.as<JonesF64>(),
                        d.cpl_sis.as<double>(),
                        d.cpl_qs.as<double>(),
                        d.list_fds.as<JonesF64>()};
    in.lmn = d.lmn.as<double4>();
    in.shape = d.shape.as<double4>();
    in.bl_uv = d.bl_uv.as<double2>();
    in.coeff_range = d.ranges.as<int2>();
    in.coeffs = d.coeffs.as<ShapeletCoeff>();
    in.coeff_count = (int)d.h.coeffs.size();
    return in;
}

// Beam responses [unique tile][unique freq][component] for one type at this LST
// (src/model/gpu.rs:884-911) and the matching BeamDesc.
const char *prepare_beam(SmModeller *m, DevType &d, double lst, double latitude, BeamDesc &b) {
    const int n = d.h.n();
    b = BeamDesc{};
    b.tile_rows = m->d_tile_rows.as<int>();
    b.freq_map = m->d_freq_map.as<int>();
    b.n_ufreq = m->n_ufreq;
    b.n_dir = n;
    b.dir_offset = 0;
    b.uniform_row = 0;
    if (m->beam_type == SM_BEAM_NONE || n == 0) return nullptr;
    std::vector<double> az(n), za(n);
    for (int i = 0; i < n; ++i) azza(lst - d.h.radec[i].ra, d.h.radec[i].dec, latitude, az[i], za[i]);
    SM_TRY(upload(d.az, az));
    SM_TRY(upload(d.za, za));
    SM_TRY(d.beam.reserve(sizeof(JonesF64) * (size_t)m->n_utile * m->n_ufreq * n));
    if (m->surrogate != nullptr) {
        std::vector<uint32_t> dl;
        std::vector<double> gn;
        for (int u = 0; u < m->n_utile; ++u) {
            const int t = m->unique_tile_total[u];
            dl.insert(dl.end(), m->delays.begin() + 16 * t, m->delays.begin() + 16 * t + 16);
            gn.insert(gn.end(), m->gains.begin() + 32 * t, m->gains.begin() + 32 * t + 32);
        }
        int32_t nfb = 0;
        if (const char *e = sm_surrogate_beam(m->surrogate, m->n_utile, dl.data(), gn.data(), m->n_ufreq,
                                              m->ufreqs.data(), n, d.az.as<double>(), d.za.as<double>(),
                                              m->surrogate_threshold, d.beam.as<JonesF64>(), &nfb))
            return e;
        m->surrogate_fallbacks += nfb;
    }
    // All unique tiles in one launch (or a few, for the grid's y extent), from their contiguous
    // delays and gains.
    if (m->surrogate == nullptr)
        SM_TRY(launch_analytic_beam(m->n_utile, m->d_udelays.as<unsigned>(), m->d_ugains.as<double>(), m->n_ufreq,
                                    m->d_ufreqs.as<double>(), n, d.az.as<double>(), d.za.as<double>(),
                                    d.beam.as<JonesF64>(), 0));
    b.jones = d.beam.as<JonesF64>();
    if (m->n_utile == 1) {
        b.uniform_row = 0;
    } else {
        b.uniform_row = -1;
        b.bl_rows = m->d_bl_rows.as<int2>();
    }
    return nullptr;
}

const char *prepare_shapelet_uvs(SmModeller *m, DevType &d, double lst) {
    const int n = d.h.n();
    if (n == 0) return nullptr;
    std::vector<double> hd(4 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        const double ha = lst - d.h.radec[i].ra, dec = d.h.radec[i].dec;
        hd[4 * i] = std::sin(ha);
        hd[4 * i + 1] = std::cos(ha);
        hd[4 * i + 2] = std::sin(dec);
        hd[4 * i + 3] = std::cos(dec);
    }
    SM_TRY(upload(d.hadec, hd));
    SM_TRY(d.bl_uv.reserve(sizeof(double2) * (size_t)n * m->nbl));
    SM_TRY(launch_shapelet_bl_uv(n, m->nt, m->nbl, d.hadec.as<double>(), m->d_xyz.as<double>(),
                                 m->d_bl_tiles.as<int2>(), d.bl_uv.as<double2>(), 0));
    return nullptr;
}

void mask_pols(int pols, JonesF32 *v, JonesF64 *v64, size_t n) {
    if (pols == SM_POLS_XX_XY_YX_YY) return;
    bool keep[4] = {false, false, false, false};
    if (pols == SM_POLS_XX) keep[0] = true;
    if (pols == SM_POLS_YY) keep[3] = true;
    if (pols == SM_POLS_XX_YY) keep[0] = keep[3] = true;
    if (pols == SM_POLS_XX_YY_XY) keep[0] = keep[1] = keep[3] = true;
    for (size_t i = 0; i < n; ++i) {
        float *f = reinterpret_cast<float *>(&v[i]);
        double *d = v64 ? reinterpret_cast<double *>(&v64[i]) : nullptr;
        for (int k = 0; k < 4; ++k)
            if (!keep[k]) {
                f[2 * k] = f[2 * k + 1] = 0.0f;
                if (d) d[2 * k] = d[2 * k + 1] = 0.0;
            }
    }
}

}  // namespace

extern "C" const char *sm_new(const SmObsParams *p, const SmComponent *comps, const int32_t *source_offsets,
                              int32_t num_sources, SmModeller **out) {
    if (p == nullptr || out == nullptr) return "null argument";
    if (p->num_unflagged_tiles < 1) return "need at least one unflagged tile";
    if (p->num_freqs < 1) return "need at least one frequency";
    if (p->unflagged_tile_xyzs == nullptr || p->freqs == nullptr) return "null tile positions or frequencies";
    if (num_sources < 0 || (num_sources > 0 && (source_offsets == nullptr || comps == nullptr)))
        return "bad source list";
    if (p->beam_type != SM_BEAM_NONE && p->beam_type != SM_BEAM_ANALYTIC) return "unknown beam type";
    SmModeller *m = new (std::nothrow) SmModeller();
    if (m == nullptr) return "out of host memory";
    m->nt = p->num_unflagged_tiles;
    m->n_total = p->num_unflagged_tiles + p->num_flagged_tiles;
    m->nbl = m->nt * (m->nt - 1) / 2;
    m->nchan = p->num_freqs;
    m->nchan_pad = pad_channels(m->nchan);
    m->xyz.assign(p->unflagged_tile_xyzs, p->unflagged_tile_xyzs + 3 * (size_t)m->nt);
    m->freqs.assign(p->freqs, p->freqs + m->nchan);
    m->freqs_pad.resize(m->nchan_pad);
    pad_freqs(m->freqs.data(), m->nchan, m->freqs_pad.data());
    m->even = chunks_even(m->freqs_pad.data(), m->nchan);
    if (p->num_flagged_tiles > 0) m->flagged.assign(p->flagged_tiles, p->flagged_tiles + p->num_flagged_tiles);
    m->pc_ra = p->phase_centre_ra;
    m->pc_dec = p->phase_centre_dec;
    m->lon = p->array_longitude;
    m->lat = p->array_latitude;
    m->dut1 = p->dut1;
    m->precess = p->apply_precession != 0;
    m->beam_type = p->beam_type;
    m->pols = p->pols;
    m->delays.assign((size_t)m->n_total * 16, 0u);
    m->gains.assign((size_t)m->n_total * 32, 1.0);
    if (p->dipole_delays) std::memcpy(m->delays.data(), p->dipole_delays, sizeof(unsigned) * m->delays.size());
    if (p->dipole_gains) std::memcpy(m->gains.data(), p->dipole_gains, sizeof(double) * m->gains.size());
    unique_tiles(m);
    // Unique beam frequencies: the analytic beam is defined at every frequency, so only exact
    // duplicates are merged (src/model/cpu.rs:145-161).
    m->freq_map.resize(m->nchan);
    for (int f = 0; f < m->nchan; ++f) {
        int idx = -1;
        for (size_t u = 0; u < m->ufreqs.size(); ++u)
            if (m->ufreqs[u] == m->freqs[f]) idx = (int)u;
        if (idx < 0) {
            idx = (int)m->ufreqs.size();
            m->ufreqs.push_back(m->freqs[f]);
        }
        m->freq_map[f] = idx;
    }
    m->n_ufreq = (int)m->ufreqs.size();
    std::vector<int2> bl_tiles(m->nbl), bl_rows(m->nbl);
    for (int i = 0, b = 0; i < m->nt; ++i)
        for (int j = i + 1; j < m->nt; ++j, ++b) {
            bl_tiles[b] = make_int2(i, j);
            bl_rows[b] = make_int2(m->tile_rows[i], m->tile_rows[j]);
        }
    auto fail = [m](const char *e) {
        sm_free(m);
        return e;
    };
#define SM_NEW_TRY(call)                                        \
    do {                                                        \
        hipError_t e_ = (call);                                 \
        if (e_ != hipSuccess) return fail(hipGetErrorString(e_)); \
    } while (0)
    SM_NEW_TRY(upload(m->d_freqs, m->freqs_pad));
    SM_NEW_TRY(upload(m->d_xyz, m->xyz));
    SM_NEW_TRY(upload(m->d_bl_tiles, bl_tiles));
    SM_NEW_TRY(upload(m->d_bl_rows, bl_rows));
    SM_NEW_TRY(upload(m->d_tile_rows, m->tile_rows));
    SM_NEW_TRY(upload(m->d_freq_map, m->freq_map));
    SM_NEW_TRY(upload(m->d_ufreqs, m->ufreqs));
    SM_NEW_TRY(upload(m->d_delays, m->delays));
    SM_NEW_TRY(upload(m->d_gains, m->gains));
    {
        std::vector<unsigned> ud;
        std::vector<double> ug;
        for (int t : m->unique_tile_total) {
            if (16 * (size_t)t + 16 > m->delays.size() || 32 * (size_t)t + 32 > m->gains.size()) break;
            ud.insert(ud.end(), m->delays.begin() + 16 * t, m->delays.begin() + 16 * t + 16);
            ug.insert(ug.end(), m->gains.begin() + 32 * t, m->gains.begin() + 32 * t + 32);
        }
        if (!ud.empty()) {
            SM_NEW_TRY(upload(m->d_udelays, ud));
            SM_NEW_TRY(upload(m->d_ugains, ug));
        }
    }
    SM_NEW_TRY(upload(m->d_sbf, shapelet_table()));
    {
        const std::vector<int2> c4 = even_chunks(m->freqs_pad.data(), m->nchan, 4);
        const std::vector<int2> c6 = even_chunks(m->freqs_pad.data(), m->nchan, 6);
        const std::vector<int2> c8 = even_chunks(m->freqs_pad.data(), m->nchan, 8);
        m->nchunks4 = (int)c4.size();
        m->nchunks6 = (int)c6.size();
        m->nchunks8 = (int)c8.size();
        SM_NEW_TRY(upload(m->d_chunks4, c4));
        SM_NEW_TRY(upload(m->d_chunks6, c6));
        SM_NEW_TRY(upload(m->d_chunks8, c8));
    }
    SM_NEW_TRY(m->d_sbf32.reserve(sizeof(float) * (size_t)SM_SBF_N * SM_SBF_L));
    SM_NEW_TRY(launch_to_f32((size_t)SM_SBF_N * SM_SBF_L, m->d_sbf.as<double>(), m->d_sbf32.as<float>(), 0));
    {
        const std::vector<int4> pairs = tile_pairs(m->nt);
        m->npairs = (int)pairs.size();
        SM_NEW_TRY(upload(m->d_pairs, pairs));
        std::vector<int> block_of;
        const std::vector<int2> fblocks = factored_blocks(m->nt, block_of);
        m->nfblocks = (int)fblocks.size();
        SM_NEW_TRY(upload(m->d_fblocks, fblocks));
        SM_NEW_TRY(upload(m->d_fblock_of, block_of));
        std::vector<int4> fmap(m->nbl);
        for (int b = 0; b < m->nbl; ++b) fmap[b] = make_int4(b, bl_tiles[b].x, bl_tiles[b].y, 0);
        SM_NEW_TRY(upload(m->d_fmap, fmap));
        if (m->n_utile > 1) {
            // Class 0 = the beam row shared by the most tiles.
            std::vector<int> count(m->n_utile, 0);
            for (int t = 0; t < m->nt; ++t) count[m->tile_rows[t]]++;
            int r0 = 0;
            for (int r = 1; r < m->n_utile; ++r)
                if (count[r] > count[r0]) r0 = r;
            if (count[r0] >= 2) {
                std::vector<int> class_of_row(m->n_utile, -1), class_row;
                class_of_row[r0] = 0;
                class_row.push_back(r0);
                for (int r = 0; r < m->n_utile; ++r)
                    if (class_of_row[r] < 0 && count[r] > 0) {
                        class_of_row[r] = (int)class_row.size();
                        class_row.push_back(r);
                    }
                // Factored numbering: class-0 tiles, padding to a block boundary, the rest.
                std::vector<int> ftile, local(m->nt, -1);
                for (int t = 0; t < m->nt; ++t)
                    if (m->tile_rows[t] == r0) {
                        local[t] = (int)ftile.size();
                        ftile.push_back(t);
                    }
                const int n0 = (int)ftile.size();
                const int g0b = (n0 + 63) / 64;
                ftile.resize((size_t)g0b * 64, -1);
                for (int t = 0; t < m->nt; ++t)
                    if (m->tile_rows[t] != r0) {
                        local[t] = (int)ftile.size();
                        ftile.push_back(t);
                    }
                const int nft = (int)ftile.size();
                const int G = (nft + 63) / 64;
                std::vector<int> rclass((size_t)G * 64, 0);
                for (int t = 0; t < nft; ++t)
                    if (ftile[t] >= 0) rclass[t] = class_of_row[m->tile_rows[ftile[t]]];
                std::vector<int2> blocks;
                std::vector<int> block_of((size_t)G * G, 0);
                for (int p = 0; p < G; ++p)
                    for (int q = 0; q < g0b; ++q) {
                        if (p < g0b && q < p) continue;
                        block_of[(size_t)p * G + q] = (int)blocks.size();
                        blocks.push_back(make_int2(p, q));
                    }
                {
                    // Baselines with a tile outside class 0, by segment: (class r, class 0),
                    // (class 0, class r) for r = 1 .. R-1, then (other, other).
                    const std::vector<int4> p0 = tile_pairs(n0);
                    const int R = (int)class_row.size();
                    std::vector<std::vector<int>> seg(2 * (R - 1) + 1);
                    for (int b = 0; b < m->nbl; ++b) {
                        const int ci = class_of_row[m->tile_rows[bl_tiles[b].x]];
                        const int cj = class_of_row[m->tile_rows[bl_tiles[b].y]];
                        if (ci == 0 && cj == 0) continue;
                        if (cj == 0)
                            seg[2 * (ci - 1)].push_back(b);
                        else if (ci == 0)
                            seg[2 * (cj - 1) + 1].push_back(b);
                        else
                            seg[2 * (R - 1)].push_back(b);
                    }
                    std::vector<int> notg0;
                    m->h_seg_off.assign(1, 0);
                    for (const auto &sg : seg) {
                        notg0.insert(notg0.end(), sg.begin(), sg.end());
                        m->h_seg_off.push_back((int)notg0.size());
                    }
                    m->h_class_row = class_row;
                    m->g0_n0 = n0;
                    m->g0_npairs = (int)p0.size();
                    SM_NEW_TRY(upload(m->d_g0_pairs, p0));
                    SM_NEW_TRY(upload(m->d_notg0_bl, notg0));
                }
                std::vector<int4> map;
                std::vector<int> rest;
                for (int b = 0; b < m->nbl; ++b) {
                    const int i = bl_tiles[b].x, j = bl_tiles[b].y;
                    const bool ci = m->tile_rows[i] == r0, cj = m->tile_rows[j] == r0;
                    if (ci && cj)
                        map.push_back(make_int4(b, local[i], local[j], 0));
                    else if (cj)
                        map.push_back(make_int4(b, local[i], local[j], 0));
                    else if (ci)
                        map.push_back(make_int4(b, local[j], local[i], 1));
                    else
                        rest.push_back(b);
                }
                m->g0_row = r0;
                m->g0_n = nft;
                m->g0_nblocks = (int)blocks.size();
                m->g0_nmap = (int)map.size();
                m->g0_nrest = (int)rest.size();
                m->g0_nclass = (int)class_row.size();
                SM_NEW_TRY(upload(m->d_g0_tiles, ftile));
                SM_NEW_TRY(upload(m->d_g0_blocks, blocks));
                SM_NEW_TRY(upload(m->d_g0_block_of, block_of));
                SM_NEW_TRY(upload(m->d_g0_map, map));
                SM_NEW_TRY(upload(m->d_g0_rest, rest));
                SM_NEW_TRY(upload(m->d_g0_rclass, rclass));
                {
                    std::vector<int> ot, oc;
                    for (int t = g0b * 64; t < nft; ++t) {
                        ot.push_back(ftile[t]);
                        oc.push_back(rclass[t]);
                    }
                    m->n_other = (int)ot.size();
                    if (m->n_other > 0) {
                        SM_NEW_TRY(upload(m->d_other_tiles, ot));
                        SM_NEW_TRY(upload(m->d_other_cls, oc));
                    }
                }
                SM_NEW_TRY(upload(m->d_g0_class_row, class_row));
                SM_NEW_TRY(upload(m->d_class_of_row, class_of_row));
            }
        }
    }
#undef SM_NEW_TRY
    // Reverse the source order (src/model/gpu.rs:230-241, src/model/cpu.rs:86-93).
    std::vector<const SmComponent *> order;
    for (int s = num_sources - 1; s >= 0; --s)
        for (int i = source_offsets[s]; i < source_offsets[s + 1]; ++i) order.push_back(&comps[i]);
    if (const char *e = set_components(m, order.data(), (int)order.size())) return fail(e);
    *out = m;
    return nullptr;
}

extern "C" void sm_free(SmModeller *m) {
    if (m == nullptr) return;
    (void)hipDeviceSynchronize();
    for (auto &t : m->pending) {
        (void)hipEventDestroy(t.e0);
        (void)hipEventDestroy(t.e1);
    }
    for (hipEvent_t e : m->pool) (void)hipEventDestroy(e);
    delete m;
}

extern "C" const char *sm_update_with_a_source(SmModeller *m, const SmComponent *comps, int32_t num_comps,
                                               double pc_ra, double pc_dec) {
    if (m == nullptr || (num_comps > 0 && comps == nullptr)) return "null argument";
    m->pc_ra = pc_ra;
    m->pc_dec = pc_dec;
    std::vector<const SmComponent *> order;
    for (int i = 0; i < num_comps; ++i) order.push_back(&comps[i]);
    return set_components(m, order.data(), (int)order.size());
}

// Steps of 16 components per FP32 accumulation span of the MFMA point kernel.
static int g_factored_seg_steps = 64;

extern "C" const char *sm_set_factored_segment(int32_t steps) {
    if (steps < 1 || steps > 4096) return "segment must be 1..4096 steps";
    g_factored_seg_steps = steps;
    return nullptr;
}

extern "C" const char *sm_set_factored_kernel(int32_t kernel) {
    if (kernel < 0 || kernel > 2) return "factored kernel must be 0, 1 or 2";
    g_factored_kernel = kernel;
    return nullptr;
}

// One timestep on the device. factored_dtype < 0: FP64 direct kernels for every type;
// otherwise points go through the split-precision MFMA path in that operand type.
// uvw_xyz (nullable): the tile XYZs the cross UVWs come from, when not the modeller's own
// (the precessed XYZs of a timestamp call); shapelet UVs always use the modeller's own.
static const char *model_timestep_device(SmModeller *m, double lst, double latitude, const UVW *d_uvws,
                                         JonesF32 *d_vis_fb, JonesF64 *d_vis_fb_f64, int factored_dtype,
                                         const std::vector<double> *uvw_xyz = nullptr) {
    if (m == nullptr || d_vis_fb == nullptr) return "null argument";
    if (m->nbl == 0) return nullptr;
    const UVW *uvw = d_uvws;
    const double *tile_uvw = nullptr;
    if (uvw == nullptr) {
        cross_uvws(uvw_xyz != nullptr ? *uvw_xyz : m->xyz, m->nt, lst - m->pc_ra, m->pc_dec, m->h_uvw,
                   &m->h_tile_uvw);
        SM_TRY(upload(m->d_uvw, m->h_uvw));
        SM_TRY(upload(m->d_tile_uvw, m->h_tile_uvw));
        uvw = m->d_uvw.as<UVW>();
        tile_uvw = m->d_tile_uvw.as<double>();
    } else if (m->nt >= 2) {
        SM_TRY(m->d_tile_uvw.reserve(sizeof(double) * 3 * m->nt));
        SM_TRY(launch_tile_uvw_from_baselines(m->nt, uvw, m->d_tile_uvw.as<double>(), 0));
        tile_uvw = m->d_tile_uvw.as<double>();
    }
    ModelCtx ctx;
    ctx.nbl = m->nbl;
    ctx.nchan = m->nchan;
    ctx.nchan_pad = m->nchan_pad;
    ctx.ntiles = m->nt;
    ctx.d_freqs_pad = m->d_freqs.as<double>();
    // Chunks are runs of evenly spaced channels: every kernel takes its evenly spaced form
    // whatever channels are flagged.
    ctx.chunks.c4 = m->d_chunks4.as<int2>();
    ctx.chunks.n4 = m->nchunks4;
    ctx.chunks.c6 = m->d_chunks6.as<int2>();
    ctx.chunks.n6 = m->nchunks6;
    ctx.chunks.c8 = m->d_chunks8.as<int2>();
    ctx.chunks.n8 = m->nchunks8;
    ctx.even = true;
    ctx.d_uvw = uvw;
    ctx.d_sbf = m->d_sbf.as<double>();
    ctx.d_tile_uvw = tile_uvw;
    ctx.d_pairs = m->d_pairs.as<int4>();
    ctx.npairs = m->npairs;
    ctx.fp32 = factored_dtype >= 0;
    ctx.d_sbf32 = m->d_sbf32.as<float>();
    if (m->g0_row >= 0) {
        ctx.nclass = m->g0_nclass;
        ctx.class_row0 = m->g0_row;
        ctx.d_class_row = m->d_g0_class_row.as<int>();
        ctx.d_class_of_row = m->d_class_of_row.as<int>();
        if (tile_uvw != nullptr && m->g0_n0 >= 2) {
            // The first g0_n0 factored tiles are the class-0 tiles, ascending.
            SM_TRY(m->d_g0_only_uvw.reserve(sizeof(double) * 3 * m->g0_n0));
            SM_TRY(launch_gather_rows3(m->g0_n0, m->d_g0_tiles.as<int>(), tile_uvw, m->d_g0_only_uvw.as<double>(), 0));
            ctx.g0_n = m->g0_n0;
            ctx.g0_npairs = m->g0_npairs;
            ctx.h_seg_off = m->h_seg_off.data();
            ctx.h_class_row = m->h_class_row.data();
            ctx.d_g0_tiles = m->d_g0_tiles.as<int>();
            ctx.d_g0_uvw = m->d_g0_only_uvw.as<double>();
            ctx.d_g0_pairs = m->d_g0_pairs.as<int4>();
            ctx.d_notg0_bl = m->d_notg0_bl.as<int>();
            if (m->n_other > 0) {
                ctx.d_other_tiles = m->d_other_tiles.as<int>();
                ctx.d_other_cls = m->d_other_cls.as<int>();
                ctx.n_other = m->n_other;
            }
        }
    }
    const Kind kinds[3] = {KIND_POINT, KIND_GAUSSIAN, KIND_SHAPELET};
    for (int ty = 0; ty < 3; ++ty) {
        DevType &d = m->types[ty];
        if (d.h.n() == 0) continue;
        if (const char *e = prepare_beam(m, d, lst, latitude, ctx.beam)) return e;
        if (ty == SM_SHAPELET)
            if (const char *e = prepare_shapelet_uvs(m, d, lst)) return e;
        TypeInputs in = type_inputs(d);
        if (ty == SM_SHAPELET) {
            in.uv_hadec = d.hadec.as<double>();
            in.uv_xyz = m->d_xyz.as<double>();
            in.uv_bl_tiles = m->d_bl_tiles.as<int2>();
        }
        const bool have_uvw = tile_uvw != nullptr && m->nt >= 2;
        const bool fact_uniform = ty == SM_POINT && factored_dtype >= 0 && ctx.beam.uniform_row >= 0 && have_uvw;
        const bool fact_group = ty == SM_POINT && factored_dtype >= 0 && ctx.beam.uniform_row < 0 &&
                                ctx.beam.jones != nullptr && m->g0_row >= 0 && have_uvw;
        auto timed = [&](int type, hipEvent_t &e0, hipEvent_t &e1) -> const char * {
            e0 = e1 = nullptr;
            if (!m->timing) return nullptr;
            for (hipEvent_t *e : {&e0, &e1}) {
                if (!m->pool.empty()) {
                    *e = m->pool.back();
                    m->pool.pop_back();
                } else {
                    SM_TRY(hipEventCreate(e));
                }
            }
            m->pending.push_back(SmModeller::Timed{e0, e1, type});
            return nullptr;
        };
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (!fact_uniform && !fact_group) {
            if (const char *e = timed(ty, e0, e1)) return e;
            SM_TRY(run_type(kinds[ty], in, ctx, d.coh, m->scratch, d_vis_fb, d_vis_fb_f64, 0, e0, e1));
            continue;
        }
        const int n = d.h.n();
        const double *coh = nullptr;
        const double *fuvw = tile_uvw;
        int fnt = m->nt, fnblocks = m->nfblocks, nmap = m->nbl;
        const int2 *fblocks = m->d_fblocks.as<int2>();
        const int *fblock_of = m->d_fblock_of.as<int>();
        const int4 *fmap = m->d_fmap.as<int4>();
        int nclass = 1;
        if (fact_uniform) {
            // Coherency J S J^H for the shared beam row.
            SM_TRY(d.coh.reserve(sizeof(double) * 8 * (size_t)n * m->nchan_pad));
            SM_TRY(launch_coherency(n, m->nchan, m->nchan_pad, ctx.d_freqs_pad, in.fb, ctx.beam, true,
                                    d.coh.as<double>(), 0));
            coh = d.coh.as<double>();
        } else {
            // Baselines outside group 0 (pairs of two tiles outside beam class 0): FP64 with
            // per-tile beams, on the own-beam tile-pair kernel when they are all the pairs of the
            // listed tiles, else on the baseline kernel's subset form.
            ModelCtx rest = ctx;
            rest.d_bl_subset = m->d_g0_rest.as<int>();
            rest.nsub = m->g0_nrest;
            rest.nclass = 0;  // no baseline of the subset has a class-0 tile
            const bool own_pairs = g_direct_variant == 0 && m->n_other >= 2 && ctx.d_other_tiles != nullptr &&
                                   (long)m->g0_nrest == (long)m->n_other * (m->n_other - 1) / 2;
            if (m->g0_nrest > 0 && own_pairs) {
                if (const char *e = timed(ty, e0, e1)) return e;
                SM_TRY(d.coh.reserve(sizeof(double) * 8 * (size_t)n * m->nchan_pad));
                SM_TRY(launch_coherency(n, m->nchan, m->nchan_pad, ctx.d_freqs_pad, in.fb, ctx.beam, false,
                                        d.coh.as<double>(), 0));
                PairJonesArgs pj;
                pj.ntiles_list = m->n_other;
                pj.ntiles_bl = m->nt;
                pj.nbl = m->nbl;
                pj.nchan = m->nchan;
                pj.nchan_pad = m->nchan_pad;
                pj.ncomp = n;
                pj.tiles = ctx.d_other_tiles;
                pj.tile_uvw = tile_uvw;
                pj.freqs = ctx.d_freqs_pad;
                pj.lmn = in.lmn;
                pj.shape = in.shape;
                pj.coh = d.coh.as<double>();
                pj.beam = ctx.beam;
                pj.fp32 = true;  // the precision path: FP32 with FP64 phases, as its Gaussians
                pj.vis32 = d_vis_fb;
                pj.vis64 = d_vis_fb_f64;
                pj.chunks = ctx.chunks;
                if (e0 != nullptr) (void)hipEventRecord(e0, 0);
                SM_TRY(launch_pair_jones(kinds[ty], pj, &m->scratch.accum, 0));
                if (e1 != nullptr) (void)hipEventRecord(e1, 0);
            } else if (m->g0_nrest > 0) {
                if (const char *e = timed(ty, e0, e1)) return e;
                SM_TRY(run_type(kinds[ty], in, rest, d.coh, m->scratch, d_vis_fb, d_vis_fb_f64, 0, e0, e1));
            }
            // Pairs with a class-0 tile: S here, J_r S J_0^H formed per chunk by the factored
            // path; UVWs of the factored tiles (zeros for padding).
            SM_TRY(m->d_g0_coh.reserve(sizeof(double) * 8 * (size_t)n * m->nchan_pad));
            SM_TRY(launch_coherency(n, m->nchan, m->nchan_pad, ctx.d_freqs_pad, in.fb, ctx.beam, false,
                                    m->d_g0_coh.as<double>(), 0));
            coh = m->d_g0_coh.as<double>();
            nclass = m->g0_nclass;
            SM_TRY(m->d_g0_uvw.reserve(sizeof(double) * 3 * m->g0_n));
            SM_TRY(launch_gather_rows3(m->g0_n, m->d_g0_tiles.as<int>(), tile_uvw, m->d_g0_uvw.as<double>(), 0));
            fuvw = m->d_g0_uvw.as<double>();
            fnt = m->g0_n;
            fnblocks = m->g0_nblocks;
            nmap = m->g0_nmap;
            fblocks = m->d_g0_blocks.as<int2>();
            fblock_of = m->d_g0_block_of.as<int>();
            fmap = m->d_g0_map.as<int4>();
        }
        if (const char *e = timed(3, e0, e1)) return e;
        FactoredArgs fa;
        fa.ntiles = fnt;
        fa.nchan = m->nchan;
        fa.nchan_pad = m->nchan_pad;
        fa.ncomp = n;
        fa.nblocks = fnblocks;
        fa.seg_steps = g_factored_seg_steps;
        fa.even = m->even;
        fa.blocks = fblocks;
        fa.freqs = ctx.d_freqs_pad;
        fa.tile_uvw = fuvw;
        fa.lmn = in.lmn;
        fa.coh = coh;
        fa.scale = nullptr;
        fa.nclass = nclass;
        if (nclass > 1) {
            fa.rclass = m->d_g0_rclass.as<int>();
            fa.class_row = m->d_g0_class_row.as<int>();
            fa.beam = ctx.beam;
        }
        SM_TRY(m->d_fout.reserve(sizeof(double) * factored_out_doubles(m->nchan, fnblocks)));
        fa.out = m->d_fout.as<double>();
        SM_TRY(m->d_fscratch.reserve(64));
        unsigned long long *bits = m->d_fscratch.as<unsigned long long>();
        double *scale = reinterpret_cast<double *>(bits + 1);
        const int chunk = factored_chunk(m->nchan, fnt, n, fa.seg_steps);
        SM_TRY(m->d_fkbuf.reserve(factored_kbuf_bytes(m->nchan, fnt, chunk)));
        fa.kbuf = m->d_fkbuf.as<float2>();
        if (g_factored_kernel == 2) {
            // The hi and lo planes take 4 bytes each per phasor, half the FP32 buffer each.
            SM_TRY(m->d_fkhi.reserve(factored_kbuf_bytes(m->nchan, fnt, chunk) / 2));
            SM_TRY(m->d_fklo.reserve(factored_kbuf_bytes(m->nchan, fnt, chunk) / 2));
            fa.khi = m->d_fkhi.as<unsigned>();
            fa.klo = m->d_fklo.as<unsigned>();
            fa.planes_dtype = factored_dtype;
        }
        SM_TRY(m->d_fcohf.reserve(factored_cohf_bytes(m->nchan, chunk, nclass)));
        fa.cohf = m->d_fcohf.as<float4>();
        SM_TRY(launch_factored(factored_dtype, fa, chunk, coh, (size_t)n * m->nchan_pad * 8, bits, scale, fmap, nmap,
                               m->nbl, fblock_of, d_vis_fb, d_vis_fb_f64, 0, e0, e1));
    }
    return nullptr;
}

extern "C" const char *sm_model_timestep_with_device(SmModeller *m, double lst, double latitude, const UVW *d_uvws,
                                                     JonesF32 *d_vis_fb, JonesF64 *d_vis_fb_f64) {
    return model_timestep_device(m, lst, latitude, d_uvws, d_vis_fb, d_vis_fb_f64, -1);
}

extern "C" const char *sm_model_timestep_factored(SmModeller *m, double lst, double latitude, int32_t dtype,
                                                  JonesF32 *d_vis_fb, JonesF64 *d_vis_fb_f64) {
    if (dtype != SM_MFMA_FP16 && dtype != SM_MFMA_BF16) return "dtype must be SM_MFMA_FP16 or SM_MFMA_BF16";
    return model_timestep_device(m, lst, latitude, nullptr, d_vis_fb, d_vis_fb_f64, dtype);
}

// The host-buffer form over model_timestep_device: copies in, models, copies out, masks.
static const char *model_timestep_host(SmModeller *m, double lst, double latitude, const std::vector<double> *uvw_xyz,
                                       UVW *uvws_out, JonesF32 *vis_fb, JonesF64 *vis_fb_f64) {
    if (m == nullptr || vis_fb == nullptr) return "null argument";
    const size_t nvis = (size_t)m->nchan * m->nbl;
    cross_uvws(uvw_xyz != nullptr ? *uvw_xyz : m->xyz, m->nt, lst - m->pc_ra, m->pc_dec, m->h_uvw);
    if (uvws_out != nullptr) std::memcpy(uvws_out, m->h_uvw.data(), sizeof(UVW) * m->h_uvw.size());
    if (nvis == 0) return nullptr;
    SM_TRY(m->d_vis32.reserve(sizeof(JonesF32) * nvis));
    SM_TRY(hipMemcpy(m->d_vis32.ptr, vis_fb, sizeof(JonesF32) * nvis, hipMemcpyHostToDevice));
    JonesF64 *d64 = nullptr;
    if (vis_fb_f64 != nullptr) {
        SM_TRY(m->d_vis64.reserve(sizeof(JonesF64) * nvis));
        SM_TRY(hipMemcpy(m->d_vis64.ptr, vis_fb_f64, sizeof(JonesF64) * nvis, hipMemcpyHostToDevice));
        d64 = m->d_vis64.as<JonesF64>();
    }
    if (const char *e = model_timestep_device(m, lst, latitude, nullptr, m->d_vis32.as<JonesF32>(), d64, -1, uvw_xyz))
        return e;
    SM_TRY(hipMemcpy(vis_fb, m->d_vis32.ptr, sizeof(JonesF32) * nvis, hipMemcpyDeviceToHost));
    if (vis_fb_f64 != nullptr)
        SM_TRY(hipMemcpy(vis_fb_f64, m->d_vis64.ptr, sizeof(JonesF64) * nvis, hipMemcpyDeviceToHost));
    mask_pols(m->pols, vis_fb, vis_fb_f64, nvis);
    return nullptr;
}

extern "C" const char *sm_model_timestep_with(SmModeller *m, double lst, double latitude, UVW *uvws_out,
                                              JonesF32 *vis_fb, JonesF64 *vis_fb_f64) {
    return model_timestep_host(m, lst, latitude, nullptr, uvws_out, vis_fb, vis_fb_f64);
}

// get_lst_uvws_latitude (src/model/cpu.rs:731-769): the LST and latitude a timestamp models
// at, and the tile XYZs its cross UVWs come from: with precession, the J2000 coordinates of
// the zenith of date and the precessed XYZs (kept in m->xyz_j2000).
static const char *lst_latitude_xyz(SmModeller *m, double gps_seconds, double &lst, double &latitude,
                                    const std::vector<double> *&xyz) {
    if (!std::isfinite(gps_seconds)) return "non-finite timestamp";
    if (!m->precess) {
        lst = sm_get_lmst(m->lon, gps_seconds, m->dut1);
        latitude = m->lat;
        xyz = &m->xyz;
        return nullptr;
    }
    SmPrecessionInfo info;
    if (const char *e = sm_precess_time(m->lon, m->lat, m->pc_ra, m->pc_dec, gps_seconds, m->dut1, &info)) return e;
    m->xyz_j2000.resize(m->xyz.size());
    if (const char *e = sm_precess_xyz(&info, reinterpret_cast<const XYZ *>(m->xyz.data()),
                                       reinterpret_cast<XYZ *>(m->xyz_j2000.data()), m->nt))
        return e;
    lst = info.lmst_j2000;
    latitude = info.array_latitude_j2000;
    xyz = &m->xyz_j2000;
    return nullptr;
}

extern "C" const char *sm_lst_uvws_latitude(SmModeller *m, double gps_seconds, double *lst, double *latitude,
                                            UVW *uvws_out) {
    if (m == nullptr) return "null argument";
    double l = 0.0, b = 0.0;
    const std::vector<double> *xyz = nullptr;
    if (const char *e = lst_latitude_xyz(m, gps_seconds, l, b, xyz)) return e;
    if (lst != nullptr) *lst = l;
    if (latitude != nullptr) *latitude = b;
    if (uvws_out != nullptr && m->nbl > 0) {
        std::vector<UVW> u;
        cross_uvws(*xyz, m->nt, l - m->pc_ra, m->pc_dec, u);
        std::memcpy(uvws_out, u.data(), sizeof(UVW) * u.size());
    }
    return nullptr;
}

extern "C" const char *sm_model_timestep_epoch_with(SmModeller *m, double gps_seconds, UVW *uvws_out,
                                                    JonesF32 *vis_fb, JonesF64 *vis_fb_f64) {
    if (m == nullptr || vis_fb == nullptr) return "null argument";
    double lst = 0.0, lat = 0.0;
    const std::vector<double> *xyz = nullptr;
    if (const char *e = lst_latitude_xyz(m, gps_seconds, lst, lat, xyz)) return e;
    return model_timestep_host(m, lst, lat, xyz, uvws_out, vis_fb, vis_fb_f64);
}

extern "C" const char *sm_model_timestep_epoch_with_device(SmModeller *m, double gps_seconds, JonesF32 *d_vis_fb,
                                                           JonesF64 *d_vis_fb_f64) {
    if (m == nullptr || d_vis_fb == nullptr) return "null argument";
    double lst = 0.0, lat = 0.0;
    const std::vector<double> *xyz = nullptr;
    if (const char *e = lst_latitude_xyz(m, gps_seconds, lst, lat, xyz)) return e;
    return model_timestep_device(m, lst, lat, nullptr, d_vis_fb, d_vis_fb_f64, -1, xyz);
}

extern "C" const char *sm_model_timestep_autos_epoch_with(SmModeller *m, double gps_seconds, JonesF32 *vis_fb,
                                                          JonesF64 *vis_fb_f64) {
    if (m == nullptr || vis_fb == nullptr) return "null argument";
    double lst = 0.0, lat = 0.0;
    const std::vector<double> *xyz = nullptr;
    if (const char *e = lst_latitude_xyz(m, gps_seconds, lst, lat, xyz)) return e;
    return sm_model_timestep_autos_with(m, lst, lat, vis_fb, vis_fb_f64);
}

extern "C" const char *sm_model_timestep_autos_with(SmModeller *m, double lst, double latitude, JonesF32 *vis_fb,
                                                    JonesF64 *vis_fb_f64) {
    if (m == nullptr || vis_fb == nullptr) return "null argument";
    const size_t nvis = (size_t)m->nchan * m->nt;
    SM_TRY(m->d_vis32.reserve(sizeof(JonesF32) * nvis));
    SM_TRY(hipMemcpy(m->d_vis32.ptr, vis_fb, sizeof(JonesF32) * nvis, hipMemcpyHostToDevice));
    JonesF64 *d64 = nullptr;
    if (vis_fb_f64 != nullptr) {
        SM_TRY(m->d_vis64.reserve(sizeof(JonesF64) * nvis));
        SM_TRY(hipMemcpy(m->d_vis64.ptr, vis_fb_f64, sizeof(JonesF64) * nvis, hipMemcpyHostToDevice));
        d64 = m->d_vis64.as<JonesF64>();
    }
    // One demotion per component type, as SkyModellerCpu does (src/model/cpu.rs:906-932).
    for (int ty = 0; ty < 3; ++ty) {
        DevType &d = m->types[ty];
        const int n = d.h.n();
        if (n == 0) continue;
        BeamDesc b;
        if (const char *e = prepare_beam(m, d, lst, latitude, b)) return e;
        SM_TRY(d.coh.reserve(sizeof(double) * 8 * (size_t)n * m->nchan_pad));
        TypeInputs in = type_inputs(d);
        BeamDesc none = b;
        none.jones = nullptr;
        SM_TRY(launch_coherency(n, m->nchan, m->nchan_pad, m->d_freqs.as<double>(), in.fb, none, false,
                                d.coh.as<double>(), 0));
        b.uniform_row = -1;
        SM_TRY(launch_autos(m->nt, m->nchan, n, d.coh.as<double>(), m->nchan_pad, b, false, m->d_vis32.as<JonesF32>(),
                            d64, nullptr, true, 0));
    }
    SM_TRY(hipMemcpy(vis_fb, m->d_vis32.ptr, sizeof(JonesF32) * nvis, hipMemcpyDeviceToHost));
    if (vis_fb_f64 != nullptr)
        SM_TRY(hipMemcpy(vis_fb_f64, m->d_vis64.ptr, sizeof(JonesF64) * nvis, hipMemcpyDeviceToHost));
    mask_pols(m->pols, vis_fb, vis_fb_f64, nvis);
    return nullptr;
}

// Folds the completed event pairs into the per-type totals.
static const char *resolve_timing(SmModeller *m) {
    for (auto &t : m->pending) {
        SM_TRY(hipEventSynchronize(t.e1));
        float ms = 0.0f;
        SM_TRY(hipEventElapsedTime(&ms, t.e0, t.e1));
        m->ms_type[t.type] += ms;
        m->ms_total += ms;
        if (t.type == 3)
            m->launches_factored++;
        else
            m->launches++;
        m->pool.push_back(t.e0);
        m->pool.push_back(t.e1);
    }
    m->pending.clear();
    return nullptr;
}

extern "C" const char *sm_get_info(SmModeller *m, SmInfo *info) {
    if (m == nullptr || info == nullptr) return "null argument";
    if (const char *e = resolve_timing(m)) return e;
    info->num_baselines = m->nbl;
    info->num_freqs = m->nchan;
    info->num_unflagged_tiles = m->nt;
    info->num_unique_tiles = m->n_utile;
    info->num_points = m->types[0].h.n();
    info->num_gaussians = m->types[1].h.n();
    info->num_shapelets = m->types[2].h.n();
    info->num_kernel_launches = m->launches;
    info->kernel_ms = m->ms_total;
    info->kernel_ms_points = m->ms_type[0];
    info->kernel_ms_gaussians = m->ms_type[1];
    info->kernel_ms_shapelets = m->ms_type[2];
    info->kernel_ms_factored = m->ms_type[3];
    info->num_factored_launches = m->launches_factored;
    return nullptr;
}

extern "C" const char *sm_get_flux_params(SmModeller *m, int32_t comp_type, int32_t flux_type, int32_t index,
                                          int32_t chan, JonesF64 *fd, double *si, double *q) {
    if (m == nullptr || fd == nullptr) return "null argument";
    if (comp_type < SM_POINT || comp_type > SM_SHAPELET) return "unknown component type";
    const DevType &d = m->types[comp_type];
    const HostType &h = d.h;
    const JonesF64 *src = nullptr;
    const double *d_si = nullptr, *d_q = nullptr;
    size_t at = (size_t)index;
    if (flux_type == SM_POWER_LAW) {
        if (index < 0 || index >= h.n_pl) return "index outside the power-law bucket";
        src = d.pl_fds.as<JonesF64>();
        d_si = d.pl_sis.as<double>();
    } else if (flux_type == SM_CURVED_POWER_LAW) {
        if (index < 0 || index >= h.n_cpl) return "index outside the curved-power-law bucket";
        src = d.cpl_fds.as<JonesF64>();
        d_si = d.cpl_sis.as<double>();
        d_q = d.cpl_qs.as<double>();
    } else if (flux_type == SM_LIST) {
        if (index < 0 || index >= h.n_list) return "index outside the list bucket";
        if (chan < 0 || chan >= m->nchan) return "channel out of range";
        src = d.list_fds.as<JonesF64>();
        at = (size_t)chan * h.n_list + index;
    } else {
        return "unknown flux-density type";
    }
    // Read back from the device buffers the kernels use.
    SM_TRY(hipMemcpy(fd, src + at, sizeof(JonesF64), hipMemcpyDeviceToHost));
    if (si != nullptr) *si = 0.0;
    if (q != nullptr) *q = 0.0;
    if (si != nullptr && d_si != nullptr) SM_TRY(hipMemcpy(si, d_si + index, sizeof(double), hipMemcpyDeviceToHost));
    if (q != nullptr && d_q != nullptr) SM_TRY(hipMemcpy(q, d_q + index, sizeof(double), hipMemcpyDeviceToHost));
    return nullptr;
}

extern "C" const char *sm_set_beam_surrogate(SmModeller *m, SmSurrogate *s, double threshold,
                                             int32_t *fallbacks_so_far) {
    if (m == nullptr) return "null argument";
    if (fallbacks_so_far != nullptr) *fallbacks_so_far = m->surrogate_fallbacks;
    if (s != nullptr && m->beam_type != SM_BEAM_ANALYTIC) return "the surrogate replaces the analytic beam only";
    m->surrogate = s;
    m->surrogate_threshold = threshold;
    m->surrogate_fallbacks = 0;
    return nullptr;
}

extern "C" const char *sm_set_timing(SmModeller *m, int32_t enabled) {
    if (m == nullptr) return "null argument";
    if (const char *e = resolve_timing(m)) return e;
    m->timing = enabled != 0;
    m->launches = 0;
    m->ms_total = 0;
    m->ms_type[0] = m->ms_type[1] = m->ms_type[2] = m->ms_type[3] = 0;
    m->launches_factored = 0;
    return nullptr;
}

extern "C" const char *sm_synchronize(void) {
    SM_TRY(hipDeviceSynchronize());
    return nullptr;
}
