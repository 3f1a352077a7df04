/*
 * CPU oracle for the sky-model visibility predictor.
 *
 * TEST INFRASTRUCTURE ONLY. This library is a plain-C restatement of
 * mwa_hyperdrive's `SkyModellerCpu` (src/model/cpu.rs) and the arithmetic it
 * calls (flux-density estimation, LMN/UVW/az-el geometry, the shapelet basis
 * table). It is the checker for the HIP product path; only tests, the smoke
 * check and the CPU-baseline leg of the benchmark may load it. Nothing in the
 * product links or calls it.
 */
#pragma once

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One flux density at one frequency (src/srclist/types/flux_density/mod.rs:20-41). */
typedef struct OrcFluxDensity {
    double freq; /* Hz */
    double i, q, u, v; /* Jy */
} OrcFluxDensity;

/* Shapelet coefficient, same layout as the FFI `ShapeletCoeff` (src/gpu/types.h:105-109). */
typedef struct OrcShapeletCoeff {
    double value;
    uint8_t n1;
    uint8_t n2;
} OrcShapeletCoeff;

enum { ORC_POINT = 0, ORC_GAUSSIAN = 1, ORC_SHAPELET = 2 };
enum { ORC_LIST = 0, ORC_POWER_LAW = 1, ORC_CURVED_POWER_LAW = 2 };

/* One sky-model component (src/srclist/types/components/mod.rs:21-31,58-97). */
typedef struct OrcComponent {
    double ra, dec;          /* radians */
    int32_t comp_type;       /* ORC_POINT / ORC_GAUSSIAN / ORC_SHAPELET */
    int32_t flux_type;       /* ORC_LIST / ORC_POWER_LAW / ORC_CURVED_POWER_LAW */
    double maj, min, pa;     /* radians */
    double si, q;            /* power-law spectral index and curvature */
    const OrcFluxDensity *fds; /* power laws use fds[0]; lists are sorted by freq */
    int32_t num_fds;
    int32_t num_coeffs;
    const OrcShapeletCoeff *coeffs;
} OrcComponent;

enum { ORC_BEAM_NONE = 0, ORC_BEAM_ANALYTIC = 1 };
enum { ORC_POLS_ALL = 0, ORC_POLS_XX = 1, ORC_POLS_YY = 2, ORC_POLS_XX_YY = 3, ORC_POLS_XX_YY_XY = 4 };

/* Observation description: the arguments of SkyModellerCpu::new (src/model/cpu.rs:73-85)
 * minus the Epoch machinery (LST and latitude are passed directly). */
typedef struct OrcObs {
    int32_t num_unflagged_tiles;
    const double *unflagged_xyzs;    /* [num_unflagged_tiles][3] geodetic XYZ metres */
    int32_t num_flagged_tiles;
    const int32_t *flagged_tiles;    /* total-array tile indices */
    int32_t num_freqs;
    const double *freqs;             /* unflagged fine-channel freqs, Hz */
    double phase_centre_ra, phase_centre_dec;
    double lst, latitude;            /* radians */
    int32_t beam_type;               /* ORC_BEAM_* */
    const uint32_t *dipole_delays;   /* [total_tiles][16] or NULL (all 0) */
    const double *dipole_gains;      /* [total_tiles][32] or NULL (all 1) */
    int32_t pols;                    /* ORC_POLS_* */
    /* Tile XYZs the cross UVWs come from, or NULL for unflagged_xyzs. With precession the
     * reference forms the cross UVWs from the precessed XYZs and the shapelet UVs from the
     * unprecessed ones (src/model/cpu.rs:741-785). */
    const double *uvw_xyzs;
} OrcObs;

/* --- geometry (marlu 0.17.1 restatements; call sites src/model/cpu.rs:767, components/mod.rs:214,394) --- */
void orc_radec_to_lmn_rime(double ra, double dec, double pc_ra, double pc_dec, double out_lmn[3]);
void orc_xyzs_to_cross_uvws(int32_t n, const double *xyzs, double ha, double dec, double *out_uvws);
void orc_hadec_to_azel(double ha, double dec, double lat, double *az, double *el);

/* --- flux densities (src/srclist/types/flux_density/mod.rs:158-275) --- */
void orc_estimate_at_freq(const OrcComponent *c, double freq, OrcFluxDensity *out);
/* CPL spectral index re-referenced to 150 MHz (src/model/gpu.rs:664-680). */
double orc_cpl_si_at_150mhz(const OrcComponent *c);

/* --- shapelet basis table (src/model/make_shapelet_basis_values.py:12-34), 101 x 10001 f64 --- */
void orc_shapelet_basis_table(double *out);
const double *orc_shapelet_basis_table_cached(void);

/* --- analytic MWA-like beam: the CPU restatement of the product's beam --- */
void orc_analytic_beam_jones(double az, double za, double freq_hz, const uint32_t delays[16],
                             const double gains[32], double out_jones[8]);

/* --- modelling (src/model/cpu.rs:249-352,371-495,518-726,787-812,828-936) ---
 * `vis_f32` [num_freqs][num_baselines][8] is accumulated into exactly as the reference does
 * (per component type: vis += (f32) f64_accum). `vis_f64` (nullable) receives the sum of
 * the f64 accumulators. `comp_type_mask` selects which types to model (bit 0 points, bit 1
 * Gaussians, bit 2 shapelets). Returns 0 on success. */
int orc_model_timestep(const OrcObs *obs, const OrcComponent *comps, const int32_t *source_offsets,
                       int32_t num_sources, int32_t comp_type_mask, float *vis_f32, double *vis_f64,
                       int32_t num_threads);

/* Same, but only over a subset of channels [chan_start, chan_start + chan_count) and with the
 * supplied uvws (used by the CPU baseline timer on a bounded slice). */
int orc_model_timestep_slice(const OrcObs *obs, const OrcComponent *comps, const int32_t *source_offsets,
                             int32_t num_sources, int32_t chan_start, int32_t chan_count, float *vis_f32,
                             double *vis_f64, int32_t num_threads);

/* Same as the slice, for the listed baselines only: vis [chan_count][num_listed][8], row k
 * holding baseline baselines[k] (the numbering of src/math/mod.rs:96-109). Used to check
 * full-size skies on a bounded sample of baselines. */
int orc_model_timestep_baselines(const OrcObs *obs, const OrcComponent *comps, const int32_t *source_offsets,
                                 int32_t num_sources, int32_t chan_start, int32_t chan_count,
                                 const int32_t *baselines, int32_t num_listed, float *vis_f32, double *vis_f64,
                                 int32_t num_threads);

int orc_model_timestep_autos(const OrcObs *obs, const OrcComponent *comps, const int32_t *source_offsets,
                             int32_t num_sources, float *vis_f32, double *vis_f64, int32_t num_threads);

/* Unique-tile / unique-frequency maps of SkyModellerCpu::new (src/model/cpu.rs:105-161).
 * Returns the number of unique tiles; fills tile_map[total_tiles]. */
int32_t orc_unique_tiles(int32_t total_tiles, const int32_t *flagged, int32_t num_flagged,
                         const uint32_t *delays, const double *gains, int32_t *tile_map);

/* Baseline index -> tile pair over unflagged tiles (src/math/mod.rs:96-109). */
void orc_baseline_to_tiles(int32_t num_tiles, int32_t bl, int32_t *t1, int32_t *t2);

#ifdef __cplusplus
}
#endif
