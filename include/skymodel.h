/*
 * C ABI of the MI355X (gfx950) sky-model visibility predictor.
 *
 * Three layers, all returning NULL on success or a static error string (the
 * reference's convention, src/gpu/common.cuh:552-561 and src/gpu/mod.rs:74-98):
 *
 *  1. Kernel-level drop-ins with the reference's exact signatures
 *     (src/gpu/model.h:33-56). Every pointer inside the structs is device memory,
 *     visibilities are accumulated into (+=) as JonesF32 [freq][baseline].
 *  2. f64 variants of the same: identical inputs plus a nullable JonesF64 output
 *     that receives the f64 accumulators before demotion (the parity path).
 *  3. A handle-level SkyModeller (src/model/mod.rs:123-204, src/model/gpu.rs:150-1527):
 *     construction from a source list, per-timestep modelling from a timestamp (GPS
 *     seconds, with the DUT1 and precession setting of the handle, as the trait's Epoch
 *     methods do through marlu, src/model/cpu.rs:731-769) or from (LST, latitude),
 *     update_with_a_source, autos.
 *
 * All device work is enqueued on the HIP null stream of the calling thread's current
 * device; entry points that take host buffers synchronise before returning.
 */
#pragma once

#include "skymodel_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ 1. kernel drop-ins */

/* Replaces `model_points` (src/gpu/model.h:33-34; launched at src/model/gpu.rs:913-945).
 * d_beam_jones is [num_unique_tiles][num_unique_beam_freqs][directions] with directions
 * ordered power law, curved power law, list (src/model/gpu.rs:889-910). */
const char *model_points(const Points *comps, const Addresses *a, const UVW *d_uvws, const JonesF64 *d_beam_jones,
                         JonesF32 *d_vis_fb);

/* Replaces `model_gaussians` (src/gpu/model.h:41-42). */
const char *model_gaussians(const Gaussians *comps, const Addresses *a, const UVW *d_uvws,
                            const JonesF64 *d_beam_jones, JonesF32 *d_vis_fb);

/* Replaces `model_shapelets` (src/gpu/model.h:49-50). The shapelet UVs inside `comps`
 * are [baseline][component] metres, as the reference's host computes them. */
const char *model_shapelets(const Shapelets *comps, const Addresses *a, const UVW *d_uvws,
                            const JonesF64 *d_beam_jones, JonesF32 *d_vis_fb);

/* Replaces `model_autos` (src/gpu/model.h:56). vis_fb is [freq][tile]. */
const char *model_autos(int num_tiles, int num_freqs, const double *freqs, const Points *points,
                        const Gaussians *gaussians, const Shapelets *shapelets, const JonesF64 *beam_jones,
                        const int *tile_map, const int *freq_map, int num_fee_freqs,
                        const int *tile_index_to_unflagged_tile_index_map, JonesF32 *vis_fb);

/* ------------------------------------------------------------------ 2. f64 variants
 * Same as the drop-ins; d_vis_fb_f64 (nullable, device, [freq][baseline]) is accumulated
 * with the un-demoted f64 sums. No counterpart in the reference: it exposes the f64
 * accumulator that SkyModellerCpu keeps before `Jones::from` (src/model/cpu.rs:335-348). */
const char *sm_model_points_f64(const Points *comps, const Addresses *a, const UVW *d_uvws,
                                const JonesF64 *d_beam_jones, JonesF32 *d_vis_fb, JonesF64 *d_vis_fb_f64);
const char *sm_model_gaussians_f64(const Gaussians *comps, const Addresses *a, const UVW *d_uvws,
                                   const JonesF64 *d_beam_jones, JonesF32 *d_vis_fb, JonesF64 *d_vis_fb_f64);
const char *sm_model_shapelets_f64(const Shapelets *comps, const Addresses *a, const UVW *d_uvws,
                                   const JonesF64 *d_beam_jones, JonesF32 *d_vis_fb, JonesF64 *d_vis_fb_f64);

/* No counterpart in the reference. For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving or peel loops, harnesses that forbid hipMalloc): the same
 * inputs, outputs and results as sm_model_points_f64, bit for bit, with every work buffer
 * (padded frequencies, tile UVWs, tile-group pairs, beam rows, LMN and the coherency table)
 * taken from one fixed 64 MiB device array per GPU instead of hipMalloc. It calls no
 * allocating routine. A call whose work buffers exceed the array (about 64 B per component
 * and padded channel, plus 8 B per baseline with per-tile beams) returns an error. Calls on one
 * GPU share the array: calls that overlap on the same GPU (other threads, other streams) race. */
const char *sm_model_points_noalloc(const Points *comps, const Addresses *a, const UVW *d_uvws,
                                    const JonesF64 *d_beam_jones, JonesF32 *d_vis_fb, JonesF64 *d_vis_fb_f64);

/* ------------------------------------------------------------------ 3. SkyModeller handle */

enum { SM_BEAM_NONE = 0, SM_BEAM_ANALYTIC = 1 };
/* Polarisations (src/context/mod.rs:26-33), used by mask_pols (src/model/mod.rs:264-285). */
enum { SM_POLS_XX_XY_YX_YY = 0, SM_POLS_XX = 1, SM_POLS_YY = 2, SM_POLS_XX_YY = 3, SM_POLS_XX_YY_XY = 4 };

/* Arguments of new_sky_modeller (src/model/mod.rs:215-227). */
typedef struct SmObsParams {
    int32_t num_unflagged_tiles;
    const double *unflagged_tile_xyzs; /* host [num_unflagged_tiles][3], geodetic metres */
    int32_t num_flagged_tiles;
    const int32_t *flagged_tiles;      /* host, total-array tile indices */
    int32_t num_freqs;
    const double *freqs;               /* host, unflagged fine-channel frequencies, Hz */
    double phase_centre_ra, phase_centre_dec;
    double array_longitude, array_latitude;
    int32_t beam_type;                 /* SM_BEAM_* */
    const uint32_t *dipole_delays;     /* host [total tiles][16] or NULL (all 0) */
    const double *dipole_gains;        /* host [total tiles][32] or NULL (all 1) */
    int32_t pols;                      /* SM_POLS_* */
    double dut1;                       /* UT1 - UTC [s] for the timestamp entry points */
    int32_t apply_precession;          /* nonzero: timestamps are precessed to J2000 */
} SmObsParams;

typedef struct SmModeller SmModeller;

/* Replaces new_sky_modeller / SkyModellerGpu::new (src/model/mod.rs:215-261,
 * src/model/gpu.rs:150-519). comps are in source-list order; source_offsets has
 * num_sources + 1 entries. The source order is reversed internally, as the reference does
 * (src/model/gpu.rs:230-241). */
const char *sm_new(const SmObsParams *params, const SmComponent *comps, const int32_t *source_offsets,
                   int32_t num_sources, SmModeller **out);

/* Frees the modeller and its device memory. */
void sm_free(SmModeller *m);

/* Replaces SkyModeller::update_with_a_source (src/model/mod.rs:185-193,
 * src/model/gpu.rs:1498-1504): the sky becomes this one source's components, modelled
 * about the new phase centre. */
const char *sm_update_with_a_source(SmModeller *m, const SmComponent *comps, int32_t num_comps, double pc_ra,
                                    double pc_dec);

/* Replaces SkyModeller::model_timestep_with (src/model/mod.rs:150-163) for a given LST and
 * latitude (no precession). vis_fb: host [num_freqs][num_baselines] JonesF32, accumulated
 * into; vis_fb_f64 (nullable, host) receives the f64 sums. uvws_out (nullable, host)
 * receives the baseline UVWs in metres. mask_pols is applied to vis_fb afterwards. */
const char *sm_model_timestep_with(SmModeller *m, double lst, double latitude, UVW *uvws_out, JonesF32 *vis_fb,
                                   JonesF64 *vis_fb_f64);

/* Device-buffer form used by device-resident callers (src/model/gpu.rs:1152-1166,
 * src/params/peel/gpu.rs:447-459): d_vis_fb / d_vis_fb_f64 are device memory; the call
 * returns once the work is enqueued. d_uvws (nullable) overrides the UVWs computed from the
 * modeller's tile positions. No mask_pols, like the reference's device method. */
const char *sm_model_timestep_with_device(SmModeller *m, double lst, double latitude, const UVW *d_uvws,
                                          JonesF32 *d_vis_fb, JonesF64 *d_vis_fb_f64);

/* Replaces SkyModeller::model_timestep_autos_with (src/model/mod.rs:195-203). vis_fb: host
 * [num_freqs][num_unflagged_tiles]. */
const char *sm_model_timestep_autos_with(SmModeller *m, double lst, double latitude, JonesF32 *vis_fb,
                                         JonesF64 *vis_fb_f64);

/* Sizes and the accumulated device time of the modelling kernels (HIP events on the
 * launch stream). No counterpart in the reference. */
typedef struct SmInfo {
    int32_t num_baselines;
    int32_t num_freqs;
    int32_t num_unflagged_tiles;
    int32_t num_unique_tiles;
    int32_t num_points, num_gaussians, num_shapelets;
    int32_t num_kernel_launches; /* direct modelling kernels timed since the last reset */
    double kernel_ms;            /* their summed duration */
    double kernel_ms_points, kernel_ms_gaussians, kernel_ms_shapelets;
    double kernel_ms_factored;   /* split-precision MFMA point kernel */
    int32_t num_factored_launches;
} SmInfo;

/* ------------------------------------------------------------------ split-precision points
 * No counterpart in the reference. Models one timestep like sm_model_timestep_with_device,
 * but point components go through the antenna-factorised MFMA path: per channel
 * V^e = (K (.) C^e) K^H with K_pc = exp(i 2 pi (u_p l + v_p m + w_p (n - 1)) f / c), phases and
 * their range reduction in FP64, operands split into hi + lo halves of `dtype`, products on
 * MFMA with FP32 accumulation promoted to FP64 every 1024 components. Gaussians and
 * shapelets run in FP32 with their phases, phase reductions, exponents and table positions in
 * FP64. With per-tile beams, every baseline with a tile of the most common beam row takes the
 * MFMA path (row coherency J_r S J_0^H), and the points of pairs of two tiles of other rows
 * the own-beam pair kernel in FP32 with FP64 phases. */
enum { SM_MFMA_FP16 = 0, SM_MFMA_BF16 = 1 };
const char *sm_model_timestep_factored(SmModeller *m, double lst, double latitude, int32_t dtype,
                                       JonesF32 *d_vis_fb, JonesF64 *d_vis_fb_f64);

/* FP32 accumulation span of the split-precision kernel, in steps of 16 components (default
 * 64 = 1024 components): after each span the partial sums are added into FP64. Process-wide. */
const char *sm_set_factored_segment(int32_t steps);

/* Work-group organisation of the split-precision kernel. 1 (default): wave-specialised, four
 * waves issue only MFMAs while eight generate the next step's operands; 0: four waves that
 * each do both. Both compute the same operands and products in the same order, so their
 * results agree bit for bit. Process-wide. */
const char *sm_set_factored_kernel(int32_t kernel);

/* Self-test of the MFMA lane layout the factored kernel assumes: D (32 x 32 f32, row-major)
 * = A (32 x 16 f16, row-major) . B (16 x 32 f16, row-major), one wave, device pointers. */
const char *sm_selftest_mfma_layout(const void *d_a, const void *d_b, float *d_d);
const char *sm_get_info(SmModeller *m, SmInfo *info);
/* Enables (1) or disables (0) per-kernel HIP-event timing and resets the counters. */
const char *sm_set_timing(SmModeller *m, int32_t enabled);
const char *sm_synchronize(void);
/* Kernel choice for the direct path: 0 (default) picks the fastest kernel for the case
 * (tile-pair phasor kernel for points and Gaussians with evenly spaced channels and one
 * shared beam, channel-per-lane kernels for shapelets, two channels per lane with a shared
 * beam); 1 forces the baseline-per-lane
 * kernels everywhere. Process-wide; used to compare the two. */
const char *sm_set_direct_variant(int32_t variant);

/* Limit in bytes on the per-beam-row-class coherency tables the per-tile-beam paths build
 * (J_r S J_0^H for every class; 64 B per class, component and padded channel). Above it, or
 * when the allocation fails, those paths form the coherency per baseline segment instead,
 * with identical results. Negative (default): the device's free memory less 1 GiB. 0 always
 * takes the per-segment path. Process-wide. No counterpart in the reference. */
const char *sm_set_class_table_limit(int64_t bytes);

/* Reads back the flux parameters the modeller holds on the device for one component of a
 * bucket (type x flux kind, components in the modeller's order): power laws and curved power
 * laws as the instrumental flux density at 150 MHz with the spectral index (and curvature q),
 * the curved-power-law index re-referenced to 150 MHz; lists as the instrumental flux density
 * estimated at channel `chan`. si and q are nullable. The reference's regression test of the
 * re-reference reads the same device buffers (src/model/tests/gpu.rs:558-601). */
const char *sm_get_flux_params(SmModeller *m, int32_t comp_type, int32_t flux_type, int32_t index, int32_t chan,
                               JonesF64 *fd, double *si, double *q);

/* ------------------------------------------------------------------ timestamps
 * The SkyModeller methods take a hifitime::Epoch (src/model/mod.rs:141-203); these take its
 * GPS seconds. The modeller turns a timestamp into (LST, cross UVWs, latitude) as
 * get_lst_uvws_latitude does (src/model/cpu.rs:731-769, src/model/gpu.rs:1357-1398), with
 * the dut1 and apply_precession of SmObsParams (new_sky_modeller's, src/model/mod.rs:215-227).
 * With precession the LST and latitude are the J2000 coordinates of the zenith of date, the
 * cross UVWs come from the precessed tile XYZs at the J2000 phase centre, and the shapelet
 * UVs from the unprecessed XYZs at that LST, as the reference computes them
 * (src/model/cpu.rs:782-785). */

/* marlu 0.17.1 precession::PrecessionInfo, the fields the reference reads. rotation_matrix
 * takes vectors in the true equator and equinox of date to J2000; hadec_j2000 is the
 * aberrated phase centre at lmst_j2000. */
typedef struct SmPrecessionInfo {
    double rotation_matrix[3][3];
    HADec hadec_j2000;
    double lmst, lmst_j2000, array_latitude_j2000;
} SmPrecessionInfo;

/* marlu precession::get_lmst (src/model/cpu.rs:752): local mean sidereal time [rad], IAU 2006
 * GMST at UT1 = UTC + dut1 plus the longitude. */
double sm_get_lmst(double array_longitude_rad, double gps_seconds, double dut1_seconds);
/* marlu precession::precess_time (src/model/cpu.rs:733-739). */
const char *sm_precess_time(double array_longitude_rad, double array_latitude_rad, double pc_ra, double pc_dec,
                            double gps_seconds, double dut1_seconds, SmPrecessionInfo *out);
/* PrecessionInfo::precess_xyz (src/model/cpu.rs:741): host XYZs rotated into the J2000 frame. */
const char *sm_precess_xyz(const SmPrecessionInfo *info, const XYZ *xyzs, XYZ *out, int32_t n);

/* get_lst_uvws_latitude (src/model/cpu.rs:731-769) for the modeller's settings: the LST and
 * latitude the timestamp models at, and (nullable, host) the cross UVWs in metres. */
const char *sm_lst_uvws_latitude(SmModeller *m, double gps_seconds, double *lst, double *latitude, UVW *uvws_out);
/* SkyModeller::model_timestep_with(Epoch) (src/model/mod.rs:156-163, src/model/cpu.rs:796-823):
 * like sm_model_timestep_with, at a timestamp. */
const char *sm_model_timestep_epoch_with(SmModeller *m, double gps_seconds, UVW *uvws_out, JonesF32 *vis_fb,
                                         JonesF64 *vis_fb_f64);
/* The device-buffer form (src/model/gpu.rs:1476-1496 without the copies): d_vis_fb /
 * d_vis_fb_f64 are device memory, accumulated into; returns once the work is enqueued. */
const char *sm_model_timestep_epoch_with_device(SmModeller *m, double gps_seconds, JonesF32 *d_vis_fb,
                                                JonesF64 *d_vis_fb_f64);
/* SkyModeller::model_timestep_autos_with(Epoch) (src/model/mod.rs:195-203). */
const char *sm_model_timestep_autos_epoch_with(SmModeller *m, double gps_seconds, JonesF32 *vis_fb,
                                               JonesF64 *vis_fb_f64);

/* ------------------------------------------------------------------ uvfits output
 * The vis-simulate writer path (src/io/write/mod.rs:113-260, src/params/vis_simulate.rs:241-300):
 * modelled visibilities written as AIPS random-groups uvfits, as marlu 0.17.1's UvfitsWriter
 * writes them for the reference. The writer forms each timestep's UVWs itself from the
 * un-precessed tile positions (src/io/write/mod.rs:79-90): precessed to J2000 at precess_time's
 * hadec_j2000, or at the LMST without precession. Host code: no device is needed. */
typedef struct SmUvfitsParams {
    int32_t num_timesteps;         /* groups in the file = num_timesteps x num_baselines */
    int32_t num_baselines;         /* rows per timestep */
    int32_t num_chans;
    int32_t num_tiles;             /* rows of the AN table */
    const char *const *tile_names; /* [num_tiles] */
    const double *tile_xyzs;       /* [num_tiles][3] un-precessed geodetic XYZ, metres */
    const int32_t *baseline_tiles; /* [num_baselines][2] tile indices, tile1 <= tile2 (equal: an auto) */
    double first_freq_hz, freq_res_hz;
    double phase_centre_ra, phase_centre_dec;
    double array_longitude, array_latitude, array_height;
    double dut1;                   /* seconds */
    int32_t apply_precession;
    double int_time_s;
    double first_gps_seconds;      /* sets the file's whole-day date reference */
    const char *obs_name;          /* nullable */
} SmUvfitsParams;
typedef struct SmUvfitsWriter SmUvfitsWriter;

/* Opens the file and writes the primary header. */
const char *sm_uvfits_create(const char *path, const SmUvfitsParams *params, SmUvfitsWriter **out);
/* One timestep's groups: vis_fb host [num_chans][num_baselines] JonesF32, weights_fb host
 * [num_chans][num_baselines] (nullable: all 1), at the timestep's centroid GPS time. */
const char *sm_uvfits_write_timestep(SmUvfitsWriter *w, double gps_seconds, const JonesF32 *vis_fb,
                                     const float *weights_fb);
/* Writes the AIPS AN table, closes the file and frees the writer (also after an error). */
const char *sm_uvfits_finish(SmUvfitsWriter *w);

/* ------------------------------------------------------------------ peel-loop neighbours
 * The device work around the modeller in the reference's peel loop
 * (src/params/peel/gpu.rs:441-470). Same signatures as src/gpu/peel.h:13-35 with
 * FLOAT = double; device pointers; visibilities [time][freq][baseline] JonesF32, UVWs
 * [time][baseline], XYZs [time][tile]. Work is enqueued on the null stream. */

/* Replaces xyzs_to_uvws (src/gpu/peel.h:13-14): baseline UVWs (metres) of the tile XYZs at
 * each timestep's LMST for the pointing centre. */
const char *xyzs_to_uvws(const XYZ *d_xyzs, const double *d_lmsts, UVW *d_uvws, RADec pointing_centre, int num_tiles,
                         int num_baselines, int num_timesteps);

/* Replaces rotate (src/gpu/peel.h:16-17): vis *= exp(-i 2 pi (w_to - w_from) / lambda), formed
 * in FP64 and rounded to FP32. */
const char *rotate(JonesF32 *d_vis_fb, const int num_timesteps, const int num_baselines, const int num_freqs,
                   const UVW *d_uvws_from, const UVW *d_uvws_to, const double *d_lambdas);

/* Replaces add_model (src/gpu/peel.h:33-35): residual += model * gain * exp(-i 2 pi (u alpha +
 * v beta) lambda) (no phase when alpha = beta = 0). */
const char *add_model(JonesF32 *d_vis_residual, const JonesF32 *d_vis_model, const IonoConsts iono_consts,
                      const double *d_lambdas_m, const UVW *d_uvws, const int num_timesteps, const int num_baselines,
                      const int num_freqs);

/* ------------------------------------------------------------------ beam */
/* Analytic MWA-like beam on the device: out[tile][freq][dir] JonesF64 for the given unique
 * tiles (delays [n_tiles][16], gains [n_tiles][32], host), frequencies (host, Hz) and
 * directions (device az/za, radians). Identity at zenith for an ideal tile. Stands in for
 * BeamGpu::calc_jones_pair (src/beam/mod.rs:143-177). */
const char *sm_analytic_beam(int32_t n_tiles, const uint32_t *delays, const double *gains, int32_t n_freqs,
                             const double *freqs, int32_t n_dirs, const double *d_az, const double *d_za,
                             JonesF64 *d_out);

/* ------------------------------------------------------------------ beam surrogate
 * No counterpart in the reference: an MLP stand-in for the beam Jones (BeamGpu::calc_jones_pair,
 * src/beam/mod.rs:143-177). 53 inputs (sin az, cos az, sin za, cos za, (f - 182 MHz) / 15 MHz,
 * delays / 31 [16], gains [32]) -> 4 x 256 SiLU -> 8 outputs (J re/im, row-major). Parameters
 * flat in the order W0 [256][53], b0 [256], W1 [256][256], b1, W2, b2, W3, b3, W4 [8][256], b4 [8]. */
typedef struct SmSurrogate SmSurrogate;
int32_t sm_surrogate_num_params(void);
const char *sm_surrogate_new(uint64_t seed, SmSurrogate **out);
/* A network of `depth` hidden layers (1..8) of `hidden` SiLU units (1..4096), for studies of
 * what accuracy a size reaches. Shapes other than 4 x 256 train and evaluate in FP32: the
 * FP16 forward (dtype 1) is the fused 4 x 256 kernel, and sm_surrogate_beam uses FP32 for
 * them. Parameters are flat like the default network's, layer by layer. */
const char *sm_surrogate_new_shaped(uint64_t seed, int32_t hidden, int32_t depth, SmSurrogate **out);
/* The network's hidden width, depth and parameter count. */
const char *sm_surrogate_shape(const SmSurrogate *s, int32_t *hidden, int32_t *depth, int64_t *num_params);
void sm_surrogate_free(SmSurrogate *s);
const char *sm_surrogate_get_params(SmSurrogate *s, float *host_params);
const char *sm_surrogate_set_params(SmSurrogate *s, const float *host_params);
/* Training samples from the exact (analytic) beam at seeded random inputs: az ~ U[0, 2 pi),
 * cos za ~ U[0, 1], f ~ U[167, 197] MHz, gains 0 with probability 0.02, and delays either
 * independent U{0..31} (mode 0) or those pointing the tile within 60 deg of zenith (mode 1).
 * Sample i is a pure function of (seed, mode, first + i). Device outputs (each nullable):
 * features [n][53] f32, Jones [n][8] f32, raw inputs [n][51] f64 (az, za, f, delays[16],
 * gains[32]). */
const char *sm_surrogate_samples(uint64_t seed, int64_t first, int32_t n, int32_t mode, float *d_x, float *d_y,
                                 double *d_raw);
/* Mean-squared-error loss and its gradients (host, flat like the parameters) on a batch. */
const char *sm_surrogate_gradients(SmSurrogate *s, int32_t n, const float *d_x, const float *d_y, double *loss,
                                   float *host_grads);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on a batch; loss is the batch MSE before it. */
const char *sm_surrogate_train_step(SmSurrogate *s, int32_t n, const float *d_x, const float *d_y, float lr,
                                    double *loss);
/* Forward pass: dtype 0 = FP32 (the training arithmetic), 1 = FP16 MFMA with FP32 accumulation. */
const char *sm_surrogate_forward(SmSurrogate *s, int32_t n, const float *d_x, int32_t dtype, float *d_y);
/* Beam Jones out[tile][freq][dir] like sm_analytic_beam, from the FP16 surrogate. Each tile is
 * judged against the exact beam on up to 512 directions x 8 frequencies; a tile whose max
 * |J_surrogate - J_exact| exceeds `threshold` takes the exact beam. n_fallback (nullable)
 * receives the number of such tiles. */
const char *sm_surrogate_beam(SmSurrogate *s, int32_t n_tiles, const uint32_t *delays, const double *gains,
                              int32_t n_freqs, const double *freqs, int32_t n_dirs, const double *d_az,
                              const double *d_za, double threshold, JonesF64 *d_out, int32_t *n_fallback);
/* Makes the modeller take its beam Jones from the surrogate (with the exact beam as per-tile
 * fallback above `threshold`), in place of the analytic beam; s = NULL restores the analytic
 * beam. fallbacks_so_far (nullable) receives the number of tile fallbacks since the last call.
 * The surrogate is not owned and must outlive its use. */
const char *sm_set_beam_surrogate(SmModeller *m, SmSurrogate *s, double threshold, int32_t *fallbacks_so_far);

#ifdef __cplusplus
} /* extern "C" */
#endif
