// Shared driver for one component type, used by the kernel-level drop-ins (abi.hip) and
// the SkyModeller handle (modeller.hip).
#pragma once

#include "sm_internal.h"

namespace sm {

// Device-resident description of one component type.
struct TypeInputs {
    FluxBuckets fb;
    const double4 *lmn = nullptr;
    const double4 *shape = nullptr;
    const double2 *bl_uv = nullptr;
    // What bl_uv is formed from, when known (the modeller): per component sin/cos of HA and
    // Dec [n][4], tile XYZs [ntiles][3] and the tiles of each baseline. The channel-per-lane
    // shapelet kernel then forms each UV in registers instead of reading bl_uv.
    const double *uv_hadec = nullptr;
    const double *uv_xyz = nullptr;
    const int2 *uv_bl_tiles = nullptr;
    const int2 *coeff_range = nullptr;
    const ShapeletCoeff *coeffs = nullptr;
    int coeff_count = 0;  // total coefficients behind coeffs
};

// Per-call modelling context.
struct ModelCtx {
    int nbl = 0, nchan = 0, nchan_pad = 0, ntiles = 0;
    const double *d_freqs_pad = nullptr;
    bool even = false;   // every chunk evenly spaced (always, with chunk tables)
    ChunkTabs chunks;    // channel chunks of evenly spaced runs, or null (fixed chunks)
    const UVW *d_uvw = nullptr;
    const double *d_sbf = nullptr;
    BeamDesc beam{};
    const double *d_tile_uvw = nullptr;  // [ntiles][3], enables the tile-pair kernel
    const int4 *d_pairs = nullptr;
    int npairs = 0;
    const int *d_bl_subset = nullptr;  // nullable: the direct kernels model only these baselines
    int nsub = 0;
    // Per-tile beams, optional: beam-row classes (class 0 = row class_row[0]); the per-tile
    // kernels then read J_r S J_0^H from a class table for baselines with a class-0 tile.
    int nclass = 0, class_row0 = -1;     // class_row0 = the beam row of class 0
    // Per-tile beams, optional: pairs of two class-0 tiles share J_0 S J_0^H and take the
    // tile-pair kernel on the class-0 tiles; every other baseline takes the per-tile kernel.
    // Baselines with a tile outside class 0, in segments: for each class r = 1 .. nclass-1,
    // first those whose first tile is in class r and second in class 0 (coherency
    // J_r S J_0^H, shared by the segment), then the reverse (J_0 S J_r^H); last, the pairs of
    // two tiles outside class 0 (per-tile J_p S J_q^H). h_seg_off (host) holds the
    // 2 (nclass - 1) + 2 segment offsets into d_notg0_bl, h_class_row the beam row per class.
    int g0_n = 0, g0_npairs = 0;
    const int *d_g0_tiles = nullptr;    // [g0_n] class-0 array tiles, ascending
    const double *d_g0_uvw = nullptr;   // [g0_n][3] their tile UVWs
    const int4 *d_g0_pairs = nullptr;   // tile-group pairs over the class-0 tiles
    const int *d_notg0_bl = nullptr;
    // Array tiles outside class 0 and their classes (pairs with a class-0 tile take the
    // row-by-column kernel).
    const int *d_other_tiles = nullptr;
    const int *d_other_cls = nullptr;
    int n_other = 0;
    const int *h_seg_off = nullptr;
    const int *h_class_row = nullptr;
    const int *d_class_row = nullptr;   // [nclass] beam row of each class
    const int *d_class_of_row = nullptr;  // [n_utile]
    // Precision path: Gaussians and shapelets on the tile-pair, row-by-column and
    // channel-per-lane kernels run in FP32 with FP64 phases (d_sbf32 = the basis table in
    // FP32). Points, the per-tile baseline kernel and the per-segment fallback stay FP64.
    bool fp32 = false;
    const float *d_sbf32 = nullptr;
};

// 0: pick the fastest kernel for the case; 1: always the baseline-per-lane kernels.
extern int g_direct_variant;

// Device work buffers of run_type beyond the coherency table. The owner (a modeller handle,
// or the calling thread for the kernel-level drop-ins) keeps them between calls and frees
// them with itself.
struct Scratch {
    DevBuf accum;    // split component loop of a baseline subset: [nchan][nsub][8] FP64
    DevBuf cls;      // shapelets: J_r S J_0^H per beam-row class
    DevBuf prange, packed;  // shapelets: packed coefficients
    DevBuf coh0;     // J_0 S J_0^H of beam-row class 0
    DevBuf cseg;     // one segment's shared coherency (fallback of the row-by-column kernel)
    DevBuf ccls;     // row-by-column kernel: J_r S J_0^H per beam-row class
    // FP32 copies for the FP32-arithmetic kernels (ModelCtx::fp32)
    DevBuf coh32, coh0_32, ccls32, cls32;
};

// Largest per-class coherency table (bytes) run_type builds; above it, or when the
// allocation fails, it takes the per-segment path. Negative (default): the device's free
// memory less 1 GiB, read when the table is needed. Process-wide.
extern long long g_class_table_limit;

// Coherency table + direct kernel for one type; ev0/ev1 (nullable) bracket the direct kernel.
hipError_t run_type(Kind kind, const TypeInputs &in, const ModelCtx &ctx, DevBuf &coh, Scratch &scratch,
                    JonesF32 *vis32, JonesF64 *vis64, hipStream_t stream, hipEvent_t ev0, hipEvent_t ev1);

double4 host_gaussian_shape(GaussianParams g);
double4 host_shapelet_shape(GaussianParams g);

}  // namespace sm
