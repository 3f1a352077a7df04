// Arguments of the tiled point/Gaussian kernel and the channel-per-lane shapelet kernel.
#pragma once

#include <vector>

#include "sm_internal.h"
#include "sm_jones.h"

namespace sm {

struct TiledArgs {
    int ntiles, nbl, nchan, nchan_pad, ncomp, npairs, nchunks;
    const int4 *pairs;        // [npairs] tile-group pairs (see tile_pairs)
    const double *freqs;      // [nchan_pad]
    const double *tile_uvw;   // [ntiles][3] metres
    const double4 *lmn;
    const double4 *shape;
    // The Hermitian form of the coherency table, [ncomp][nchan_pad][4] = (Re XX, Re XY, Im XY,
    // Re YY) of J S J^H (launch_coherency with `hermitian`). Valid because the same J stands on
    // both sides (one shared beam, or class 0's beam for pairs of two class-0 tiles) and S is
    // Hermitian, so XX and YY are real and YX = conj(XY).
    const double *coh;
    // FP32 arithmetic (phases and their range reduction still FP64): the same table in FP32.
    // Null selects the FP64 kernels.
    const float *coh32 = nullptr;
    JonesF32 *vis32;
    JonesF64 *vis64;
    // Optional tile subset: kernel tile i is array tile tile_map[i] (ascending), and output
    // baselines are numbered over ntiles_bl array tiles.
    const int *tile_map = nullptr;
    int ntiles_bl = 0;
    ChunkTabs chunks;  // channel chunks; the fixed chunks must then be evenly spaced
};

struct ShapeletChanArgs {
    // ngroups: groups of 64 channels of the per-tile kernels (the shared-beam kernel takes
    // groups of 128, two channels per lane, and forms its own count).
    int nbl, nchan, nchan_pad, ncomp, ngroups;
    const double *freqs;
    const UVW *uvw;           // [nbl]
    const double4 *lmn;
    const double4 *shape;
    // Shared beam: the Hermitian form of J S J^H, [ncomp][nchan_pad][4] (Re XX, Re XY, Im XY,
    // Re YY). Per-tile beams: S, [ncomp][nchan_pad][8].
    const double *coh;
    const double2 *bl_uv;     // [ncomp][nbl]
    // With hadec set, the UVs are formed in the kernel from hadec [ncomp][4] (sin HA, cos HA,
    // sin Dec, cos Dec), xyz [ntiles][3] and bl_tiles [nbl], as k_shapelet_bl_uv forms bl_uv.
    const double *hadec = nullptr;
    const double *xyz = nullptr;
    const int2 *bl_tiles = nullptr;
    const int4 *prange;
    const PackedCoeff *packed;
    const double *sbf;
    BeamDesc beam;
    JonesF32 *vis32;
    JonesF64 *vis64;
    // FP32 arithmetic with FP64 phases and table positions: coherency, per-class coherency
    // and basis table in FP32. Null coh32 selects the FP64 kernels.
    const float *coh32 = nullptr;
    const float *coh_cls32 = nullptr;  // [nclass][ncomp][nchan_pad][8] when beam.coh_cls is set
    const float *sbf32 = nullptr;      // [SM_SBF_N][SM_SBF_L]
};

// Pairs of a tile outside beam class 0 (rows) with a class-0 tile (columns): one wave per
// row tile and 64 column tiles, so the row's class, and with it the coherency
// J_r S J_0^H, is wave-uniform.
struct RectArgs {
    int n_row, n_col, nbl, ntiles_bl, nchan, nchan_pad, ncomp;
    const int *row_tiles;     // [n_row] array tiles outside class 0
    const int *row_cls;       // [n_row] their classes (1 .. nclass - 1)
    const int *col_tiles;     // [n_col] class-0 array tiles
    const double *tile_uvw;   // [ntiles_bl][3] metres
    const double *freqs;      // [nchan_pad]
    const double4 *lmn;
    const double4 *shape;
    const double *coh_cls;    // [nclass][ncomp][nchan_pad][8], J_r S J_0^H
    const float *coh_cls32 = nullptr;  // the same in FP32 (FP32 arithmetic); null: FP64 kernels
    JonesF32 *vis32;
    JonesF64 *vis64;
    ChunkTabs chunks;  // channel chunks; the fixed chunks must then be evenly spaced
};

// Pairs of two tiles with their own beam rows (tiles outside beam class 0): a block owns a
// 16 x 16 block of tile pairs of the listed tiles and a chunk of channels, and per component
// and channel stages J_p S in LDS for its 16 row tiles and J_q for its 16 column tiles, so
// that each lane forms (J_p S) J_q^H from LDS instead of reading both tiles' Jones from
// memory per visibility.
struct PairJonesArgs {
    int ntiles_list, ntiles_bl, nbl, nchan, nchan_pad, ncomp;
    const int *tiles;         // [ntiles_list] array tiles, ascending
    const double *tile_uvw;   // [ntiles_bl][3]
    const double *freqs;      // [nchan_pad]
    const double4 *lmn;
    const double4 *shape;
    const double *coh;        // [ncomp][nchan_pad][8] instrumental S
    BeamDesc beam;            // per-tile rows: tile_rows, freq_map, jones
    bool fp32 = false;        // FP32 arithmetic with FP64 phases
    JonesF32 *vis32;
    JonesF64 *vis64;
    ChunkTabs chunks;
    // Set by the launcher: channel chunks, and the component split. With csplit > 1 each
    // block sums one slice of the components into accum [nchan][pairs][8] (FP64), and a
    // finishing kernel adds those sums to the visibilities.
    int nchunks = 0, csplit = 1;
    double *accum = nullptr;
};

std::vector<int4> tile_pairs(int ntiles);
// accum (nullable) backs the component split, which a grid of fewer than about 2048 blocks
// takes (few tiles outside beam class 0); without it every block runs all components.
hipError_t launch_pair_jones(Kind kind, const PairJonesArgs &a, DevBuf *accum, hipStream_t stream);
hipError_t launch_direct_rect(Kind kind, const RectArgs &a, hipStream_t stream);
hipError_t launch_direct_tiled(Kind kind, const TiledArgs &a, hipStream_t stream);
hipError_t launch_shapelet_chan(bool per_tile, const ShapeletChanArgs &a, hipStream_t stream);
hipError_t launch_tile_uvw_from_baselines(int ntiles, const UVW *d_uvw, double *d_out, hipStream_t stream);
// FP64 -> FP32 copy of n values (coherency tables, basis table).
hipError_t launch_to_f32(size_t n, const double *d_in, float *d_out, hipStream_t stream);

}  // namespace sm
