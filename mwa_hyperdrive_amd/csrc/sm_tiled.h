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
    const double *coh;        // [ncomp][nchan_pad][8]
    JonesF32 *vis32;
    JonesF64 *vis64;
    // Optional tile subset: kernel tile i is array tile tile_map[i] (ascending), and output
    // baselines are numbered over ntiles_bl array tiles.
    const int *tile_map = nullptr;
    int ntiles_bl = 0;
};

struct ShapeletChanArgs {
    int nbl, nchan, nchan_pad, ncomp, ngroups;
    const double *freqs;
    const UVW *uvw;           // [nbl]
    const double4 *lmn;
    const double4 *shape;
    const double *coh;
    const double2 *bl_uv;     // [ncomp][nbl]
    const int4 *prange;
    const PackedCoeff *packed;
    const double *sbf;
    BeamDesc beam;
    JonesF32 *vis32;
    JonesF64 *vis64;
};

std::vector<int4> tile_pairs(int ntiles);
hipError_t launch_direct_tiled(Kind kind, const TiledArgs &a, hipStream_t stream);
hipError_t launch_shapelet_chan(bool per_tile, const ShapeletChanArgs &a, hipStream_t stream);
hipError_t launch_tile_uvw_from_baselines(int ntiles, const UVW *d_uvw, double *d_out, hipStream_t stream);

}  // namespace sm
