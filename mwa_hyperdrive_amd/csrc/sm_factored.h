// Split-precision MFMA point path: arguments and launchers (factored.hip).
#pragma once

#include <vector>

#include "sm_internal.h"

namespace sm {

struct FactoredArgs {
    int ntiles, nchan, nchan_pad, ncomp, nblocks, seg_steps;
    int tpad;                 // ntiles rounded up to the 64-tile block size
    int c_begin, c_count;     // the chunk of components in flight
    int cpad;                 // c_count rounded up to the 16-component step
    bool even;                // channels evenly spaced within each 8-channel group
    const int2 *blocks;       // [nblocks] (p-block, q-block), p-block <= q-block, 64 tiles each
    const double *freqs;      // [nchan_pad]
    const double *tile_uvw;   // [ntiles][3]
    const double4 *lmn;       // [ncomp]
    const double *coh;        // [ncomp][nchan_pad][8], J S J^H
    const double *scale;      // device scalar, set by launch_factored
    float2 *kbuf;             // [nchan][tpad][cpad] phasors of the chunk
    // Kernel 2 only: the same phasors split into hi and lo halves of the MFMA input type,
    // (re, im) packed in one 32-bit word, [nchan][tpad][cpad] each. Null otherwise.
    unsigned *khi = nullptr, *klo = nullptr;
    int planes_dtype = -1;    // SM_MFMA_FP16 / SM_MFMA_BF16 for the planes, -1 without them
    float4 *cohf;             // [nchan][nclass][cpad][2] scaled FP32 coherencies of the chunk
    double *out;              // [nchan][nblocks][4][64][128] FP64 block results
    // Per-tile beams: the A-side coherency of row tile p is J_r S J_0^H with r the beam row of
    // p's class, J_0 that of class 0 (every B-side tile belongs to class 0), and `coh` holds
    // S. With nclass == 1, `coh` holds the coherency itself and the fields below are unused.
    int nclass = 1;
    const int *rclass = nullptr;     // [tpad] class of each factored tile
    const int *class_row = nullptr;  // [nclass] beam row of each class
    BeamDesc beam{};
};

// 1 (default): wave-specialised MFMA kernel; 0: every wave generates operands and runs MFMAs;
// 2: wave-specialised with the K operand read by the MFMA waves from pre-split planes.
extern int g_factored_kernel;

std::vector<int2> factored_blocks(int ntiles, std::vector<int> &block_of);
size_t factored_out_doubles(int nchan, int nblocks);
// Components per chunk (a multiple of 16 * seg_steps unless one chunk holds them all), so
// that the phasor buffer stays within a fixed cap, and that buffer's size.
int factored_chunk(int nchan, int ntiles, int ncomp, int seg_steps);
size_t factored_kbuf_bytes(int nchan, int ntiles, int chunk);
size_t factored_cohf_bytes(int nchan, int chunk, int nclass);
hipError_t launch_gather_rows3(int n, const int *d_idx, const double *d_src, double *d_dst, hipStream_t stream);
// map[nmap] = (output baseline, p, q, h) in the factored tile numbering (a.ntiles tiles):
// the baseline takes block result (p, q), or with h = 1 its conjugate transpose; nbl = the
// visibility buffer's baseline count. coh_n = the element count of `coh` (nclass == 1).
hipError_t launch_factored(int dtype, FactoredArgs a, int chunk, const double *coh_for_scale, size_t coh_n,
                           unsigned long long *d_max_bits, double *d_scale, const int4 *map, int nmap, int nbl,
                           const int *block_of, JonesF32 *vis32, JonesF64 *vis64, hipStream_t stream,
                           hipEvent_t ev0, hipEvent_t ev1);

}  // namespace sm
