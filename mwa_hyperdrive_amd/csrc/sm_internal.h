// Internal interfaces of the MI355X sky-model predictor (not part of the C ABI).
//
// Device data layout (per component type, components concatenated in the reference's
// bucket order power law -> curved power law -> list, src/model/gpu.rs:889-910):
//   lmn4  [n]                    double4 (2pi l, 2pi m, 2pi (n-1), 0)
//   shape [n]                    double4: Gaussians (A, B, C, 0) of the envelope exponent
//                                A u^2 + B u v + C v^2 (u, v in metres; includes the
//                                -(pi/2)^2/ln2 factor and 1/c^2); shapelets
//                                (sin pa, cos pa, maj*K/c, -min*K/c) with K = sqrt(pi^2/(2 ln2))/0.01
//   coh   [n][nchan_pad][8]      f64: J1 S J2^H when every tile shares one beam (or there is
//                                no beam), otherwise the instrumental flux S alone
//   bl_uv [n][nbl]               double2 shapelet (u, v) metres, component as phase centre
//   vis   [nchan][nbl]           JonesF32 (+=) and optionally JonesF64 (+=)
// Channels are processed in chunks of at most SM_NCH (kernels pick 4 or 8); nchan_pad rounds
// nchan up to a whole chunk of SM_NCH and adds one more, so every chunk's reads stay inside.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/skymodel.h"

#define SM_NCH 8
#define SM_VEL_C 299792458.0
// Size of the fixed work array of sm_model_points_noalloc (bytes, per GPU).
#define SM_NOALLOC_SCRATCH_BYTES (64u << 20)

namespace sm {

enum Kind { KIND_POINT = 0, KIND_GAUSSIAN = 1, KIND_SHAPELET = 2 };

// Flux description of one component type as the reference's GPU structs hold it.
struct FluxBuckets {
    int n_pl, n_cpl, n_list;
    const JonesF64 *pl_fds;
    const double *pl_sis;
    const JonesF64 *cpl_fds;
    const double *cpl_sis;
    const double *cpl_qs;
    const JonesF64 *list_fds;  // [nchan][n_list]
};

// How the beam enters: none (identity), one shared beam row, or per-tile rows.
struct BeamDesc {
    const JonesF64 *jones;  // [n_utile][n_ufreq][n_dir] or nullptr (identity)
    const int *freq_map;    // [nchan] -> unique freq
    int n_ufreq;
    int n_dir;              // directions per (tile, freq) row of `jones`
    int dir_offset;         // first direction of this component type within a row
    int uniform_row;        // unique-tile row used by every tile (uniform mode), else -1
    const int2 *bl_rows;    // [nbl] unique-tile rows of (tile1, tile2) (per-tile mode)
    const int *tile_rows;   // [ntiles] unique-tile row per unflagged tile (autos)
    // Per-tile mode, optional: J_r S J_0^H for every beam-row class r, where class 0 is row
    // row0 ([nclass][ncomp][nchan_pad][8]). In the channel-per-lane shapelet kernel, a
    // baseline with one tile in row row0 reads its coherency here (conjugate-transposed when
    // row0 is its first tile's row) instead of forming J_p S J_q^H per visibility.
    const double *coh_cls = nullptr;
    const int *class_of_row = nullptr;  // [n_utile]
    int row0 = -1;
};

// Channel chunks of the direct kernels. With tables: runs of at most 4 (c4), 6 (c6) or 8 (c8)
// evenly spaced channels, (first channel, count), so that a flagged channel only shortens
// the runs around it; every chunk is evenly spaced. Null tables: fixed chunks of 4, 6 or 8
// consecutive channels, evenly spaced or not. A chunk reads up to 8 channels from its first
// one, inside the padded channel count (pad_channels adds 8).
struct ChunkTabs {
    const int2 *c4 = nullptr;
    int n4 = 0;
    const int2 *c6 = nullptr;
    int n6 = 0;
    const int2 *c8 = nullptr;
    int n8 = 0;
};

// Runs of at most `width` channels f0, f0 + 1, ... with freqs[f0 + k] == freqs[f0] + k (freqs[f0 + 1]
// - freqs[f0]) exactly, covering all nchan channels in order.
std::vector<int2> even_chunks(const double *freqs, int nchan, int width);

// Number of chunks of width nch (4, 6 or 8).
__host__ __device__ inline int chunk_count(const ChunkTabs &t, int nch, int nchan) {
    const int2 *tab = nch == 4 ? t.c4 : (nch == 6 ? t.c6 : t.c8);
    if (tab != nullptr) return nch == 4 ? t.n4 : (nch == 6 ? t.n6 : t.n8);
    return (nchan + nch - 1) / nch;
}

// First channel and channel count of chunk `chunk` of width NCH.
template <int NCH>
__device__ __forceinline__ void chunk_range(const ChunkTabs &t, int chunk, int nchan, int &f0, int &cnt) {
    static_assert(NCH == 4 || NCH == 6 || NCH == 8, "chunk tables exist for widths 4, 6 and 8");
    const int2 *tab = NCH == 4 ? t.c4 : (NCH == 6 ? t.c6 : t.c8);
    if (tab != nullptr) {
        const int2 c = tab[chunk];
        f0 = c.x;
        cnt = c.y;
    } else {
        f0 = chunk * NCH;
        cnt = nchan - f0 < NCH ? nchan - f0 : NCH;
    }
}

// Index (i, j), i < j, of entry b of the upper triangle of an n x n grid in row-major order
// (the baseline order of src/math/mod.rs:96-109): row i holds n - 1 - i entries starting at
// i n - i (i + 1) / 2. The float estimate is corrected by integer tests, so the result is
// exact for every n.
__device__ __forceinline__ void triangle_pair(int b, int n, int &i, int &j) {
    const double nn = (double)n - 0.5;
    int r = (int)floor(nn - sqrt(nn * nn - 2.0 * (double)b));
    if (r < 0) r = 0;
    if (r > n - 2) r = n - 2;
    auto start = [n](int row) { return row * n - row * (row + 1) / 2; };
    while (r > 0 && start(r) > b) --r;
    while (r < n - 2 && start(r + 1) <= b) ++r;
    i = r;
    j = b - start(r) + r + 1;
}

// Shapelet (u, v) of baseline (tile 1 - tile 2) with the component as phase centre
// (ShapeletComponentParams::get_shapelet_uvws, src/srclist/types/components/mod.rs:299-315):
// each tile's (u, v) at the component's (HA, Dec), then their difference. hd = (sin HA,
// cos HA, sin Dec, cos Dec). No contraction, so every caller rounds alike.
__device__ __forceinline__ double2 shapelet_uv(double4 hd, double x1, double y1, double z1, double x2, double y2,
                                               double z2) {
#pragma clang fp contract(off)
    const double u1 = hd.x * x1 + hd.y * y1;
    const double v1 = -hd.z * hd.y * x1 + hd.z * hd.x * y1 + hd.w * z1;
    const double u2 = hd.x * x2 + hd.y * y2;
    const double v2 = -hd.z * hd.y * x2 + hd.z * hd.x * y2 + hd.w * z2;
    return make_double2(u1 - u2, v1 - v2);
}

// A shapelet coefficient as the kernels read it: the value with the sign of
// i^((n1 + n2) mod 4) folded in, and the basis-table rows n1, n2. Per component the terms
// that land in the real part of the envelope come first, then those that land in the
// imaginary part, each group in source order (launch_pack_coeffs).
struct PackedCoeff {
    double value;
    int n1, n2;
};

struct DirectArgs {
    int nbl, nchan, nchan_pad, ncomp;
    const double *freqs;      // [nchan_pad]
    const UVW *uvw;           // [nbl]
    const double4 *lmn;       // [ncomp]
    const double4 *shape;     // [ncomp]
    const double *coh;        // [ncomp][nchan_pad][8]
    const double2 *bl_uv;     // [ncomp][nbl]
    const int4 *prange;       // [ncomp] (offset, real-part terms, imaginary-part terms, rows used)
    const PackedCoeff *packed;
    const double *sbf;        // [SM_SBF_N][SM_SBF_L]
    BeamDesc beam;            // per-tile mode only
    JonesF32 *vis32;          // [nchan][nbl]
    JonesF64 *vis64;          // nullable
    const int *bl_subset = nullptr;  // nullable: model only these baselines (nsub of them)
    int nsub = 0;
    // Subset mode only: a few baselines give too few lanes to fill the GPU, so the component
    // loop is split into csplit ranges whose partial sums meet in accum ([nchan][nsub][8],
    // FP64, set up by launch_direct) before one pass adds them to the visibilities.
    int csplit = 1;
    double *accum = nullptr;
    ChunkTabs chunks;  // channel chunks (evenly spaced runs), or fixed chunks when null
};

// Launchers (direct.hip). All enqueue on `stream` and return hipSuccess or an error.
hipError_t launch_pack_coeffs(int ncomp, const int2 *d_range, const ShapeletCoeff *d_coeffs, int4 *d_prange,
                              PackedCoeff *d_packed, hipStream_t stream);
// coh_cls[r][c][f] = J_{class_row[r]} S J_{class_row[0]}^H from S = coh ([ncomp][nchan_pad][8]).
hipError_t launch_class_coherency(int nclass, int ncomp, int nchan, int nchan_pad, const double *d_s,
                                  const BeamDesc &beam, const int *d_class_row, double *d_coh_cls,
                                  hipStream_t stream);
// out[c][f] = J_{row_a} S J_{row_b}^H from S = d_s ([ncomp][nchan_pad][8]).
hipError_t launch_pair_coherency(int ncomp, int nchan, int nchan_pad, const double *d_s, const BeamDesc &beam,
                                 int row_a, int row_b, double *d_out, hipStream_t stream);
// Component ranges for a baseline subset of nsub baselines (1 when they fill the GPU).
int direct_subset_split(int nsub, int nchan, int ncomp);
hipError_t launch_coherency(int ncomp, int nchan, int nchan_pad, const double *d_freqs, FluxBuckets fb,
                            BeamDesc beam, bool apply_beam, double *d_coh, hipStream_t stream, bool hermitian = false);
hipError_t launch_direct(Kind kind, bool per_tile, bool even, const DirectArgs &a, hipStream_t stream);
bool chunks_even(const double *freqs_padded, int nchan);
hipError_t launch_autos(int ntiles, int nchan, int ncomp, const double *d_coh, int nchan_pad, BeamDesc beam,
                        bool coh_has_beam, JonesF32 *vis32, JonesF64 *vis64, double *accum64, bool demote_now,
                        hipStream_t stream);
hipError_t launch_autos_demote(int n, const double *accum64, JonesF32 *vis32, JonesF64 *vis64, hipStream_t stream);
hipError_t launch_shapelet_bl_uv(int ncomp, int ntiles, int nbl, const double *d_comp_hadec, const double *d_xyz,
                                 const int2 *d_bl_tiles, double2 *d_bl_uv, hipStream_t stream);
hipError_t launch_repack_shapelet_uv(int nbl, int n_pl, int n_cpl, int n_list, const ShapeletUV *pl,
                                     const ShapeletUV *cpl, const ShapeletUV *list, double2 *d_bl_uv,
                                     hipStream_t stream);
hipError_t launch_analytic_beam(int n_tiles, const unsigned *d_delays, const double *d_gains, int n_freqs,
                                const double *d_freqs, int n_dirs, const double *d_az, const double *d_za,
                                JonesF64 *d_out, hipStream_t stream);

// Host helpers (abi.hip).
int pad_channels(int nchan);
// Pads a frequency list to whole channel chunks by continuing its last step.
void pad_freqs(const double *freqs, int nchan, double *out_padded);

// A grow-only device buffer. A borrowed buffer views memory it does not own (set up with
// borrow()): it never allocates or frees, and reserve() beyond its capacity fails with
// hipErrorOutOfMemory.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    bool borrowed = false;
    hipError_t reserve(size_t n);
    void release();
    void borrow(void *p, size_t n) {
        release();
        ptr = p;
        bytes = n;
        borrowed = true;
    }
    ~DevBuf() { release(); }
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

}  // namespace sm

#define SM_TRY(call)                                   \
    do {                                               \
        hipError_t sm_err_ = (call);                   \
        if (sm_err_ != hipSuccess) return hipGetErrorString(sm_err_); \
    } while (0)
