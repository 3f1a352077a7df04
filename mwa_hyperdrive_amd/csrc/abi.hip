// Kernel-level drop-ins with the reference's signatures (src/gpu/model.h:33-56) and the
// shared per-component-type driver used by both these and the SkyModeller handle.
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "sm_abi_common.h"
#include "sm_tiled.h"

namespace sm {

hipError_t DevBuf::reserve(size_t n) {
    if (n <= bytes && ptr != nullptr) return hipSuccess;
    if (borrowed) return hipErrorOutOfMemory;
    release();
    size_t want = n < 256 ? 256 : n;
    hipError_t e = hipMalloc(&ptr, want);
    if (e != hipSuccess) {
        ptr = nullptr;
        bytes = 0;
        return e;
    }
    bytes = want;
    return hipSuccess;
}

void DevBuf::release() {
    if (ptr != nullptr && !borrowed) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
    borrowed = false;
}

int pad_channels(int nchan) { return ((nchan + SM_NCH - 1) / SM_NCH) * SM_NCH + SM_NCH; }

void pad_freqs(const double *freqs, int nchan, double *out) {
    const int npad = pad_channels(nchan);
    for (int f = 0; f < nchan; ++f) out[f] = freqs[f];
    const double step = nchan >= 2 ? freqs[nchan - 1] - freqs[nchan - 2] : 1.0;
    for (int f = nchan; f < npad; ++f) out[f] = freqs[nchan - 1] + (double)(f - nchan + 1) * step;
}

// ------------------------------------------------------------------ parameter gathers
__global__ void k_gather_lmn(int n_pl, const LmnRime *pl, int n_cpl, const LmnRime *cpl, int n_list,
                             const LmnRime *list, double4 *out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = n_pl + n_cpl + n_list;
    if (c >= n) return;
    const LmnRime l = c < n_pl ? pl[c] : (c < n_pl + n_cpl ? cpl[c - n_pl] : list[c - n_pl - n_cpl]);
    out[c] = make_double4(l.l, l.m, l.n, 0.0);
}

// Gaussian envelope exponent coefficients (src/model/cpu.rs:456-465 expanded as a quadratic
// form in the baseline's (u, v) metres, with the 1/c^2 of the wavelength scaling folded in).
__device__ __host__ inline double4 gaussian_shape(GaussianParams g) {
    const double K = -(1.5707963267948966 * 1.5707963267948966) / 0.6931471805599453;
    const double s = sin(g.pa), c = cos(g.pa);
    const double maj2 = g.maj * g.maj, min2 = g.min * g.min;
    const double ic2 = 1.0 / (SM_VEL_C * SM_VEL_C);
    return make_double4(K * (maj2 * s * s + min2 * c * c) * ic2, K * (2.0 * s * c * (maj2 - min2)) * ic2,
                        K * (maj2 * c * c + min2 * s * s) * ic2, 0.0);
}

// Shapelet position scales (src/model/cpu.rs:638-644): (sin pa, cos pa, maj K / c, -min K / c).
__device__ __host__ inline double4 shapelet_shape(GaussianParams g) {
    const double K = 2.6682231283184983 / SM_SBF_DX;
    return make_double4(sin(g.pa), cos(g.pa), g.maj * K / SM_VEL_C, -g.min * K / SM_VEL_C);
}

__global__ void k_gather_shape(int kind, int n_pl, const GaussianParams *pl, int n_cpl, const GaussianParams *cpl,
                               int n_list, const GaussianParams *list, double4 *out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = n_pl + n_cpl + n_list;
    if (c >= n) return;
    const GaussianParams g = c < n_pl ? pl[c] : (c < n_pl + n_cpl ? cpl[c - n_pl] : list[c - n_pl - n_cpl]);
    out[c] = kind == KIND_GAUSSIAN ? gaussian_shape(g) : shapelet_shape(g);
}

double4 host_gaussian_shape(GaussianParams g) { return gaussian_shape(g); }
double4 host_shapelet_shape(GaussianParams g) { return shapelet_shape(g); }

// ------------------------------------------------------------------ shared driver
// The baseline kernel on the subset a.bl_subset / a.nsub, splitting the component loop when
// the subset fills too few lanes (FP64 partial sums in accum).
static hipError_t launch_subset(Kind kind, bool per_tile, bool even, DirectArgs a, DevBuf &accum,
                                hipStream_t stream) {
    a.csplit = direct_subset_split(a.nsub, a.nchan, a.ncomp);
    if (a.csplit > 1) {
        hipError_t e = accum.reserve(sizeof(double) * 8 * (size_t)a.nchan * a.nsub);
        if (e != hipSuccess) return e;
        a.accum = accum.as<double>();
    }
    return launch_direct(kind, per_tile, even, a, stream);
}

long long g_class_table_limit = -1;

// An FP32 copy of n FP64 values in buf.
static hipError_t to_f32(DevBuf &buf, size_t n, const double *src, hipStream_t stream) {
    hipError_t e = buf.reserve(sizeof(float) * (n > 0 ? n : 1));
    if (e != hipSuccess) return e;
    return launch_to_f32(n, src, buf.as<float>(), stream);
}

// Reserves a per-class coherency table of `bytes` in buf if it is within the limit and the
// allocation succeeds; false sends the caller to its per-segment path. A failed allocation
// is cleared from the runtime's last-error state, so it does not surface later.
static bool reserve_class_table(DevBuf &buf, size_t bytes) {
    if (g_class_table_limit >= 0 && bytes > (size_t)g_class_table_limit) return false;
    if (buf.ptr != nullptr && buf.bytes >= bytes) return true;
    size_t limit;
    if (g_class_table_limit >= 0) {
        limit = (size_t)g_class_table_limit;
    } else {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        // What buf already holds is released before the larger allocation.
        free_b += buf.bytes;
        limit = free_b > (size_t(1) << 30) ? free_b - (size_t(1) << 30) : 0;
    }
    if (bytes > limit) return false;
    if (buf.reserve(bytes) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return true;
}

hipError_t run_type(Kind kind, const TypeInputs &in, const ModelCtx &ctx, DevBuf &coh, Scratch &scratch,
                    JonesF32 *vis32, JonesF64 *vis64, hipStream_t stream, hipEvent_t ev0, hipEvent_t ev1) {
    const int n = in.fb.n_pl + in.fb.n_cpl + in.fb.n_list;
    if (n == 0 || ctx.nbl == 0 || ctx.nchan == 0) return hipSuccess;
    hipError_t e = coh.reserve((size_t)n * ctx.nchan_pad * 8 * sizeof(double));
    if (e != hipSuccess) return e;
    const bool per_tile = ctx.beam.jones != nullptr && ctx.beam.uniform_row < 0;
    // Shared beam, evenly spaced channels: the tile-pair kernel, which reads the table in its
    // Hermitian form (4 values per entry).
    const bool shared_tiled = ctx.d_bl_subset == nullptr && g_direct_variant == 0 && kind != KIND_SHAPELET &&
                              !per_tile && ctx.even && ctx.d_tile_uvw != nullptr && ctx.d_pairs != nullptr &&
                              ctx.ntiles >= 2;
    // Shared-beam shapelets on the two-channels-per-lane kernel read the Hermitian form too.
    const bool shared_shapelet =
        ctx.d_bl_subset == nullptr && g_direct_variant != 1 && kind == KIND_SHAPELET && !per_tile;
    const bool hermitian = shared_tiled || shared_shapelet;
    e = launch_coherency(n, ctx.nchan, ctx.nchan_pad, ctx.d_freqs_pad, in.fb, ctx.beam, !per_tile, coh.as<double>(),
                         stream, hermitian);
    if (e != hipSuccess) return e;
    // Per-tile shapelets with beam-row classes: J_r S J_0^H per class for the channel-per-lane
    // kernel, when the table fits (otherwise the kernel forms J_p S J_q^H per visibility).
    // (Measured on the 256-tile case: that kernel runs 28% faster. The baseline-per-lane
    // kernels ran 2.2x slower with the same lookup, whether compiled into them or as a
    // separate instantiation, so they keep forming J_p S J_q^H.)
    BeamDesc beam = ctx.beam;
    const size_t cls_bytes = (size_t)ctx.nclass * n * ctx.nchan_pad * 8 * sizeof(double);
    if (kind == KIND_SHAPELET && g_direct_variant != 1 && ctx.d_bl_subset == nullptr && per_tile &&
        ctx.nclass > 0 && ctx.d_class_row != nullptr && ctx.class_row0 >= 0 &&
        reserve_class_table(scratch.cls, cls_bytes)) {
        e = launch_class_coherency(ctx.nclass, n, ctx.nchan, ctx.nchan_pad, coh.as<double>(), ctx.beam, ctx.d_class_row,
                                   scratch.cls.as<double>(), stream);
        if (e != hipSuccess) return e;
        beam.coh_cls = scratch.cls.as<double>();
        beam.class_of_row = ctx.d_class_of_row;
        beam.row0 = ctx.class_row0;
    }
    const size_t coh_n = (size_t)n * ctx.nchan_pad * 8;
    const bool f32 = ctx.fp32 && kind != KIND_POINT && ctx.d_bl_subset == nullptr && g_direct_variant == 0;
    // Shapelet coefficients in the kernels' packed form.
    DevBuf &prange = scratch.prange, &packed = scratch.packed;
    const int4 *d_prange = nullptr;
    const PackedCoeff *d_packed = nullptr;
    if (kind == KIND_SHAPELET) {
        e = packed.reserve(sizeof(PackedCoeff) * (in.coeff_count > 0 ? in.coeff_count : 1));
        if (e == hipSuccess) e = prange.reserve(sizeof(int4) * n);
        if (e == hipSuccess)
            e = launch_pack_coeffs(n, in.coeff_range, in.coeffs, prange.as<int4>(), packed.as<PackedCoeff>(), stream);
        if (e != hipSuccess) return e;
        d_prange = prange.as<int4>();
        d_packed = packed.as<PackedCoeff>();
    }
    DirectArgs a;
    a.nbl = ctx.nbl;
    a.nchan = ctx.nchan;
    a.nchan_pad = ctx.nchan_pad;
    a.ncomp = n;
    a.freqs = ctx.d_freqs_pad;
    a.uvw = ctx.d_uvw;
    a.lmn = in.lmn;
    a.shape = in.shape;
    a.coh = coh.as<double>();
    a.bl_uv = in.bl_uv;
    a.prange = d_prange;
    a.packed = d_packed;
    a.sbf = ctx.d_sbf;
    a.beam = beam;
    a.vis32 = vis32;
    a.vis64 = vis64;
    a.bl_subset = ctx.d_bl_subset;
    a.nsub = ctx.nsub;
    a.chunks = ctx.chunks;
    const bool g0_tiled = per_tile && ctx.d_bl_subset == nullptr && g_direct_variant == 0 && kind != KIND_SHAPELET &&
                          ctx.even && ctx.g0_n >= 2 && ctx.d_g0_uvw != nullptr && ctx.class_row0 >= 0;
    if (g0_tiled) {
        // Pairs of two class-0 tiles: J_0 S J_0^H, shared by all of them, on the tile-pair
        // kernel over the class-0 tiles; the other baselines: per-tile kernel on a subset.
        // J_0 S J_0^H in the Hermitian form the tile-pair kernel reads (4 values per entry).
        DevBuf &coh0 = scratch.coh0;
        e = coh0.reserve((size_t)n * ctx.nchan_pad * 4 * sizeof(double));
        if (e != hipSuccess) return e;
        BeamDesc b0 = ctx.beam;
        b0.uniform_row = ctx.class_row0;
        e = launch_coherency(n, ctx.nchan, ctx.nchan_pad, ctx.d_freqs_pad, in.fb, b0, true, coh0.as<double>(), stream,
                             true);
        if (e == hipSuccess && f32) e = to_f32(scratch.coh0_32, coh_n / 2, coh0.as<double>(), stream);
        if (e != hipSuccess) return e;
        TiledArgs t;
        t.ntiles = ctx.g0_n;
        t.nbl = ctx.nbl;
        t.nchan = ctx.nchan;
        t.nchan_pad = ctx.nchan_pad;
        t.ncomp = n;
        t.npairs = ctx.g0_npairs;
        t.nchunks = (ctx.nchan + SM_NCH - 1) / SM_NCH;
        t.pairs = ctx.d_g0_pairs;
        t.freqs = ctx.d_freqs_pad;
        t.tile_uvw = ctx.d_g0_uvw;
        t.lmn = in.lmn;
        t.shape = in.shape;
        t.coh = coh0.as<double>();
        if (f32) t.coh32 = scratch.coh0_32.as<float>();
        t.vis32 = vis32;
        t.vis64 = vis64;
        t.tile_map = ctx.d_g0_tiles;
        t.ntiles_bl = ctx.ntiles;
        t.chunks = ctx.chunks;
        if (ev0 != nullptr) (void)hipEventRecord(ev0, stream);
        e = launch_direct_tiled(kind, t, stream);
        // Pairs of a class-r tile with a class-0 tile. With a table of J_r S J_0^H for every
        // class (when it fits, g_class_table_limit): one launch of the row-by-column kernel.
        // Otherwise one shared coherency per segment and the shared-beam form of the
        // baseline kernel.
        DevBuf &cseg = scratch.cseg, &ccls = scratch.ccls;
        const int nseg = 2 * (ctx.nclass - 1);
        const size_t ccls_bytes = (size_t)ctx.nclass * n * ctx.nchan_pad * 8 * sizeof(double);
        const bool rect = e == hipSuccess && ctx.n_other > 0 && ctx.d_other_tiles != nullptr && ctx.nclass > 1 &&
                          reserve_class_table(ccls, ccls_bytes);
        if (rect) {
            e = launch_class_coherency(ctx.nclass, n, ctx.nchan, ctx.nchan_pad, coh.as<double>(), ctx.beam,
                                       ctx.d_class_row, ccls.as<double>(), stream);
            if (e == hipSuccess && f32) e = to_f32(scratch.ccls32, (size_t)ctx.nclass * coh_n, ccls.as<double>(), stream);
            if (e == hipSuccess) {
                RectArgs r;
                r.n_row = ctx.n_other;
                r.n_col = ctx.g0_n;
                r.nbl = ctx.nbl;
                r.ntiles_bl = ctx.ntiles;
                r.nchan = ctx.nchan;
                r.nchan_pad = ctx.nchan_pad;
                r.ncomp = n;
                r.row_tiles = ctx.d_other_tiles;
                r.row_cls = ctx.d_other_cls;
                r.col_tiles = ctx.d_g0_tiles;
                r.tile_uvw = ctx.d_tile_uvw;
                r.freqs = ctx.d_freqs_pad;
                r.lmn = in.lmn;
                r.shape = in.shape;
                r.coh_cls = ccls.as<double>();
                if (f32) r.coh_cls32 = scratch.ccls32.as<float>();
                r.vis32 = vis32;
                r.vis64 = vis64;
                r.chunks = ctx.chunks;
                e = launch_direct_rect(kind, r, stream);
            }
        }
        for (int sgi = 0; sgi < nseg && e == hipSuccess && !rect; ++sgi) {
            const int off = ctx.h_seg_off[sgi], cnt = ctx.h_seg_off[sgi + 1] - off;
            if (cnt <= 0) continue;
            const int r = 1 + sgi / 2;
            const bool fwd = (sgi & 1) == 0;
            e = cseg.reserve((size_t)n * ctx.nchan_pad * 8 * sizeof(double));
            if (e != hipSuccess) break;
            const int ra = fwd ? ctx.h_class_row[r] : ctx.class_row0, rb = fwd ? ctx.class_row0 : ctx.h_class_row[r];
            e = launch_pair_coherency(n, ctx.nchan, ctx.nchan_pad, coh.as<double>(), ctx.beam, ra, rb,
                                      cseg.as<double>(), stream);
            if (e != hipSuccess) break;
            DirectArgs u = a;
            u.coh = cseg.as<double>();
            u.bl_subset = ctx.d_notg0_bl + off;
            u.nsub = cnt;
            e = launch_subset(kind, false, ctx.even, u, scratch.accum, stream);
        }
        // Pairs of two tiles outside class 0: per-tile J_p S J_q^H, on the own-beam tile-pair
        // kernel (both tiles' Jones staged in LDS), else on the baseline kernel's subset form.
        const int off = ctx.h_seg_off[nseg], cnt = ctx.h_seg_off[nseg + 1] - off;
        const bool pairs_ok = ctx.n_other >= 2 && ctx.d_other_tiles != nullptr && ctx.beam.tile_rows != nullptr &&
                              (long)cnt == (long)ctx.n_other * (ctx.n_other - 1) / 2;
        if (e == hipSuccess && cnt > 0 && pairs_ok) {
            PairJonesArgs pj;
            pj.ntiles_list = ctx.n_other;
            pj.ntiles_bl = ctx.ntiles;
            pj.nbl = ctx.nbl;
            pj.nchan = ctx.nchan;
            pj.nchan_pad = ctx.nchan_pad;
            pj.ncomp = n;
            pj.tiles = ctx.d_other_tiles;
            pj.tile_uvw = ctx.d_tile_uvw;
            pj.freqs = ctx.d_freqs_pad;
            pj.lmn = in.lmn;
            pj.shape = in.shape;
            pj.coh = coh.as<double>();
            pj.beam = ctx.beam;
            pj.fp32 = f32;
            pj.vis32 = vis32;
            pj.vis64 = vis64;
            pj.chunks = ctx.chunks;
            e = launch_pair_jones(kind, pj, &scratch.accum, stream);
        } else if (e == hipSuccess && cnt > 0) {
            DirectArgs u = a;
            u.bl_subset = ctx.d_notg0_bl + off;
            u.nsub = cnt;
            e = launch_subset(kind, true, ctx.even, u, scratch.accum, stream);
        }
        if (ev1 != nullptr) (void)hipEventRecord(ev1, stream);
        return e != hipSuccess ? e : hipGetLastError();
    }
    if (ev0 != nullptr) (void)hipEventRecord(ev0, stream);
    if (ctx.d_bl_subset != nullptr) {
        e = launch_subset(kind, per_tile, ctx.even, a, scratch.accum, stream);
    } else if (shared_tiled) {
        TiledArgs t;
        t.ntiles = ctx.ntiles;
        t.nbl = ctx.nbl;
        t.nchan = ctx.nchan;
        t.nchan_pad = ctx.nchan_pad;
        t.ncomp = n;
        t.npairs = ctx.npairs;
        t.nchunks = (ctx.nchan + SM_NCH - 1) / SM_NCH;
        t.pairs = ctx.d_pairs;
        t.freqs = ctx.d_freqs_pad;
        t.tile_uvw = ctx.d_tile_uvw;
        t.lmn = in.lmn;
        t.shape = in.shape;
        t.coh = coh.as<double>();
        if (f32) {
            e = to_f32(scratch.coh32, coh_n / 2, coh.as<double>(), stream);
            if (e != hipSuccess) return e;
            t.coh32 = scratch.coh32.as<float>();
        }
        t.vis32 = vis32;
        t.vis64 = vis64;
        t.chunks = ctx.chunks;
        e = launch_direct_tiled(kind, t, stream);
    } else if (g_direct_variant != 1 && kind == KIND_SHAPELET) {
        ShapeletChanArgs s;
        s.nbl = ctx.nbl;
        s.nchan = ctx.nchan;
        s.nchan_pad = ctx.nchan_pad;
        s.ncomp = n;
        s.ngroups = (ctx.nchan + 63) / 64;
        s.freqs = ctx.d_freqs_pad;
        s.uvw = ctx.d_uvw;
        s.lmn = in.lmn;
        s.shape = in.shape;
        s.coh = coh.as<double>();
        s.bl_uv = in.bl_uv;
        s.hadec = in.uv_hadec;
        s.xyz = in.uv_xyz;
        s.bl_tiles = in.uv_bl_tiles;
        s.prange = d_prange;
        s.packed = d_packed;
        s.sbf = ctx.d_sbf;
        s.beam = beam;
        s.vis32 = vis32;
        s.vis64 = vis64;
        if (f32 && ctx.d_sbf32 != nullptr) {
            e = to_f32(scratch.coh32, hermitian ? coh_n / 2 : coh_n, coh.as<double>(), stream);
            if (e == hipSuccess && beam.coh_cls != nullptr)
                e = to_f32(scratch.cls32, (size_t)ctx.nclass * coh_n, beam.coh_cls, stream);
            if (e != hipSuccess) return e;
            s.coh32 = scratch.coh32.as<float>();
            if (beam.coh_cls != nullptr) s.coh_cls32 = scratch.cls32.as<float>();
            s.sbf32 = ctx.d_sbf32;
        }
        e = launch_shapelet_chan(per_tile, s, stream);
    } else {
        e = launch_direct(kind, per_tile, ctx.even, a, stream);
    }
    if (ev1 != nullptr) (void)hipEventRecord(ev1, stream);
    return e;
}

int g_direct_variant = 0;

// ------------------------------------------------------------------ ABI context
namespace {

// Work buffers of the kernel-level drop-ins. One set per calling thread, so that threads
// calling the drop-ins on their own buffers never share scratch memory.
struct AbiState {
    DevBuf freqs_pad, lmn, shape, coh, bl_rows, bl_uv, coeffs, coeff_range, sbf, accum, tile_rows, tile_uvw, pairs;
    Scratch scratch;
    std::vector<double> h_freqs_pad;
    int pairs_ntiles = -1, npairs = 0;
};
AbiState &abi_state() {
    thread_local AbiState s;
    return s;
}

// Builds the per-call context from an Addresses struct: frequencies (padded to whole
// chunks), and whether every unflagged tile uses the same beam row.
const char *abi_ctx(AbiState &st, const Addresses *a, const UVW *d_uvws, const JonesF64 *d_beam, int n_dir,
                    ModelCtx &ctx) {
    ctx = ModelCtx{};
    ctx.nbl = a->num_baselines;
    ctx.nchan = a->num_freqs;
    if (ctx.nbl <= 0 || ctx.nchan <= 0) return nullptr;
    const int ntiles = (int)llround((1.0 + std::sqrt(1.0 + 8.0 * (double)ctx.nbl)) / 2.0);
    if (ntiles * (ntiles - 1) / 2 != ctx.nbl) return "num_baselines is not a triangular number";
    ctx.nchan_pad = pad_channels(ctx.nchan);
    std::vector<double> f(ctx.nchan);
    SM_TRY(hipMemcpy(f.data(), a->d_freqs, sizeof(double) * ctx.nchan, hipMemcpyDeviceToHost));
    st.h_freqs_pad.assign(ctx.nchan_pad, 0.0);
    pad_freqs(f.data(), ctx.nchan, st.h_freqs_pad.data());
    SM_TRY(st.freqs_pad.reserve(sizeof(double) * ctx.nchan_pad));
    SM_TRY(hipMemcpy(st.freqs_pad.ptr, st.h_freqs_pad.data(), sizeof(double) * ctx.nchan_pad, hipMemcpyHostToDevice));
    ctx.d_freqs_pad = st.freqs_pad.as<double>();
    ctx.even = chunks_even(st.h_freqs_pad.data(), ctx.nchan);
    ctx.d_uvw = d_uvws;
    ctx.d_sbf = a->d_shapelet_basis_values;
    ctx.ntiles = ntiles;
    if (d_uvws != nullptr && ntiles >= 2) {
        SM_TRY(st.tile_uvw.reserve(sizeof(double) * 3 * ntiles));
        SM_TRY(launch_tile_uvw_from_baselines(ntiles, d_uvws, st.tile_uvw.as<double>(), 0));
        ctx.d_tile_uvw = st.tile_uvw.as<double>();
        if (st.pairs_ntiles != ntiles) {
            const std::vector<int4> pairs = tile_pairs(ntiles);
            SM_TRY(st.pairs.reserve(sizeof(int4) * pairs.size()));
            SM_TRY(hipMemcpy(st.pairs.ptr, pairs.data(), sizeof(int4) * pairs.size(), hipMemcpyHostToDevice));
            st.pairs_ntiles = ntiles;
            st.npairs = (int)pairs.size();
        }
        ctx.d_pairs = st.pairs.as<int4>();
        ctx.npairs = st.npairs;
    }
    BeamDesc &b = ctx.beam;
    b.jones = d_beam;
    b.freq_map = a->d_freq_map;
    b.n_ufreq = a->num_unique_beam_freqs;
    b.n_dir = n_dir;
    b.dir_offset = 0;
    b.uniform_row = 0;
    if (d_beam == nullptr) return nullptr;
    std::vector<int> unflagged_to_total(ntiles);
    SM_TRY(hipMemcpy(unflagged_to_total.data(), a->d_tile_index_to_unflagged_tile_index_map, sizeof(int) * ntiles,
                     hipMemcpyDeviceToHost));
    int max_total = 0;
    for (int t : unflagged_to_total) {
        if (t < 0) return "negative tile index in tile_index_to_unflagged_tile_index_map";
        max_total = t > max_total ? t : max_total;
    }
    std::vector<int> tile_map(max_total + 1);
    SM_TRY(hipMemcpy(tile_map.data(), a->d_tile_map, sizeof(int) * (max_total + 1), hipMemcpyDeviceToHost));
    std::vector<int> rows(ntiles);
    bool uniform = true;
    for (int t = 0; t < ntiles; ++t) {
        rows[t] = tile_map[unflagged_to_total[t]];
        if (rows[t] != rows[0]) uniform = false;
    }
    SM_TRY(st.tile_rows.reserve(sizeof(int) * ntiles));
    SM_TRY(hipMemcpy(st.tile_rows.ptr, rows.data(), sizeof(int) * ntiles, hipMemcpyHostToDevice));
    b.tile_rows = st.tile_rows.as<int>();
    if (uniform) {
        b.uniform_row = rows[0];
        return nullptr;
    }
    b.uniform_row = -1;
    std::vector<int2> bl_rows(ctx.nbl);
    int k = 0;
    for (int i = 0; i < ntiles; ++i)
        for (int j = i + 1; j < ntiles; ++j) bl_rows[k++] = make_int2(rows[i], rows[j]);
    SM_TRY(st.bl_rows.reserve(sizeof(int2) * ctx.nbl));
    SM_TRY(hipMemcpy(st.bl_rows.ptr, bl_rows.data(), sizeof(int2) * ctx.nbl, hipMemcpyHostToDevice));
    b.bl_rows = st.bl_rows.as<int2>();
    return nullptr;
}

const char *gather_lmn(int n_pl, const LmnRime *pl, int n_cpl, const LmnRime *cpl, int n_list, const LmnRime *list,
                       DevBuf &out) {
    const int n = n_pl + n_cpl + n_list;
    SM_TRY(out.reserve(sizeof(double4) * n));
    hipLaunchKernelGGL(k_gather_lmn, dim3((n + 255) / 256), dim3(256), 0, 0, n_pl, pl, n_cpl, cpl, n_list, list,
                       out.as<double4>());
    SM_TRY(hipGetLastError());
    return nullptr;
}

const char *gather_shape(Kind kind, int n_pl, const GaussianParams *pl, int n_cpl, const GaussianParams *cpl,
                         int n_list, const GaussianParams *list, DevBuf &out) {
    const int n = n_pl + n_cpl + n_list;
    SM_TRY(out.reserve(sizeof(double4) * n));
    hipLaunchKernelGGL(k_gather_shape, dim3((n + 255) / 256), dim3(256), 0, 0, (int)kind, n_pl, pl, n_cpl, cpl,
                       n_list, list, out.as<double4>());
    SM_TRY(hipGetLastError());
    return nullptr;
}

}  // namespace
}  // namespace sm

using namespace sm;

// The body of sm_model_points_f64 / sm_model_points_noalloc over the work buffers in st.
static const char *model_points_with(AbiState &st, const Points *p, const Addresses *a, const UVW *d_uvws,
                                     const JonesF64 *d_beam_jones, JonesF32 *d_vis_fb, JonesF64 *d_vis_fb_f64) {
    const int n = p->num_power_laws + p->num_curved_power_laws + p->num_lists;
    ModelCtx ctx;
    if (const char *err = abi_ctx(st, a, d_uvws, d_beam_jones, n, ctx)) return err;
    if (ctx.nbl == 0 || ctx.nchan == 0) return nullptr;
    if (const char *err = gather_lmn(p->num_power_laws, p->power_law_lmns, p->num_curved_power_laws,
                                     p->curved_power_law_lmns, p->num_lists, p->list_lmns, st.lmn))
        return err;
    TypeInputs in{};
    in.fb = FluxBuckets{p->num_power_laws, p->num_curved_power_laws, p->num_lists, p->power_law_fds,
                        p->power_law_sis, p->curved_power_law_fds, p->curved_power_law_sis, p->curved_power_law_qs,
                        p->list_fds};
    in.lmn = st.lmn.as<double4>();
    SM_TRY(run_type(KIND_POINT, in, ctx, st.coh, st.scratch, d_vis_fb, d_vis_fb_f64, 0, nullptr, nullptr));
    SM_TRY(hipStreamSynchronize(0));
    return nullptr;
}

extern "C" const char *sm_model_points_f64(const Points *p, const Addresses *a, const UVW *d_uvws,
                                           const JonesF64 *d_beam_jones, JonesF32 *d_vis_fb, JonesF64 *d_vis_fb_f64) {
    if (p == nullptr || a == nullptr) return "null argument";
    if (p->num_power_laws + p->num_curved_power_laws + p->num_lists == 0) return nullptr;
    return model_points_with(abi_state(), p, a, d_uvws, d_beam_jones, d_vis_fb, d_vis_fb_f64);
}

// ------------------------------------------------------------------ allocation-free entry
// Work memory of sm_model_points_noalloc: one array per GPU, held until the process ends.
// Host code never passes `scratch` itself (its host-side name is not a device address):
// noalloc_scratch() reads the device address with hipGetSymbolAddress once per device.
static __device__ __align__(256) unsigned char scratch[SM_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[kMaxDevices] = {};

// The current device's copy of `scratch`, looked up on the first call on that device.
const char *noalloc_scratch(void **out) {
    int dev = 0;
    SM_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return "device index beyond the supported range";
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[dev] == nullptr) {
        void *ptr = nullptr;
        SM_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (ptr == nullptr) return "hipGetSymbolAddress returned no address for the scratch array";
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return nullptr;
}

size_t round_up256(size_t b) { return ((b < 256 ? 256 : b) + 255) / 256 * 256; }

}  // namespace

extern "C" const char *sm_model_points_noalloc(const Points *p, const Addresses *a, const UVW *d_uvws,
                                               const JonesF64 *d_beam_jones, JonesF32 *d_vis_fb,
                                               JonesF64 *d_vis_fb_f64) {
    if (p == nullptr || a == nullptr) return "null argument";
    const int n = p->num_power_laws + p->num_curved_power_laws + p->num_lists;
    if (n == 0 || a->num_baselines <= 0 || a->num_freqs <= 0) return nullptr;
    if (n < 0) return "negative component count";
    const long long nbl = a->num_baselines;
    const int ntiles = (int)llround((1.0 + std::sqrt(1.0 + 8.0 * (double)nbl)) / 2.0);
    if ((long long)ntiles * (ntiles - 1) / 2 != nbl) return "num_baselines is not a triangular number";
    const size_t nchan_pad = (size_t)pad_channels(a->num_freqs);
    const size_t groups = (size_t)(ntiles + 15) / 16;
    const size_t npairs = groups * (groups - 1) / 2 + (groups + 1) / 2;
    // Every buffer the shared body reserves, in the order it reserves them.
    const size_t want[7] = {
        sizeof(double) * nchan_pad,                                  // padded frequencies
        d_uvws != nullptr ? sizeof(double) * 3 * (size_t)ntiles : 0,  // tile UVWs
        d_uvws != nullptr ? sizeof(int4) * npairs : 0,               // tile-group pairs
        d_beam_jones != nullptr ? sizeof(int) * (size_t)ntiles : 0,  // beam row per tile
        d_beam_jones != nullptr ? sizeof(int2) * (size_t)nbl : 0,    // beam rows per baseline
        sizeof(double4) * (size_t)n,                                 // LMN
        sizeof(double) * 8 * (size_t)n * nchan_pad,                  // coherency table
    };
    size_t total = 0;
    for (size_t w : want) total += w > 0 ? round_up256(w) : 0;
    if (total > (size_t)SM_NOALLOC_SCRATCH_BYTES)
        return "sm_model_points_noalloc: this call needs more work memory than the fixed scratch array "
               "holds (SM_NOALLOC_SCRATCH_BYTES; about 64 B per component and padded channel)";
    void *base = nullptr;
    if (const char *e = noalloc_scratch(&base)) return e;
    AbiState st;
    DevBuf *bufs[7] = {&st.freqs_pad, &st.tile_uvw, &st.pairs, &st.tile_rows, &st.bl_rows, &st.lmn, &st.coh};
    size_t off = 0;
    for (int i = 0; i < 7; ++i) {
        if (want[i] == 0) {
            bufs[i]->borrow(nullptr, 0);
            continue;
        }
        bufs[i]->borrow(static_cast<unsigned char *>(base) + off, round_up256(want[i]));
        off += round_up256(want[i]);
    }
    // Buffers the point path never reserves: borrowed with no capacity, so that any use
    // fails with an error instead of allocating.
    for (DevBuf *b : {&st.shape, &st.bl_uv, &st.coeffs, &st.coeff_range, &st.sbf, &st.accum, &st.scratch.accum,
                      &st.scratch.cls, &st.scratch.prange, &st.scratch.packed, &st.scratch.coh0, &st.scratch.cseg,
                      &st.scratch.ccls})
        b->borrow(nullptr, 0);
    return model_points_with(st, p, a, d_uvws, d_beam_jones, d_vis_fb, d_vis_fb_f64);
}

extern "C" const char *model_points(const Points *comps, const Addresses *a, const UVW *d_uvws,
                                    const JonesF64 *d_beam_jones, JonesF32 *d_vis_fb) {
    return sm_model_points_f64(comps, a, d_uvws, d_beam_jones, d_vis_fb, nullptr);
}

extern "C" const char *sm_model_gaussians_f64(const Gaussians *g, const Addresses *a, const UVW *d_uvws,
                                              const JonesF64 *d_beam_jones, JonesF32 *d_vis_fb,
                                              JonesF64 *d_vis_fb_f64) {
    if (g == nullptr || a == nullptr) return "null argument";
    const int n = g->num_power_laws + g->num_curved_power_laws + g->num_lists;
    if (n == 0) return nullptr;
    AbiState &st = abi_state();
    ModelCtx ctx;
    if (const char *err = abi_ctx(st, a, d_uvws, d_beam_jones, n, ctx)) return err;
    if (ctx.nbl == 0 || ctx.nchan == 0) return nullptr;
    if (const char *err = gather_lmn(g->num_power_laws, g->power_law_lmns, g->num_curved_power_laws,
                                     g->curved_power_law_lmns, g->num_lists, g->list_lmns, st.lmn))
        return err;
    if (const char *err = gather_shape(KIND_GAUSSIAN, g->num_power_laws, g->power_law_gps, g->num_curved_power_laws,
                                       g->curved_power_law_gps, g->num_lists, g->list_gps, st.shape))
        return err;
    TypeInputs in{};
    in.fb = FluxBuckets{g->num_power_laws, g->num_curved_power_laws, g->num_lists, g->power_law_fds,
                        g->power_law_sis, g->curved_power_law_fds, g->curved_power_law_sis, g->curved_power_law_qs,
                        g->list_fds};
    in.lmn = st.lmn.as<double4>();
    in.shape = st.shape.as<double4>();
    SM_TRY(run_type(KIND_GAUSSIAN, in, ctx, st.coh, st.scratch, d_vis_fb, d_vis_fb_f64, 0, nullptr, nullptr));
    SM_TRY(hipStreamSynchronize(0));
    return nullptr;
}

extern "C" const char *model_gaussians(const Gaussians *comps, const Addresses *a, const UVW *d_uvws,
                                       const JonesF64 *d_beam_jones, JonesF32 *d_vis_fb) {
    return sm_model_gaussians_f64(comps, a, d_uvws, d_beam_jones, d_vis_fb, nullptr);
}

extern "C" const char *sm_model_shapelets_f64(const Shapelets *s, const Addresses *a, const UVW *d_uvws,
                                              const JonesF64 *d_beam_jones, JonesF32 *d_vis_fb,
                                              JonesF64 *d_vis_fb_f64) {
    if (s == nullptr || a == nullptr) return "null argument";
    const int n_pl = s->num_power_laws, n_cpl = s->num_curved_power_laws, n_list = s->num_lists;
    const int n = n_pl + n_cpl + n_list;
    if (n == 0) return nullptr;
    if (a->d_shapelet_basis_values == nullptr) return "d_shapelet_basis_values is null";
    AbiState &st = abi_state();
    ModelCtx ctx;
    if (const char *err = abi_ctx(st, a, d_uvws, d_beam_jones, n, ctx)) return err;
    if (ctx.nbl == 0 || ctx.nchan == 0) return nullptr;
    if (const char *err = gather_lmn(n_pl, s->power_law_lmns, n_cpl, s->curved_power_law_lmns, n_list, s->list_lmns,
                                     st.lmn))
        return err;
    if (const char *err = gather_shape(KIND_SHAPELET, n_pl, s->power_law_gps, n_cpl, s->curved_power_law_gps, n_list,
                                       s->list_gps, st.shape))
        return err;
    SM_TRY(st.bl_uv.reserve(sizeof(double2) * (size_t)n * ctx.nbl));
    SM_TRY(launch_repack_shapelet_uv(ctx.nbl, n_pl, n_cpl, n_list, s->power_law_shapelet_uvs,
                                     s->curved_power_law_shapelet_uvs, s->list_shapelet_uvs, st.bl_uv.as<double2>(),
                                     0));
    // Coefficient counts per bucket -> one flattened list with (offset, count) per component.
    const int *counts_d[3] = {s->power_law_num_shapelet_coeffs, s->curved_power_law_num_shapelet_coeffs,
                              s->list_num_shapelet_coeffs};
    const ShapeletCoeff *coeffs_d[3] = {s->power_law_shapelet_coeffs, s->curved_power_law_shapelet_coeffs,
                                        s->list_shapelet_coeffs};
    const int nb[3] = {n_pl, n_cpl, n_list};
    std::vector<int2> ranges(n);
    std::vector<size_t> bucket_total(3, 0);
    int c = 0;
    size_t off = 0;
    for (int b = 0; b < 3; ++b) {
        if (nb[b] == 0) continue;
        std::vector<int> cnt(nb[b]);
        SM_TRY(hipMemcpy(cnt.data(), counts_d[b], sizeof(int) * nb[b], hipMemcpyDeviceToHost));
        for (int i = 0; i < nb[b]; ++i) {
            if (cnt[i] < 0) return "negative shapelet coefficient count";
            ranges[c++] = make_int2((int)off, cnt[i]);
            off += (size_t)cnt[i];
            bucket_total[b] += (size_t)cnt[i];
        }
    }
    SM_TRY(st.coeffs.reserve(sizeof(ShapeletCoeff) * (off > 0 ? off : 1)));
    size_t dst = 0;
    for (int b = 0; b < 3; ++b) {
        if (bucket_total[b] == 0) continue;
        SM_TRY(hipMemcpy(st.coeffs.as<ShapeletCoeff>() + dst, coeffs_d[b], sizeof(ShapeletCoeff) * bucket_total[b],
                         hipMemcpyDeviceToDevice));
        dst += bucket_total[b];
    }
    SM_TRY(st.coeff_range.reserve(sizeof(int2) * n));
    SM_TRY(hipMemcpy(st.coeff_range.ptr, ranges.data(), sizeof(int2) * n, hipMemcpyHostToDevice));
    TypeInputs in{};
    in.fb = FluxBuckets{n_pl, n_cpl, n_list, s->power_law_fds, s->power_law_sis, s->curved_power_law_fds,
                        s->curved_power_law_sis, s->curved_power_law_qs, s->list_fds};
    in.lmn = st.lmn.as<double4>();
    in.shape = st.shape.as<double4>();
    in.bl_uv = st.bl_uv.as<double2>();
    in.coeff_range = st.coeff_range.as<int2>();
    in.coeffs = st.coeffs.as<ShapeletCoeff>();
    in.coeff_count = (int)off;
    SM_TRY(run_type(KIND_SHAPELET, in, ctx, st.coh, st.scratch, d_vis_fb, d_vis_fb_f64, 0, nullptr, nullptr));
    SM_TRY(hipStreamSynchronize(0));
    return nullptr;
}

extern "C" const char *model_shapelets(const Shapelets *comps, const Addresses *a, const UVW *d_uvws,
                                       const JonesF64 *d_beam_jones, JonesF32 *d_vis_fb) {
    return sm_model_shapelets_f64(comps, a, d_uvws, d_beam_jones, d_vis_fb, nullptr);
}

// Autos over all three types with one demotion of the summed f64 accumulator, as the
// reference's GPU kernel does (src/gpu/model.cu:368-485). beam_jones holds every direction
// in the order points, Gaussians, shapelets (src/model/gpu.rs:1184-1226).
extern "C" const char *model_autos(int num_tiles, int num_freqs, const double *freqs, const Points *points,
                                   const Gaussians *gaussians, const Shapelets *shapelets,
                                   const JonesF64 *beam_jones, const int *tile_map, const int *freq_map,
                                   int num_fee_freqs, const int *tile_index_to_unflagged_tile_index_map,
                                   JonesF32 *vis_fb) {
    if (points == nullptr || gaussians == nullptr || shapelets == nullptr) return "null argument";
    if (num_tiles <= 0 || num_freqs <= 0) return nullptr;
    AbiState &st = abi_state();
    const int np = points->num_power_laws + points->num_curved_power_laws + points->num_lists;
    const int ng = gaussians->num_power_laws + gaussians->num_curved_power_laws + gaussians->num_lists;
    const int ns = shapelets->num_power_laws + shapelets->num_curved_power_laws + shapelets->num_lists;
    const int n_dir = np + ng + ns;
    if (n_dir == 0) return nullptr;
    const int nchan_pad = pad_channels(num_freqs);
    std::vector<double> f(num_freqs);
    SM_TRY(hipMemcpy(f.data(), freqs, sizeof(double) * num_freqs, hipMemcpyDeviceToHost));
    std::vector<double> fp(nchan_pad);
    pad_freqs(f.data(), num_freqs, fp.data());
    SM_TRY(st.freqs_pad.reserve(sizeof(double) * nchan_pad));
    SM_TRY(hipMemcpy(st.freqs_pad.ptr, fp.data(), sizeof(double) * nchan_pad, hipMemcpyHostToDevice));
    std::vector<int> unflagged_to_total(num_tiles);
    SM_TRY(hipMemcpy(unflagged_to_total.data(), tile_index_to_unflagged_tile_index_map, sizeof(int) * num_tiles,
                     hipMemcpyDeviceToHost));
    int max_total = 0;
    for (int t : unflagged_to_total) max_total = t > max_total ? t : max_total;
    std::vector<int> tmap(max_total + 1);
    SM_TRY(hipMemcpy(tmap.data(), tile_map, sizeof(int) * (max_total + 1), hipMemcpyDeviceToHost));
    std::vector<int> rows(num_tiles);
    for (int t = 0; t < num_tiles; ++t) rows[t] = tmap[unflagged_to_total[t]];
    SM_TRY(st.tile_rows.reserve(sizeof(int) * num_tiles));
    SM_TRY(hipMemcpy(st.tile_rows.ptr, rows.data(), sizeof(int) * num_tiles, hipMemcpyHostToDevice));
    const size_t nvis = (size_t)num_tiles * num_freqs;
    SM_TRY(st.accum.reserve(sizeof(double) * 8 * nvis));
    SM_TRY(hipMemsetAsync(st.accum.ptr, 0, sizeof(double) * 8 * nvis, 0));
    BeamDesc beam{};
    beam.jones = beam_jones;
    beam.freq_map = freq_map;
    beam.n_ufreq = num_fee_freqs;
    beam.n_dir = n_dir;
    beam.tile_rows = st.tile_rows.as<int>();
    beam.uniform_row = -1;
    const FluxBuckets fbs[3] = {
        {points->num_power_laws, points->num_curved_power_laws, points->num_lists, points->power_law_fds,
         points->power_law_sis, points->curved_power_law_fds, points->curved_power_law_sis,
         points->curved_power_law_qs, points->list_fds},
        {gaussians->num_power_laws, gaussians->num_curved_power_laws, gaussians->num_lists, gaussians->power_law_fds,
         gaussians->power_law_sis, gaussians->curved_power_law_fds, gaussians->curved_power_law_sis,
         gaussians->curved_power_law_qs, gaussians->list_fds},
        {shapelets->num_power_laws, shapelets->num_curved_power_laws, shapelets->num_lists, shapelets->power_law_fds,
         shapelets->power_law_sis, shapelets->curved_power_law_fds, shapelets->curved_power_law_sis,
         shapelets->curved_power_law_qs, shapelets->list_fds}};
    const int counts[3] = {np, ng, ns};
    int dir_off = 0;
    for (int ty = 0; ty < 3; ++ty) {
        if (counts[ty] == 0) continue;
        SM_TRY(st.coh.reserve(sizeof(double) * 8 * (size_t)counts[ty] * nchan_pad));
        SM_TRY(launch_coherency(counts[ty], num_freqs, nchan_pad, st.freqs_pad.as<double>(), fbs[ty], beam, false,
                                st.coh.as<double>(), 0));
        beam.dir_offset = dir_off;
        SM_TRY(launch_autos(num_tiles, num_freqs, counts[ty], st.coh.as<double>(), nchan_pad, beam, false, vis_fb,
                            nullptr, st.accum.as<double>(), false, 0));
        dir_off += counts[ty];
    }
    SM_TRY(launch_autos_demote((int)nvis, st.accum.as<double>(), vis_fb, nullptr, 0));
    SM_TRY(hipStreamSynchronize(0));
    return nullptr;
}

extern "C" const char *sm_analytic_beam(int32_t n_tiles, const uint32_t *delays, const double *gains,
                                        int32_t n_freqs, const double *freqs, int32_t n_dirs, const double *d_az,
                                        const double *d_za, JonesF64 *d_out) {
    if (n_tiles <= 0 || n_freqs <= 0 || n_dirs <= 0) return nullptr;
    DevBuf dd, dg, df;
    SM_TRY(dd.reserve(sizeof(unsigned) * 16 * n_tiles));
    SM_TRY(dg.reserve(sizeof(double) * 32 * n_tiles));
    SM_TRY(df.reserve(sizeof(double) * n_freqs));
    std::vector<unsigned> hd(16 * (size_t)n_tiles, 0u);
    std::vector<double> hg(32 * (size_t)n_tiles, 1.0);
    if (delays != nullptr) std::memcpy(hd.data(), delays, sizeof(unsigned) * hd.size());
    if (gains != nullptr) std::memcpy(hg.data(), gains, sizeof(double) * hg.size());
    SM_TRY(hipMemcpy(dd.ptr, hd.data(), sizeof(unsigned) * hd.size(), hipMemcpyHostToDevice));
    SM_TRY(hipMemcpy(dg.ptr, hg.data(), sizeof(double) * hg.size(), hipMemcpyHostToDevice));
    SM_TRY(hipMemcpy(df.ptr, freqs, sizeof(double) * n_freqs, hipMemcpyHostToDevice));
    SM_TRY(launch_analytic_beam(n_tiles, dd.as<unsigned>(), dg.as<double>(), n_freqs, df.as<double>(), n_dirs, d_az,
                                d_za, d_out, 0));
    SM_TRY(hipStreamSynchronize(0));
    return nullptr;
}

extern "C" const char *sm_set_class_table_limit(int64_t bytes) {
    sm::g_class_table_limit = bytes < 0 ? -1 : (long long)bytes;
    return nullptr;
}

extern "C" const char *sm_set_direct_variant(int32_t variant) {
    if (variant < 0 || variant > 1) return "variant must be 0 (fastest) or 1 (baseline per lane)";
    sm::g_direct_variant = variant;
    return nullptr;
}
