// Faster FP64 direct kernels for the common case: one beam shared by every tile (or no
// beam) and evenly spaced channels.
//
// k_direct_tiled (points, Gaussians): a block owns a 16 x 16 block of tile pairs and one
// chunk of NCH channels (kTiledNchPoint, kTiledNchGaussian). The baseline phase factorises over antennas,
//   e^{i theta_pq f/c} = e^{i theta_p f/c} e^{-i theta_q f/c},  theta_t = u_t L + v_t M + w_t N,
// (baseline uvw = tile1 - tile2, mdbook/src/defs/coords.md), so per component the block
// evaluates the start phasor and the per-channel step for its 32 tiles into LDS (64
// sincos shared by 256 baselines instead of 512), and each lane forms its baseline's
// phasors with two complex products. Diagonal tile blocks are paired so a block keeps 240
// of its 256 lanes busy.
//
// k_shapelet_chan (shapelets): one wave per baseline, one lane per channel. The basis-table
// positions x_pos = k f + 5000 of neighbouring channels fall in the same or adjacent table
// cells, so each coefficient's table reads coalesce across the wave instead of touching
// 64 unrelated cache lines.
#include "sm_internal.h"
#include "sm_jones.h"
#include "sm_math.h"
#include "sm_tiled.h"

namespace sm {

constexpr int kTG = 16;   // tiles per group
constexpr int kWin = 32;  // shapelet table window per wave and axis, cells
constexpr int kRows = 16; // shapelet table rows staged per wave
constexpr int kWinPair = 24;  // shapelet table window per wave and axis of the two-channel kernel, cells
constexpr int kCT = 32;   // components per LDS stage
// Channels per lane in k_direct_tiled. Fewer channels hold fewer
// accumulators (more waves per SIMD) but share each component's scalar coherency loads, phasor
// reads and products, and a Gaussian's three exps, over fewer channels. Measured per C2 launch
// (768 channels) with the Hermitian coherency table: points 21.4 ms with 4 channels, 19.4 ms
// with 6, 17.4 ms with 8 (167 VGPRs, 3 waves per SIMD); Gaussians 15.1 ms with 4, 12.6 ms with 6
// (161 VGPRs), 12.5 ms with 8. The FP32 Gaussians keep 6 (see launch_direct_tiled).
constexpr int kTiledNchPoint = 8;
constexpr int kTiledNchGaussian = 6;
static_assert(kTiledNchPoint <= SM_NCH && kTiledNchGaussian <= SM_NCH,
              "the padded channel count covers whole chunks of SM_NCH channels");

// The kernels below come in two arithmetic types T. double: the FP64 direct kernels. float:
// the precision path's kernels, in which the phases, their range reduction mod 2 pi, the table
// positions of shapelets and the Gaussian exponents are still formed in FP64, and everything
// after (phasor products, envelopes, coherency products, accumulation) runs in FP32.
template <class T> struct Vec4;
template <> struct Vec4<double> { using type = double4; };
template <> struct Vec4<float> { using type = float4; };

__device__ __forceinline__ double fma_t(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// e^x of an FP64 argument: FP64 exp, or the hardware base-2 exponential in FP32 (an argument
// below the FP32 range gives 0).
__device__ __forceinline__ void exp_of(double x, double &out) { out = exp(x); }
__device__ __forceinline__ void exp_of(double x, float &out) {
    out = __builtin_amdgcn_exp2f((float)(x * 1.4426950408889634));
}

// sin and cos of an FP64 phase in the kernel's arithmetic (the shapelet kernel's per-channel
// phases). FP64: table_sincos. FP32: x is reduced mod pi/2 in FP64 exactly as fast_sincos does,
// and the reduced |r| <= pi/4 goes through the Taylor series of sin to r^9 and of cos to r^10
// in FP32 (truncation below 2e-9, under the FP32 rounding).
__device__ __forceinline__ void sincos_phase(double x, double &s, double &c) { table_sincos(x, kSinCos32, s, c); }
__device__ __forceinline__ void sincos_phase(double x, float &s, float &c) {
    const double k = rint(x * 0.6366197723675814);
    double r = fma(-k, 1.5707963267948966, x);
    r = fma(-k, 6.123233995736766e-17, r);
    const float rf = (float)r, r2 = rf * rf;
    // coefficients (-1)^n / (2n+1)! and (-1)^n / (2n)!
    float ps = 2.7557319e-06f;
    ps = __builtin_fmaf(ps, r2, -1.9841270e-04f);
    ps = __builtin_fmaf(ps, r2, 8.3333333e-03f);
    ps = __builtin_fmaf(ps, r2, -1.6666667e-01f);
    const float sn = __builtin_fmaf(rf * r2, ps, rf);
    float pc = -2.7557319e-07f;
    pc = __builtin_fmaf(pc, r2, 2.4801587e-05f);
    pc = __builtin_fmaf(pc, r2, -1.3888889e-03f);
    pc = __builtin_fmaf(pc, r2, 4.1666667e-02f);
    pc = __builtin_fmaf(pc, r2, -0.5f);
    const float cn = __builtin_fmaf(r2, pc, 1.0f);
    const int q = (int)(long long)k & 3;
    const float a = (q & 1) ? cn : sn;  // sin x = sin r, cos r, -sin r, -cos r for q = 0..3
    const float b = (q & 1) ? sn : cn;  // cos x = cos r, -sin r, -cos r, sin r
    s = (q & 2) ? -a : a;
    c = ((q + 1) & 2) ? -b : b;
}

// FP32 kernels sum in two levels: a stage's components (at most 32) add into the inner
// accumulators, which are then added into the outer ones, so the rounding of a sum over N
// components grows like sqrt(N / stage) rather than sqrt(N). FP64 kernels sum in one level.
template <class T> struct TwoLevel { static constexpr bool on = false; };
template <> struct TwoLevel<float> { static constexpr bool on = true; };

template <class T, int N>
__device__ __forceinline__ void flush_stage(T (&acc)[N][8], T (&tot)[N][8]) {
    if (!TwoLevel<T>::on) return;
#pragma unroll
    for (int k = 0; k < N; ++k)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            tot[k][e] += acc[k][e];
            acc[k][e] = T(0);
        }
}

template <class T> __device__ __forceinline__ T two_level_sum(T acc, T tot) { return TwoLevel<T>::on ? tot + acc : acc; }

template <class T> __device__ __forceinline__ const T *pick(const double *d, const float *f);
template <> __device__ __forceinline__ const double *pick<double>(const double *d, const float *) { return d; }
template <> __device__ __forceinline__ const float *pick<float>(const double *, const float *f) { return f; }

template <int KIND, int NCH, class T>
__global__ __launch_bounds__(256) void k_direct_tiled(const TiledArgs a) {
    using V4 = typename Vec4<T>::type;
    __shared__ V4 s_ph[kCT][32];
    __shared__ double s_uvw[32][3];
    const int total = a.npairs * a.nchunks;
    int b = blockIdx.x;
    if ((total & 7) == 0) b = (b & 7) * (total >> 3) + (b >> 3);
    const int chunk = b / a.npairs;
    int f0, cnt;
    chunk_range<NCH>(a.chunks, chunk, a.nchan, f0, cnt);
    const int4 pr = a.pairs[b - chunk * a.npairs];
    const int tid = threadIdx.x, r = tid >> 4, cq = tid & 15;
    if (tid < 32) {
        const int g = tid < 16 ? pr.y : pr.z;
        int t = g * kTG + (tid & 15);
        if (g < 0 || t >= a.ntiles) t = 0;
        s_uvw[tid][0] = a.tile_uvw[3 * t];
        s_uvw[tid][1] = a.tile_uvw[3 * t + 1];
        s_uvw[tid][2] = a.tile_uvw[3 * t + 2];
    }
    int p, q, sp, sq;
    bool valid;
    if (pr.x == 0) {
        p = pr.y * kTG + r;
        q = pr.z * kTG + cq;
        sp = r;
        sq = 16 + cq;
        valid = q < a.ntiles;
    } else if (r < cq) {
        p = pr.y * kTG + r;
        q = pr.y * kTG + cq;
        sp = r;
        sq = cq;
        valid = q < a.ntiles;
    } else if (r > cq && pr.z >= 0) {
        p = pr.z * kTG + cq;
        q = pr.z * kTG + r;
        sp = 16 + cq;
        sq = 16 + r;
        valid = q < a.ntiles;
    } else {
        p = 0;
        q = 1;
        sp = 0;
        sq = 1;
        valid = false;
    }
    __syncthreads();
    const double bu = s_uvw[sp][0] - s_uvw[sq][0];
    const double bv = s_uvw[sp][1] - s_uvw[sq][1];

    const double *__restrict__ freqs = a.freqs;
    const double F0 = freqs[f0];
    const double DF = freqs[f0 + 1] - F0;
    const double f0c = F0 / SM_VEL_C;
    const double dfc = DF / SM_VEL_C;
    const T *__restrict__ coh = pick<T>(a.coh, a.coh32);
    const double4 *__restrict__ lmn4 = a.lmn;
    const double4 *__restrict__ shape = a.shape;

    T acc[NCH][8], tot[NCH][8];
#pragma unroll
    for (int k = 0; k < NCH; ++k)
#pragma unroll
        for (int e = 0; e < 8; ++e) acc[k][e] = tot[k][e] = T(0);

    for (int c0 = 0; c0 < a.ncomp; c0 += kCT) {
        const int nc = a.ncomp - c0 < kCT ? a.ncomp - c0 : kCT;
        // Per-tile phasors (start, step) of this stage's components, phases reduced in FP64.
        for (int e = tid; e < kCT * 32; e += 256) {
            const int slot = e & 31, cc = e >> 5;
            double ps = 0.0, pc = 1.0, qs = 0.0, qc = 1.0;
            if (cc < nc) {
                const double4 l = lmn4[c0 + cc];
                const double th = s_uvw[slot][0] * l.x + s_uvw[slot][1] * l.y + s_uvw[slot][2] * l.z;
                table_sincos(th * f0c, kSinCos32, ps, pc);
                table_sincos(th * dfc, kSinCos32, qs, qc);
            }
            V4 ph;
            ph.x = (T)pc;
            ph.y = (T)ps;
            ph.z = (T)qc;
            ph.w = (T)qs;
            s_ph[cc][slot] = ph;
        }
        __syncthreads();
        // Component cc + 1's tile phasors, coherencies and Gaussian shape are loaded while
        // component cc's products run. The explicit wait comes first: a scalar load's result
        // can only be waited for with every outstanding scalar and LDS load, so the next
        // component's loads are issued after it and are not waited for until the next
        // iteration.
        for (int cc = 0; cc < nc; ++cc) {
            const int c = c0 + cc;
            const V4 A = s_ph[cc][sp];
            const V4 B = s_ph[cc][sq];
            T zr = A.x * B.x + A.y * B.y;
            T zi = A.y * B.x - A.x * B.y;
            const T rr = A.z * B.z + A.w * B.w;
            const T ri = A.w * B.z - A.z * B.w;
            T env = T(1), er = T(1), eg = T(1);
            if (KIND == KIND_GAUSSIAN) {
                const double4 g = shape[c];
                const double qf = g.x * (bu * bu) + g.y * (bu * bv) + g.z * (bv * bv);
                exp_of(qf * (F0 * F0), env);
                exp_of(qf * (DF * (2.0 * F0 + DF)), er);
                exp_of(qf * (2.0 * DF * DF), eg);
            }
            // The table is the Hermitian form (Re XX, Re XY, Im XY, Re YY) of J S J^H (see
            // TiledArgs), 32 contiguous bytes per channel. Per visibility: V_XX = sum C_XX z,
            // A = sum Re(C_XY) z, B = sum Im(C_XY) z, V_YY = sum C_YY z (8 FMAs per component
            // instead of 16); V_XY = A + i B and V_YX = A - i B are formed once at the end.
            const T *cc4 = coh + ((size_t)c * a.nchan_pad + f0) * 4;
#pragma unroll
            for (int k = 0; k < NCH; ++k) {
                T pr_ = zr, pi_ = zi;
                if (KIND == KIND_GAUSSIAN) {
                    pr_ *= env;
                    pi_ *= env;
                }
                const T cxx = cc4[k * 4], cxyr = cc4[k * 4 + 1], cxyi = cc4[k * 4 + 2], cyy = cc4[k * 4 + 3];
                acc[k][0] = fma_t(cxx, pr_, acc[k][0]);
                acc[k][1] = fma_t(cxx, pi_, acc[k][1]);
                acc[k][2] = fma_t(cxyr, pr_, acc[k][2]);
                acc[k][3] = fma_t(cxyr, pi_, acc[k][3]);
                acc[k][4] = fma_t(cxyi, pr_, acc[k][4]);
                acc[k][5] = fma_t(cxyi, pi_, acc[k][5]);
                acc[k][6] = fma_t(cyy, pr_, acc[k][6]);
                acc[k][7] = fma_t(cyy, pi_, acc[k][7]);
                const T nr = zr * rr - zi * ri;
                zi = zr * ri + zi * rr;
                zr = nr;
                if (KIND == KIND_GAUSSIAN) {
                    env *= er;
                    er *= eg;
                }
            }
        }
        flush_stage(acc, tot);
        __syncthreads();
    }
    if (!valid) return;
    if (a.tile_map != nullptr) {
        p = a.tile_map[p];
        q = a.tile_map[q];
    }
    const size_t nt = a.tile_map != nullptr ? (size_t)a.ntiles_bl : (size_t)a.ntiles;
    const size_t bl = (size_t)p * nt - (size_t)p * (p + 1) / 2 + (size_t)(q - p - 1);
#pragma unroll
    for (int k = 0; k < NCH; ++k) {
        const int f = f0 + k;
        if (k >= cnt) break;
        const size_t o = (size_t)f * a.nbl + bl;
        T s8[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) s8[e] = two_level_sum(acc[k][e], tot[k][e]);
        // XX, XY = A + i B, YX = A - i B, YY.
        const T v[8] = {s8[0], s8[1], s8[2] - s8[5], s8[3] + s8[4], s8[2] + s8[5], s8[3] - s8[4], s8[6], s8[7]};
        float *v32 = reinterpret_cast<float *>(&a.vis32[o]);
#pragma unroll
        for (int e = 0; e < 8; ++e) v32[e] += (float)v[e];
        if (a.vis64 != nullptr) {
            double *v64 = reinterpret_cast<double *>(&a.vis64[o]);
#pragma unroll
            for (int e = 0; e < 8; ++e) v64[e] += (double)v[e];
        }
    }
}

// Tile-group pairs: off-diagonal (0, g0, g1) with g0 < g1, then diagonal groups paired from
// both ends (1, g, G-1-g), the middle one alone (1, g, -1) when G is odd.
std::vector<int4> tile_pairs(int ntiles) {
    const int G = (ntiles + kTG - 1) / kTG;
    std::vector<int4> out;
    for (int g0 = 0; g0 < G; ++g0)
        for (int g1 = g0 + 1; g1 < G; ++g1) out.push_back(make_int4(0, g0, g1, 0));
    for (int g = 0; g < G / 2; ++g) out.push_back(make_int4(1, g, G - 1 - g, 0));
    if (G % 2 == 1) out.push_back(make_int4(1, G / 2, -1, 0));
    return out;
}

hipError_t launch_direct_tiled(Kind kind, const TiledArgs &a_in, hipStream_t stream) {
    if (a_in.ncomp <= 0 || a_in.nbl <= 0 || a_in.nchan <= 0) return hipSuccess;
    TiledArgs a = a_in;
    // FP32 Gaussians take 6 channels per lane too (142 VGPRs with the two-level accumulators,
    // 3 waves per SIMD). Measured per C5 timestep: 78.4 ms with 4, 76.2 ms with 6, 115.9 ms
    // with 8.
    const int nch = kind == KIND_GAUSSIAN ? kTiledNchGaussian : kTiledNchPoint;
    a.nchunks = chunk_count(a.chunks, nch, a.nchan);
    const int total = a.npairs * a.nchunks;
    if (a.coh32 != nullptr) {
        if (kind == KIND_GAUSSIAN)
            hipLaunchKernelGGL((k_direct_tiled<KIND_GAUSSIAN, kTiledNchGaussian, float>), dim3(total), dim3(256), 0,
                               stream, a);
        else
            hipLaunchKernelGGL((k_direct_tiled<KIND_POINT, kTiledNchPoint, float>), dim3(total), dim3(256), 0, stream,
                               a);
    } else if (kind == KIND_GAUSSIAN) {
        hipLaunchKernelGGL((k_direct_tiled<KIND_GAUSSIAN, kTiledNchGaussian, double>), dim3(total), dim3(256), 0,
                           stream, a);
    } else {
        hipLaunchKernelGGL((k_direct_tiled<KIND_POINT, kTiledNchPoint, double>), dim3(total), dim3(256), 0, stream,
                           a);
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------ pairs with own beams
constexpr int kPJC = 8;  // components per LDS stage
constexpr int kPJN = 4;  // channels per lane

// 2x2 complex (A B^H)[i][j] = sum_k A[i][k] conj(B[j][k]) added to acc with phasor z:
// acc += z (A B^H).
template <class T>
__device__ __forceinline__ void acc_abh(const T *a, const T *b, T zr, T zi, T *acc) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            T mr = T(0), mi = T(0);
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const T ar = a[4 * i + 2 * k], ai = a[4 * i + 2 * k + 1];
                const T br = b[4 * j + 2 * k], bi = b[4 * j + 2 * k + 1];
                mr = fma_t(ar, br, mr);
                mr = fma_t(ai, bi, mr);
                mi = fma_t(ai, br, mi);
                mi = fma_t(-ar, bi, mi);
            }
            const int e = 2 * (2 * i + j);
            acc[e] = fma_t(mr, zr, acc[e]);
            acc[e] = fma_t(-mi, zi, acc[e]);
            acc[e + 1] = fma_t(mr, zi, acc[e + 1]);
            acc[e + 1] = fma_t(mi, zr, acc[e + 1]);
        }
}

// 2x2 complex acc += A B^H.
template <class T> __device__ __forceinline__ void acc_ab(const T *a, const T *b, T *acc) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int e = 2 * (2 * i + j);
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const T ar = a[4 * i + 2 * k], ai = a[4 * i + 2 * k + 1];
                const T br = b[4 * j + 2 * k], bi = b[4 * j + 2 * k + 1];
                acc[e] = fma_t(ar, br, acc[e]);
                acc[e] = fma_t(ai, bi, acc[e]);
                acc[e + 1] = fma_t(ai, br, acc[e + 1]);
                acc[e + 1] = fma_t(-ar, bi, acc[e + 1]);
            }
        }
}

// Points: the baseline phasor factorises over the two tiles, z = K_p conj(K_q), so the staged
// factors carry it: A = K_p J_p S (row slots) and B = K_q J_q (column slots) per channel, with
// K_t = exp(i theta_t f / c) formed in FP64 per (component, tile slot, channel), and each lane
// accumulates A B^H (32 FMAs per channel and component instead of 48). Gaussians keep z
// outside: their envelope depends on the baseline, not on the tiles.
template <int KIND, class T>
__global__ __launch_bounds__(256) void k_pair_jones(const PairJonesArgs a, int ngroups) {
    using V4 = typename Vec4<T>::type;
    __shared__ V4 s_ph[kPJC][32];
    // [component][channel][slot]: slots 0..15 hold J_p S of the row tiles, 16..31 J_q of the
    // column tiles; a wave's 16 column lanes read 16 consecutive slots.
    __shared__ T s_m[kPJC][kPJN][32][8];
    __shared__ double s_uvw[32][3];
    __shared__ int s_row[32];
    const int ngpairs = ngroups * (ngroups + 1) / 2;
    // Block = (component slice, channel chunk, tile-group pair), the pair fastest.
    const int per_slice = ngpairs * a.nchunks;
    const int slice = blockIdx.x / per_slice;
    const int in_slice = blockIdx.x - slice * per_slice;
    const int chunk = in_slice / ngpairs;
    const int slice_len = (a.ncomp + a.csplit - 1) / a.csplit;
    const int c_begin = slice * slice_len;
    const int c_end = a.ncomp < c_begin + slice_len ? a.ncomp : c_begin + slice_len;
    // Tile-group pair (g0 <= g1) of this block, in row-major order of the upper triangle.
    int2 gp = make_int2(0, 0);
    {
        int rem = in_slice - chunk * ngpairs, g0 = 0;
        while (rem >= ngroups - g0) {
            rem -= ngroups - g0;
            ++g0;
        }
        gp = make_int2(g0, g0 + rem);
    }
    int f0, cnt;
    chunk_range<kPJN>(a.chunks, chunk, a.nchan, f0, cnt);
    const int tid = threadIdx.x, r = tid >> 4, cq = tid & 15;
    if (tid < 32) {
        const int g = tid < 16 ? gp.x : gp.y;
        const int i = g * 16 + (tid & 15);
        const int t = a.tiles[i < a.ntiles_list ? i : a.ntiles_list - 1];
        s_uvw[tid][0] = a.tile_uvw[3 * t];
        s_uvw[tid][1] = a.tile_uvw[3 * t + 1];
        s_uvw[tid][2] = a.tile_uvw[3 * t + 2];
        s_row[tid] = a.beam.tile_rows[t];
    }
    const int ip = gp.x * 16 + r, iq = gp.y * 16 + cq;
    const bool valid = ip < a.ntiles_list && iq < a.ntiles_list && ip < iq;
    __syncthreads();
    const double bu = s_uvw[r][0] - s_uvw[16 + cq][0];
    const double bv = s_uvw[r][1] - s_uvw[16 + cq][1];
    const double F0 = a.freqs[f0];
    const double DF = a.freqs[f0 + 1] - F0;
    const double f0c = F0 / SM_VEL_C, dfc = DF / SM_VEL_C;

    T acc[kPJN][8], tot[kPJN][8];
#pragma unroll
    for (int k = 0; k < kPJN; ++k)
#pragma unroll
        for (int e = 0; e < 8; ++e) acc[k][e] = tot[k][e] = T(0);

    for (int c0 = c_begin; c0 < c_end; c0 += kPJC) {
        const int nc = c_end - c0 < kPJC ? c_end - c0 : kPJC;
        if (KIND != KIND_POINT) {
            // One phasor pair per (component, tile slot): 8 x 32 = 256 threads.
            const int cc = tid >> 5, slot = tid & 31;
            double ps = 0.0, pc = 1.0, qs = 0.0, qc = 1.0;
            if (cc < nc) {
                const double4 l = a.lmn[c0 + cc];
                const double th = s_uvw[slot][0] * l.x + s_uvw[slot][1] * l.y + s_uvw[slot][2] * l.z;
                table_sincos(th * f0c, kSinCos32, ps, pc);
                table_sincos(th * dfc, kSinCos32, qs, qc);
            }
            V4 ph;
            ph.x = (T)pc;
            ph.y = (T)ps;
            ph.z = (T)qc;
            ph.w = (T)qs;
            s_ph[cc][slot] = ph;
        }
        // J_p S (row slots) and J_q (column slots) per (component, slot, channel).
        for (int w = tid; w < kPJC * 32 * kPJN; w += 256) {
            const int k = w % kPJN, slot = (w / kPJN) & 31, cc = w / (kPJN * 32);
            const int f = f0 + k < a.nchan ? f0 + k : a.nchan - 1;
            T *o = s_m[cc][k][slot];
            if (cc >= nc) {
#pragma unroll
                for (int e = 0; e < 8; ++e) o[e] = T(0);
                continue;
            }
            const int c = c0 + cc;
            double j[8];
            load_jones(&a.beam.jones[((size_t)s_row[slot] * a.beam.n_ufreq + a.beam.freq_map[f]) * a.beam.n_dir +
                                     a.beam.dir_offset + c],
                       j);
            double ks = 0.0, kc = 1.0;  // the tile phasor K_t at this channel (points)
            if (KIND == KIND_POINT) {
                const double4 l = a.lmn[c];
                const double th = s_uvw[slot][0] * l.x + s_uvw[slot][1] * l.y + s_uvw[slot][2] * l.z;
                table_sincos(th * (a.freqs[f] / SM_VEL_C), kSinCos32, ks, kc);
            }
            if (slot < 16) {
                double sv[8];
                load_jones(reinterpret_cast<const JonesF64 *>(a.coh + ((size_t)c * a.nchan_pad + f) * 8), sv);
                // (J S)[i][l] = sum_k J[i][k] S[k][l]
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int l = 0; l < 2; ++l) {
                        double re = 0.0, im = 0.0;
#pragma unroll
                        for (int kk = 0; kk < 2; ++kk) {
                            const double jr = j[4 * i + 2 * kk], ji = j[4 * i + 2 * kk + 1];
                            const double sr = sv[4 * kk + 2 * l], si = sv[4 * kk + 2 * l + 1];
                            re += jr * sr - ji * si;
                            im += jr * si + ji * sr;
                        }
                        if (KIND == KIND_POINT) {
                            o[4 * i + 2 * l] = (T)(re * kc - im * ks);
                            o[4 * i + 2 * l + 1] = (T)(re * ks + im * kc);
                        } else {
                            o[4 * i + 2 * l] = (T)re;
                            o[4 * i + 2 * l + 1] = (T)im;
                        }
                    }
            } else if (KIND == KIND_POINT) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    o[2 * e] = (T)(j[2 * e] * kc - j[2 * e + 1] * ks);
                    o[2 * e + 1] = (T)(j[2 * e] * ks + j[2 * e + 1] * kc);
                }
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) o[e] = (T)j[e];
            }
        }
        __syncthreads();
        if (KIND == KIND_POINT) {
            for (int cc = 0; cc < nc; ++cc)
#pragma unroll
                for (int k = 0; k < kPJN; ++k) acc_ab<T>(s_m[cc][k][r], s_m[cc][k][16 + cq], acc[k]);
        } else
        for (int cc = 0; cc < nc; ++cc) {
            const V4 A = s_ph[cc][r];
            const V4 B = s_ph[cc][16 + cq];
            T zr = A.x * B.x + A.y * B.y;
            T zi = A.y * B.x - A.x * B.y;
            const T rr = A.z * B.z + A.w * B.w;
            const T ri = A.w * B.z - A.z * B.w;
            T env = T(1), er = T(1), eg = T(1);
            if (KIND == KIND_GAUSSIAN) {
                const double4 g = a.shape[c0 + cc];
                const double qf = g.x * (bu * bu) + g.y * (bu * bv) + g.z * (bv * bv);
                exp_of(qf * (F0 * F0), env);
                exp_of(qf * (DF * (2.0 * F0 + DF)), er);
                exp_of(qf * (2.0 * DF * DF), eg);
            }
#pragma unroll
            for (int k = 0; k < kPJN; ++k) {
                T pr = zr, pi = zi;
                if (KIND == KIND_GAUSSIAN) {
                    pr *= env;
                    pi *= env;
                }
                acc_abh<T>(s_m[cc][k][r], s_m[cc][k][16 + cq], pr, pi, acc[k]);
                const T nr = zr * rr - zi * ri;
                zi = zr * ri + zi * rr;
                zr = nr;
                if (KIND == KIND_GAUSSIAN) {
                    env *= er;
                    er *= eg;
                }
            }
        }
        flush_stage(acc, tot);
        __syncthreads();
    }
    if (!valid) return;
    if (a.csplit > 1) {
        // This slice's sums into accum[f][pair][8]; pairs in upper-triangle order of the list.
        const int n = a.ntiles_list;
        const size_t npairs = (size_t)n * (n - 1) / 2;
        const size_t t = (size_t)ip * n - (size_t)ip * (ip + 1) / 2 + (size_t)(iq - ip - 1);
#pragma unroll
        for (int k = 0; k < kPJN; ++k) {
            if (k >= cnt) break;
            double *dst = a.accum + ((size_t)(f0 + k) * npairs + t) * 8;
#pragma unroll
            for (int e = 0; e < 8; ++e) unsafeAtomicAdd(&dst[e], (double)two_level_sum(acc[k][e], tot[k][e]));
        }
        return;
    }
    const int p = a.tiles[ip], q = a.tiles[iq];
    const size_t nt = (size_t)a.ntiles_bl;
    const size_t bl = (size_t)p * nt - (size_t)p * (p + 1) / 2 + (size_t)(q - p - 1);
#pragma unroll
    for (int k = 0; k < kPJN; ++k) {
        if (k >= cnt) break;
        const size_t o = (size_t)(f0 + k) * a.nbl + bl;
        float *v32 = reinterpret_cast<float *>(&a.vis32[o]);
#pragma unroll
        for (int e = 0; e < 8; ++e) v32[e] += (float)two_level_sum(acc[k][e], tot[k][e]);
        if (a.vis64 != nullptr) {
            double *v64 = reinterpret_cast<double *>(&a.vis64[o]);
#pragma unroll
            for (int e = 0; e < 8; ++e) v64[e] += (double)two_level_sum(acc[k][e], tot[k][e]);
        }
    }
}

// vis[f][bl(pair)] += accum[f][pair] (FP64, and its FP32 rounding), pairs in upper-triangle
// order of the tile list.
__global__ void k_pair_finish(int nchan, int n, const int *__restrict__ tiles, int ntiles_bl, int nbl,
                              const double *__restrict__ accum, JonesF32 *vis32, JonesF64 *vis64) {
    const size_t npairs = (size_t)n * (n - 1) / 2;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nchan * npairs) return;
    const int f = (int)(i / npairs);
    int ip, iq;
    triangle_pair((int)(i - (size_t)f * npairs), n, ip, iq);
    const size_t p = (size_t)tiles[ip], q = (size_t)tiles[iq];
    const size_t o = (size_t)f * nbl + p * ntiles_bl - p * (p + 1) / 2 + (q - p - 1);
    const double *src = accum + i * 8;
    float *v32 = reinterpret_cast<float *>(&vis32[o]);
#pragma unroll
    for (int e = 0; e < 8; ++e) v32[e] += (float)src[e];
    if (vis64 != nullptr) {
        double *v64 = reinterpret_cast<double *>(&vis64[o]);
#pragma unroll
        for (int e = 0; e < 8; ++e) v64[e] += src[e];
    }
}

__global__ void k_zero_accum(size_t n, double *x) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        x[i] = 0.0;
}

hipError_t launch_pair_jones(Kind kind, const PairJonesArgs &a_in, DevBuf *accum, hipStream_t stream) {
    PairJonesArgs a = a_in;
    if (a.ncomp <= 0 || a.ntiles_list < 2 || a.nchan <= 0) return hipSuccess;
    if (kind == KIND_SHAPELET) return hipErrorInvalidValue;
    const int G = (a.ntiles_list + 15) / 16;
    a.nchunks = chunk_count(a.chunks, kPJN, a.nchan);
    const long blocks = (long)(G * (G + 1) / 2) * a.nchunks;
    // About 2048 blocks (8 per CU), each slice at least 64 components.
    long split = (2048 + blocks - 1) / blocks;
    split = std::min<long>(split, std::max(1, a.ncomp / (8 * kPJC)));
    split = std::max(1L, std::min(split, 64L));
    const size_t npairs = (size_t)a.ntiles_list * (a.ntiles_list - 1) / 2;
    const size_t nacc = (size_t)a.nchan * npairs * 8;
    a.csplit = 1;
    a.accum = nullptr;
    if (split > 1 && accum != nullptr) {
        if (accum->reserve(sizeof(double) * nacc) == hipSuccess) {
            a.csplit = (int)split;
            a.accum = accum->as<double>();
            hipLaunchKernelGGL(k_zero_accum, dim3(1024), dim3(256), 0, stream, nacc, a.accum);
        } else {
            (void)hipGetLastError();  // no room for the partial sums: unsplit launch
        }
    }
    const long total = blocks * a.csplit;
    if (a.fp32) {
        if (kind == KIND_GAUSSIAN)
            hipLaunchKernelGGL((k_pair_jones<KIND_GAUSSIAN, float>), dim3((unsigned)total), dim3(256), 0, stream, a, G);
        else
            hipLaunchKernelGGL((k_pair_jones<KIND_POINT, float>), dim3((unsigned)total), dim3(256), 0, stream, a, G);
    } else if (kind == KIND_GAUSSIAN) {
        hipLaunchKernelGGL((k_pair_jones<KIND_GAUSSIAN, double>), dim3((unsigned)total), dim3(256), 0, stream, a, G);
    } else {
        hipLaunchKernelGGL((k_pair_jones<KIND_POINT, double>), dim3((unsigned)total), dim3(256), 0, stream, a, G);
    }
    if (a.csplit > 1) {
        const size_t n = (size_t)a.nchan * npairs;
        hipLaunchKernelGGL(k_pair_finish, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, a.nchan,
                           a.ntiles_list, a.tiles, a.ntiles_bl, a.nbl, a.accum, a.vis32, a.vis64);
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------ row tiles x class 0
// k_direct_rect: a block owns 4 row tiles (one per wave) and 64 class-0 column tiles (one
// per lane) for one channel chunk. Every lane of a wave shares its row's class, so the
// coherency J_r S J_0^H arrives through scalar loads as in k_direct_tiled. The kernel forms
// V_pq for the ordered pair (p = row tile, q = column tile); where the array orders the
// baseline the other way (q < p) it stores V_qp = V_pq^H, which holds because the
// instrumental coherency S is Hermitian (XY and YX are conjugates).
constexpr int kRC = 16;  // components per LDS stage
constexpr int kRT = 4 + 64;  // tile slots per block: 4 rows, then 64 columns
// Gaussians here take 6 channels per lane. Measured per C5 timestep (256 tiles, 384 channels),
// FP64: 48.1 ms with 4, 42.4 ms with 6 (151 VGPRs, 3 waves per SIMD); FP32: 28.3 ms with 4,
// 26.8 ms with 6.
constexpr int kRectNchGaussian = 6;
// Points: the full-form J_r S J_0^H of a row is 64 bytes per channel of scalar loads, so wider
// chunks cost more scalar registers and waits than they share. Measured per C5 FP64 timestep:
// 138.1 ms with 4 channels per lane, 137.0 ms with 6, 149.9 ms with 8.
constexpr int kRectNchPoint = 4;

template <int KIND, int NCH, class T>
__global__ __launch_bounds__(256) void k_direct_rect(const RectArgs a) {
    using V4 = typename Vec4<T>::type;
    __shared__ V4 s_ph[kRC][kRT];
    __shared__ double s_uvw[kRT][3];
    const int nrg = (a.n_row + 3) / 4, ncg = (a.n_col + 63) / 64;
    const int b = blockIdx.x;
    const int chunk = b / (nrg * ncg);
    int f0, cnt;
    chunk_range<NCH>(a.chunks, chunk, a.nchan, f0, cnt);
    const int rem = b - chunk * (nrg * ncg);
    const int rg = rem / ncg, cg = rem - rg * ncg;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    if (tid < kRT) {
        int t;
        if (tid < 4) {
            const int i = rg * 4 + tid;
            t = a.row_tiles[i < a.n_row ? i : a.n_row - 1];
        } else {
            const int j = cg * 64 + tid - 4;
            t = a.col_tiles[j < a.n_col ? j : a.n_col - 1];
        }
        s_uvw[tid][0] = a.tile_uvw[3 * t];
        s_uvw[tid][1] = a.tile_uvw[3 * t + 1];
        s_uvw[tid][2] = a.tile_uvw[3 * t + 2];
    }
    const int ri = rg * 4 + w, cj = cg * 64 + lane;
    const bool row_ok = ri < a.n_row;
    const bool valid = row_ok && cj < a.n_col;
    const int cls = __builtin_amdgcn_readfirstlane(row_ok ? a.row_cls[ri] : 0);
    __syncthreads();
    const int sp = w, sq = 4 + lane;
    const double bu = s_uvw[sp][0] - s_uvw[sq][0];
    const double bv = s_uvw[sp][1] - s_uvw[sq][1];

    const double *__restrict__ freqs = a.freqs;
    const double F0 = freqs[f0];
    const double DF = freqs[f0 + 1] - F0;
    const double f0c = F0 / SM_VEL_C;
    const double dfc = DF / SM_VEL_C;
    const T *__restrict__ coh = pick<T>(a.coh_cls, a.coh_cls32) + (size_t)cls * a.ncomp * a.nchan_pad * 8;
    const double4 *__restrict__ lmn4 = a.lmn;
    const double4 *__restrict__ shape = a.shape;

    T acc[NCH][8], tot[NCH][8];
#pragma unroll
    for (int k = 0; k < NCH; ++k)
#pragma unroll
        for (int e = 0; e < 8; ++e) acc[k][e] = tot[k][e] = T(0);

    for (int c0 = 0; c0 < a.ncomp; c0 += kRC) {
        const int nc = a.ncomp - c0 < kRC ? a.ncomp - c0 : kRC;
        for (int e = tid; e < kRC * kRT; e += 256) {
            const int cc = e / kRT, slot = e - cc * kRT;
            double ps = 0.0, pc = 1.0, qs = 0.0, qc = 1.0;
            if (cc < nc) {
                const double4 l = lmn4[c0 + cc];
                const double th = s_uvw[slot][0] * l.x + s_uvw[slot][1] * l.y + s_uvw[slot][2] * l.z;
                table_sincos(th * f0c, kSinCos32, ps, pc);
                table_sincos(th * dfc, kSinCos32, qs, qc);
            }
            V4 ph;
            ph.x = (T)pc;
            ph.y = (T)ps;
            ph.z = (T)qc;
            ph.w = (T)qs;
            s_ph[cc][slot] = ph;
        }
        __syncthreads();
        if (row_ok) {
            for (int cc = 0; cc < nc; ++cc) {
                const int c = c0 + cc;
                const V4 A = s_ph[cc][sp];
                const V4 B = s_ph[cc][sq];
                T zr = A.x * B.x + A.y * B.y;
                T zi = A.y * B.x - A.x * B.y;
                const T rr = A.z * B.z + A.w * B.w;
                const T ri_ = A.w * B.z - A.z * B.w;
                T env = T(1), er = T(1), eg = T(1);
                if (KIND == KIND_GAUSSIAN) {
                    const double4 g = shape[c];
                    const double qf = g.x * (bu * bu) + g.y * (bu * bv) + g.z * (bv * bv);
                    exp_of(qf * (F0 * F0), env);
                    exp_of(qf * (DF * (2.0 * F0 + DF)), er);
                    exp_of(qf * (2.0 * DF * DF), eg);
                }
                const T *cc8 = coh + ((size_t)c * a.nchan_pad + f0) * 8;
#pragma unroll
                for (int k = 0; k < NCH; ++k) {
                    T pr_ = zr, pi_ = zi;
                    if (KIND == KIND_GAUSSIAN) {
                        pr_ *= env;
                        pi_ *= env;
                    }
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const T cr = cc8[k * 8 + 2 * e], ci = cc8[k * 8 + 2 * e + 1];
                        acc[k][2 * e] = fma_t(cr, pr_, acc[k][2 * e]);
                        acc[k][2 * e] = fma_t(-ci, pi_, acc[k][2 * e]);
                        acc[k][2 * e + 1] = fma_t(cr, pi_, acc[k][2 * e + 1]);
                        acc[k][2 * e + 1] = fma_t(ci, pr_, acc[k][2 * e + 1]);
                    }
                    const T nr = zr * rr - zi * ri_;
                    zi = zr * ri_ + zi * rr;
                    zr = nr;
                    if (KIND == KIND_GAUSSIAN) {
                        env *= er;
                        er *= eg;
                    }
                }
            }
        }
        flush_stage(acc, tot);
        __syncthreads();
    }
    if (!valid) return;
    const int p = a.row_tiles[ri], q = a.col_tiles[cj];
    const bool flip = q < p;  // array baseline (q, p): store V_pq^H
    const int lo = flip ? q : p, hi = flip ? p : q;
    const size_t nt = (size_t)a.ntiles_bl;
    const size_t bl = (size_t)lo * nt - (size_t)lo * (lo + 1) / 2 + (size_t)(hi - lo - 1);
#pragma unroll
    for (int k = 0; k < NCH; ++k) {
        const int f = f0 + k;
        if (k >= cnt) break;
        T v[8], w[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) w[e] = two_level_sum(acc[k][e], tot[k][e]);
        if (flip) {
            // (V^H)_ij = conj(V_ji): XX, YX, XY, YY with the imaginary parts negated.
            v[0] = w[0];
            v[1] = -w[1];
            v[2] = w[4];
            v[3] = -w[5];
            v[4] = w[2];
            v[5] = -w[3];
            v[6] = w[6];
            v[7] = -w[7];
        } else {
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = w[e];
        }
        const size_t o = (size_t)f * a.nbl + bl;
        float *v32 = reinterpret_cast<float *>(&a.vis32[o]);
#pragma unroll
        for (int e = 0; e < 8; ++e) v32[e] += (float)v[e];
        if (a.vis64 != nullptr) {
            double *v64 = reinterpret_cast<double *>(&a.vis64[o]);
#pragma unroll
            for (int e = 0; e < 8; ++e) v64[e] += (double)v[e];
        }
    }
}

hipError_t launch_direct_rect(Kind kind, const RectArgs &a, hipStream_t stream) {
    if (a.ncomp <= 0 || a.n_row <= 0 || a.n_col <= 0 || a.nchan <= 0) return hipSuccess;
    const int nch = kind == KIND_GAUSSIAN ? kRectNchGaussian : kRectNchPoint;
    const long total = (long)((a.n_row + 3) / 4) * ((a.n_col + 63) / 64) * chunk_count(a.chunks, nch, a.nchan);
    if (a.coh_cls32 != nullptr) {
        if (kind == KIND_GAUSSIAN)
            hipLaunchKernelGGL((k_direct_rect<KIND_GAUSSIAN, kRectNchGaussian, float>), dim3((unsigned)total),
                               dim3(256), 0, stream, a);
        else
            hipLaunchKernelGGL((k_direct_rect<KIND_POINT, kRectNchPoint, float>), dim3((unsigned)total), dim3(256), 0,
                               stream, a);
    } else if (kind == KIND_GAUSSIAN) {
        hipLaunchKernelGGL((k_direct_rect<KIND_GAUSSIAN, kRectNchGaussian, double>), dim3((unsigned)total),
                           dim3(256), 0, stream, a);
    } else {
        hipLaunchKernelGGL((k_direct_rect<KIND_POINT, kRectNchPoint, double>), dim3((unsigned)total), dim3(256), 0,
                           stream, a);
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------ shapelets
// Lane `lane`'s value of v (every lane active), as a wave-uniform value.
__device__ __forceinline__ double readlane_f64(double v, int lane) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), lane);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

template <bool PER_TILE, class T>
__device__ __forceinline__ void shapelet_chan(const ShapeletChanArgs &a) {
    const int nbx = (a.nbl + 3) / 4;
    const int total = nbx * a.ngroups;
    int b = blockIdx.x;
    if ((total & 7) == 0) b = (b & 7) * (total >> 3) + (b >> 3);
    const int group = b / nbx;
    const int bx = b - group * nbx;
    const int lane = threadIdx.x & 63;
    const int bl = bx * 4 + (int)(threadIdx.x >> 6);
    if (bl >= a.nbl) return;  // whole wave; the kernel has no barriers
    const int f_raw = group * 64 + lane;
    const bool valid = f_raw < a.nchan;
    const int f = valid ? f_raw : a.nchan - 1;
    const double freq = a.freqs[f];
    const double fc = freq / SM_VEL_C;
    const UVW uvw = a.uvw[bl];
    int2 rows = make_int2(0, 0);
    int uf = 0;
    int cls = -1;
    bool herm = false;
    if (PER_TILE) {
        rows = a.beam.bl_rows[bl];
        uf = a.beam.freq_map[f];
        if (a.beam.coh_cls != nullptr) {
            if (rows.y == a.beam.row0) {
                cls = a.beam.class_of_row[rows.x];
            } else if (rows.x == a.beam.row0) {
                cls = a.beam.class_of_row[rows.y];
                herm = true;
            }
        }
    }
    const PackedCoeff *__restrict__ packed = a.packed;
    const T *__restrict__ sbf = pick<T>(a.sbf, a.sbf32);
    const T *__restrict__ coh_t = pick<T>(a.coh, a.coh32);
    // Per-wave windows of the basis table (kRows rows x kWin cells per axis).
    __shared__ T wx_lds[4][kRows][kWin], wy_lds[4][kRows][kWin];
    const int wv = threadIdx.x >> 6;
    // The wave's lowest and highest channel frequency (invalid lanes repeat the last channel).
    double fmin_w = freq, fmax_w = freq;
    for (int o = 32; o > 0; o >>= 1) {
        fmin_w = fmin(fmin_w, __shfl_xor(fmin_w, o));
        fmax_w = fmax(fmax_w, __shfl_xor(fmax_w, o));
    }
    T acc[1][8] = {{0, 0, 0, 0, 0, 0, 0, 0}}, tot[1][8] = {{0, 0, 0, 0, 0, 0, 0, 0}};
    // The wave's baseline is uniform; the next component's coefficient range is loaded one
    // iteration ahead.
    const int blu = __builtin_amdgcn_readfirstlane(bl);
    // Shapelet UVs: formed here from the baseline's tile XYZs and each component's (HA, Dec)
    // when those are given, so the [component][baseline] table is not re-read by every channel
    // group; else read from bl_uv. The per-tile kernel reads bl_uv: it has no registers to spare
    // (measured on the 256-tile case: more spills, 104 -> 110 ms per timestep).
    const bool uv_here = !PER_TILE && a.hadec != nullptr;
    double x1 = 0, y1 = 0, z1 = 0, x2 = 0, y2 = 0, z2 = 0;
    if (uv_here) {
        const int2 t = a.bl_tiles[blu];
        x1 = a.xyz[3 * t.x];
        y1 = a.xyz[3 * t.x + 1];
        z1 = a.xyz[3 * t.x + 2];
        x2 = a.xyz[3 * t.y];
        y2 = a.xyz[3 * t.y + 1];
        z2 = a.xyz[3 * t.y + 2];
    }
    const double4 *__restrict__ hd4 = reinterpret_cast<const double4 *>(a.hadec);
    // The phase dot product, the UV in the component's frame (xk, yk) and the table window are
    // the same on every lane of the wave: lane l forms those of component 64 q + l once per
    // batch q of 64 components, and each component reads them from that lane. The window is
    // packed as first cell | width << 16 (both below SM_SBF_L < 2^15).
    double th_lane = 0.0, xk_lane = 0.0, yk_lane = 0.0;
    int wx_lane = 0, wy_lane = 0;
    auto load_batch = [&](int q) {
        const int cl = min((q << 6) + lane, a.ncomp - 1);
        const double4 l = a.lmn[cl];
        const double4 sh = a.shape[cl];
        const double2 suv =
            uv_here ? shapelet_uv(hd4[cl], x1, y1, z1, x2, y2, z2) : a.bl_uv[(size_t)cl * a.nbl + blu];
        th_lane = uvw.u * l.x + uvw.v * l.y + uvw.w * l.z;
        const double xk = (suv.x * sh.x + suv.y * sh.y) * sh.z;
        const double yk = (suv.x * sh.y - suv.y * sh.x) * sh.w;
        // The cells the wave touches: positions are monotone in frequency, so the extreme
        // channels bound them (the same expressions as each lane's position, so the same
        // roundings).
        const double xa = xk * fmin_w + SM_SBF_C, xb = xk * fmax_w + SM_SBF_C;
        const double ya = yk * fmin_w + SM_SBF_C, yb = yk * fmax_w + SM_SBF_C;
        const int ix0 = (int)fmin(fmax(floor(fmin(xa, xb)), 0.0), (double)(SM_SBF_L - 2));
        const int ix1 = (int)fmin(fmax(floor(fmax(xa, xb)), 0.0), (double)(SM_SBF_L - 2));
        const int iy0 = (int)fmin(fmax(floor(fmin(ya, yb)), 0.0), (double)(SM_SBF_L - 2));
        const int iy1 = (int)fmin(fmax(floor(fmax(ya, yb)), 0.0), (double)(SM_SBF_L - 2));
        xk_lane = xk;
        yk_lane = yk;
        wx_lane = ix0 | ((ix1 - ix0 + 2) << 16);
        wy_lane = iy0 | ((iy1 - iy0 + 2) << 16);
    };
    int4 pr_next = a.prange[0];
    for (int c = 0; c < a.ncomp; ++c) {
        const int4 pr4 = pr_next;
        pr_next = a.prange[c + 1 < a.ncomp ? c + 1 : c];
        if ((c & 63) == 0) load_batch(c >> 6);
        const int src = c & 63;
        const double theta = readlane_f64(th_lane, src);
        const double xk = readlane_f64(xk_lane, src), yk = readlane_f64(yk_lane, src);
        const int wxp = __builtin_amdgcn_readlane(wx_lane, src), wyp = __builtin_amdgcn_readlane(wy_lane, src);
        T zr, zi;
        sincos_phase(theta * fc, zi, zr);
        const double x_pos = xk * freq + SM_SBF_C;
        const double y_pos = yk * freq + SM_SBF_C;
        const double fx = floor(x_pos), fy = floor(y_pos);
        const int ix = (int)fmin(fmax(fx, 0.0), (double)(SM_SBF_L - 2));
        const int iy = (int)fmin(fmax(fy, 0.0), (double)(SM_SBF_L - 2));
        const T wx = (T)(x_pos - fx), wy = (T)(y_pos - fy);
        T er, ei;
        const int xw = wxp >> 16, yw = wyp >> 16;
        T cpre[8];
        const T *s_c = coh_t + ((size_t)c * a.nchan_pad + f) * 8;
        if (xw <= kWin && yw <= kWin && pr4.w <= kRows) {
            // Stage rows 0 .. pr4.w - 1 of both windows in this wave's LDS, then interpolate
            // from LDS: one 16-byte read per axis per coefficient instead of two gathers.
            const int bx0 = wxp & 0xffff, by0 = wyp & 0xffff;
            // Lane -> (row lane >> lw, cell lane & (2^lw - 1)), 2^lw the smallest power of two
            // of at least 4 cells covering both windows, so one pass stages 64 >> lw rows of both
            // axes (16 rows for windows of up to 4 cells). All loads first, then all LDS writes:
            // the loads are in flight together instead of one round trip per row. A lane past
            // the window or past row pr4.w - 1 reads an in-bounds cell (the window's last cell;
            // rows below kRows + 16 <= SM_SBF_N) that is never written. Byte offsets are 32-bit
            // (the table is 8 MB), from the uniform base.
            const int wm = max(xw, yw);
            const int lw = wm <= 4 ? 2 : (wm <= 8 ? 3 : (wm <= 16 ? 4 : 5));
            const int rstep = 64 >> lw;
            const int j = lane & ((1 << lw) - 1), rl = lane >> lw;
            const int jxc = min(j, xw - 1), jyc = min(j, yw - 1);
            static_assert(kRows + 16 <= SM_SBF_N, "staged rows must exist in the basis table");
            const unsigned ox = (unsigned)((rl * SM_SBF_L + bx0 + jxc) * sizeof(T));
            const unsigned oy = (unsigned)((rl * SM_SBF_L + by0 + jyc) * sizeof(T));
            const char *sbf_b = reinterpret_cast<const char *>(sbf);
            constexpr int kBatch = 2;  // row steps in flight at once (register budget)
            for (int i0 = 0; i0 < pr4.w; i0 += kBatch * rstep) {
                T tx[kBatch], ty[kBatch];
#pragma unroll
                for (int i = 0; i < kBatch; ++i) {
                    const int base = i0 + i * rstep;
                    const unsigned step = (unsigned)(base * SM_SBF_L * sizeof(T));
                    if (base < pr4.w) {
                        tx[i] = *reinterpret_cast<const T *>(sbf_b + (ox + step));
                        ty[i] = *reinterpret_cast<const T *>(sbf_b + (oy + step));
                    }
                }
#pragma unroll
                for (int i = 0; i < kBatch; ++i) {
                    const int base = i0 + i * rstep, r = base + rl;
                    if (base < pr4.w && r < pr4.w) {
                        if (j < xw) wx_lds[wv][r][j] = tx[i];
                        if (j < yw) wy_lds[wv][r][j] = ty[i];
                    }
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const int jx = ix - bx0, jy = iy - by0;
            if (!PER_TILE) {
#pragma unroll
                for (int e = 0; e < 8; ++e) cpre[e] = s_c[e];
            }
            T part[2];
#pragma unroll
            for (int pt = 0; pt < 2; ++pt) {
                const PackedCoeff *cp = packed + pr4.x + (pt ? pr4.y : 0);
                const int cnt = pt ? pr4.z : pr4.y;
                T accp = T(0);
#pragma unroll 2
                for (int k = 0; k < cnt; ++k) {
                    double val;
                    int n1, n2;
                    load_packed_uniform(cp + k, val, n1, n2);
                    const T xl = wx_lds[wv][n1][jx], xh = wx_lds[wv][n1][jx + 1];
                    const T yl = wy_lds[wv][n2][jy], yh = wy_lds[wv][n2][jy + 1];
                    const T u_value = xl + (xh - xl) * wx;
                    const T v_value = yl + (yh - yl) * wy;
                    accp += ((T)val * u_value) * v_value;
                }
                part[pt] = accp;
            }
            er = part[0];
            ei = part[1];
            // Every lane's reads are done before the next component's staging.
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        } else {
            // Rare (a window wider than the LDS copy): one coefficient at a time, which keeps
            // this path's registers below the staged path's.
            T part[2];
#pragma unroll
            for (int pt = 0; pt < 2; ++pt) {
                const PackedCoeff *cp = packed + pr4.x + (pt ? pr4.y : 0);
                const int cnt = pt ? pr4.z : pr4.y;
                T accp = T(0);
#pragma unroll 1
                for (int k = 0; k < cnt; ++k) {
                    double val;
                    int n1, n2;
                    load_packed_uniform(cp + k, val, n1, n2);
                    const T *bx = sbf + n1 * SM_SBF_L + ix, *by = sbf + n2 * SM_SBF_L + iy;
                    const T u_value = bx[0] + (bx[1] - bx[0]) * wx;
                    const T v_value = by[0] + (by[1] - by[0]) * wy;
                    accp += ((T)val * u_value) * v_value;
                }
                part[pt] = accp;
            }
            er = part[0];
            ei = part[1];
            if (!PER_TILE) {
#pragma unroll
                for (int e = 0; e < 8; ++e) cpre[e] = s_c[e];
            }
        }
        const T pr = zr * er - zi * ei;
        const T pi = zr * ei + zi * er;
        T cj[8];
        if (PER_TILE && cls >= 0) {
            class_coh(pick<T>(a.beam.coh_cls, a.coh_cls32) + (((size_t)cls * a.ncomp + c) * a.nchan_pad + f) * 8,
                      herm, cj);
        } else if (PER_TILE) {
            // J_p S J_q^H of the instrumental flux S (FP64 table) and the two tiles' beams.
            const double *s = a.coh + ((size_t)c * a.nchan_pad + f) * 8;
            double sv[8], j1[8], j2[8], cd[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) sv[e] = s[e];
            load_jones_t(&a.beam.jones[((size_t)rows.x * a.beam.n_ufreq + uf) * a.beam.n_dir + a.beam.dir_offset + c], j1);
            load_jones_t(&a.beam.jones[((size_t)rows.y * a.beam.n_ufreq + uf) * a.beam.n_dir + a.beam.dir_offset + c], j2);
            sandwich_t(j1, sv, j2, cd);
#pragma unroll
            for (int e = 0; e < 8; ++e) cj[e] = (T)cd[e];
        } else {
#pragma unroll
            for (int e = 0; e < 8; ++e) cj[e] = cpre[e];
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            acc[0][2 * e] = fma_t(cj[2 * e], pr, acc[0][2 * e]);
            acc[0][2 * e] = fma_t(-cj[2 * e + 1], pi, acc[0][2 * e]);
            acc[0][2 * e + 1] = fma_t(cj[2 * e], pi, acc[0][2 * e + 1]);
            acc[0][2 * e + 1] = fma_t(cj[2 * e + 1], pr, acc[0][2 * e + 1]);
        }
        if ((c & 15) == 15) flush_stage(acc, tot);
    }
    if (!valid) return;
    const size_t o = (size_t)f * a.nbl + bl;
    float *v32 = reinterpret_cast<float *>(&a.vis32[o]);
#pragma unroll
    for (int e = 0; e < 8; ++e) v32[e] += (float)two_level_sum(acc[0][e], tot[0][e]);
    if (a.vis64 != nullptr) {
        double *v64 = reinterpret_cast<double *>(&a.vis64[o]);
#pragma unroll
        for (int e = 0; e < 8; ++e) v64[e] += (double)two_level_sum(acc[0][e], tot[0][e]);
    }
}

// Shared beam: one wave per baseline and 128 channels, two adjacent channels per lane
// (channels 2 l and 2 l + 1 of the group on lane l). The table cells are read for both channels
// at once: adjacent channels' positions differ by |k| df, below one cell (checked per component),
// so both fall in cells i0 and i0 + 1 with i0 the lower of their two cells. The staged window
// holds each cell's value and its difference to the next cell, formed once per component, and
// the lane reads T[i0], T[i0 + 1] - T[i0] and T[i0 + 2] - T[i0 + 1] per axis and coefficient
// (24 bytes for two channels, against 16 per channel reading two values). A channel in cell
// i0 interpolates
//   u = T[i0] + (T[i0+1] - T[i0]) w,
// as the single-channel form does, and one in cell i0 + 1 takes
//   u = (T[i0] + (T[i0+1] - T[i0])) + (T[i0+2] - T[i0+1]) w,
// equal to T[i0+1] + (T[i0+2] - T[i0+1]) w up to the rounding of T[i0] + (T[i0+1] - T[i0]).
// The coherency table is in the Hermitian form (Re XX, Re XY, Im XY, Re YY; J S J^H with one
// beam on both sides), so each channel accumulates with 8 FMAs per component.

// A staged basis-table cell: the value T[i] and the difference T[i + 1] - T[i].
template <class T> struct __attribute__((aligned(2 * sizeof(T)))) CellTD {
    T t, d;
};

template <class T>
__device__ __forceinline__ void shapelet_pair(const ShapeletChanArgs &a) {
    const int ngroups = (a.nchan + 127) / 128;
    const int nbx = (a.nbl + 3) / 4;
    const int total = nbx * ngroups;
    int b = blockIdx.x;
    if ((total & 7) == 0) b = (b & 7) * (total >> 3) + (b >> 3);
    const int group = b / nbx;
    const int bx = b - group * nbx;
    const int lane = threadIdx.x & 63;
    const int bl = bx * 4 + (int)(threadIdx.x >> 6);
    if (bl >= a.nbl) return;  // whole wave; the kernel has no barriers
    const int fa_raw = group * 128 + 2 * lane;
    const bool valid_a = fa_raw < a.nchan, valid_b = fa_raw + 1 < a.nchan;
    const int fa = valid_a ? fa_raw : a.nchan - 1;
    const int fb = valid_b ? fa_raw + 1 : a.nchan - 1;
    const double freq_a = a.freqs[fa], freq_b = a.freqs[fb];
    const double fc_a = freq_a / SM_VEL_C, fc_b = freq_b / SM_VEL_C;
    const UVW uvw = a.uvw[bl];
    const PackedCoeff *__restrict__ packed = a.packed;
    const T *__restrict__ sbf = pick<T>(a.sbf, a.sbf32);
    const T *__restrict__ coh_t = pick<T>(a.coh, a.coh32);
    // Per-wave windows of the basis table (kRows rows x kWinPair cells per axis), each cell
    // holding T[i] and T[i + 1] - T[i]: 48 KB per block in FP64, so three blocks share a CU.
    __shared__ CellTD<T> wx_lds[4][kRows][kWinPair], wy_lds[4][kRows][kWinPair];
    const int wv = threadIdx.x >> 6;
    // The wave's lowest and highest channel frequency and its largest step between a lane's
    // two channels (channels past the last repeat it).
    double fmin_w = fmin(freq_a, freq_b), fmax_w = fmax(freq_a, freq_b), dfmax_w = fabs(freq_b - freq_a);
    for (int o = 32; o > 0; o >>= 1) {
        fmin_w = fmin(fmin_w, __shfl_xor(fmin_w, o));
        fmax_w = fmax(fmax_w, __shfl_xor(fmax_w, o));
        dfmax_w = fmax(dfmax_w, __shfl_xor(dfmax_w, o));
    }
    // FP64 with the same step between the two channels on every lane (evenly spaced channels):
    // channel b's phasor is channel a's turned by exp(i theta df / c), the turn formed once per
    // component in the batch below, instead of a second phase reduction and sincos.
    const double df_ref = __shfl(freq_b - freq_a, 0);
    const bool rot = std::is_same<T, double>::value && __all(!valid_b || freq_b - freq_a == df_ref);
    const double dfc_rot = df_ref / SM_VEL_C;
    T acc[2][8], tot[2][8];
#pragma unroll
    for (int ch = 0; ch < 2; ++ch)
#pragma unroll
        for (int e = 0; e < 8; ++e) acc[ch][e] = tot[ch][e] = T(0);
    const int blu = __builtin_amdgcn_readfirstlane(bl);
    // Shapelet UVs: formed here from the baseline's tile XYZs and each component's (HA, Dec)
    // when those are given; else read from bl_uv.
    const bool uv_here = a.hadec != nullptr;
    double x1 = 0, y1 = 0, z1 = 0, x2 = 0, y2 = 0, z2 = 0;
    if (uv_here) {
        const int2 t = a.bl_tiles[blu];
        x1 = a.xyz[3 * t.x];
        y1 = a.xyz[3 * t.x + 1];
        z1 = a.xyz[3 * t.x + 2];
        x2 = a.xyz[3 * t.y];
        y2 = a.xyz[3 * t.y + 1];
        z2 = a.xyz[3 * t.y + 2];
    }
    const double4 *__restrict__ hd4 = reinterpret_cast<const double4 *>(a.hadec);
    // As in shapelet_chan: the wave-uniform per-component values, formed by lane l for
    // component 64 q + l and read back by lane.
    double th_lane = 0.0, xk_lane = 0.0, yk_lane = 0.0, rs_lane = 0.0, rc_lane = 1.0;
    int wx_lane = 0, wy_lane = 0;
    auto load_batch = [&](int q) {
        const int cl = min((q << 6) + lane, a.ncomp - 1);
        const double4 l = a.lmn[cl];
        const double4 sh = a.shape[cl];
        const double2 suv =
            uv_here ? shapelet_uv(hd4[cl], x1, y1, z1, x2, y2, z2) : a.bl_uv[(size_t)cl * a.nbl + blu];
        th_lane = uvw.u * l.x + uvw.v * l.y + uvw.w * l.z;
        if (rot) sincos_phase(th_lane * dfc_rot, rs_lane, rc_lane);
        const double xk = (suv.x * sh.x + suv.y * sh.y) * sh.z;
        const double yk = (suv.x * sh.y - suv.y * sh.x) * sh.w;
        const double xa = xk * fmin_w + SM_SBF_C, xb = xk * fmax_w + SM_SBF_C;
        const double ya = yk * fmin_w + SM_SBF_C, yb = yk * fmax_w + SM_SBF_C;
        const int ix0 = (int)fmin(fmax(floor(fmin(xa, xb)), 0.0), (double)(SM_SBF_L - 2));
        const int ix1 = (int)fmin(fmax(floor(fmax(xa, xb)), 0.0), (double)(SM_SBF_L - 2));
        const int iy0 = (int)fmin(fmax(floor(fmin(ya, yb)), 0.0), (double)(SM_SBF_L - 2));
        const int iy1 = (int)fmin(fmax(floor(fmax(ya, yb)), 0.0), (double)(SM_SBF_L - 2));
        xk_lane = xk;
        yk_lane = yk;
        wx_lane = ix0 | ((ix1 - ix0 + 2) << 16);
        wy_lane = iy0 | ((iy1 - iy0 + 2) << 16);
    };
    int4 pr_next = a.prange[0];
    for (int c = 0; c < a.ncomp; ++c) {
        const int4 pr4 = pr_next;
        pr_next = a.prange[c + 1 < a.ncomp ? c + 1 : c];
        if ((c & 63) == 0) load_batch(c >> 6);
        const int src = c & 63;
        const double theta = readlane_f64(th_lane, src);
        const double xk = readlane_f64(xk_lane, src), yk = readlane_f64(yk_lane, src);
        const int wxp = __builtin_amdgcn_readlane(wx_lane, src), wyp = __builtin_amdgcn_readlane(wy_lane, src);
        T zr[2], zi[2];
        sincos_phase(theta * fc_a, zi[0], zr[0]);
        if (rot) {
            const double rs = readlane_f64(rs_lane, src), rc = readlane_f64(rc_lane, src);
            zr[1] = (T)((double)zr[0] * rc - (double)zi[0] * rs);
            zi[1] = (T)((double)zr[0] * rs + (double)zi[0] * rc);
        } else {
            sincos_phase(theta * fc_b, zi[1], zr[1]);
        }
        const double xpa = xk * freq_a + SM_SBF_C, xpb = xk * freq_b + SM_SBF_C;
        const double ypa = yk * freq_a + SM_SBF_C, ypb = yk * freq_b + SM_SBF_C;
        const double fxa = floor(xpa), fxb = floor(xpb), fya = floor(ypa), fyb = floor(ypb);
        const int ixa = (int)fmin(fmax(fxa, 0.0), (double)(SM_SBF_L - 2));
        const int ixb = (int)fmin(fmax(fxb, 0.0), (double)(SM_SBF_L - 2));
        const int iya = (int)fmin(fmax(fya, 0.0), (double)(SM_SBF_L - 2));
        const int iyb = (int)fmin(fmax(fyb, 0.0), (double)(SM_SBF_L - 2));
        const T wxa = (T)(xpa - fxa), wxb = (T)(xpb - fxb), wya = (T)(ypa - fya), wyb = (T)(ypb - fyb);
        const int xw = wxp >> 16, yw = wyp >> 16;
        // One component's coherencies for both channels (Hermitian form, 4 values each).
        T cq[2][4];
        {
            const T *s_a = coh_t + ((size_t)c * a.nchan_pad + fa) * 4;
            const T *s_b = coh_t + ((size_t)c * a.nchan_pad + fb) * 4;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                cq[0][e] = s_a[e];
                cq[1][e] = s_b[e];
            }
        }
        T er[2], ei[2];
        // Staged path: both windows, one extra cell for T[i0 + 2], fit the LDS copy, and the
        // two channels of a lane are less than half a cell apart on both axes (so their cells
        // differ by at most one, with a wide margin for rounding).
        const bool near = fabs(xk) * dfmax_w < 0.5 && fabs(yk) * dfmax_w < 0.5;
        if (near && xw + 1 <= kWinPair && yw + 1 <= kWinPair && pr4.w <= kRows) {
            const int bx0 = wxp & 0xffff, by0 = wyp & 0xffff;
            const int xs = xw + 1, ys = yw + 1;  // staged cells per axis
            // As in shapelet_chan: lane -> (row lane >> lw, cell lane & (2^lw - 1)). A lane past
            // the window or past row pr4.w - 1 reads an in-bounds cell that is never written; the
            // last cell is at most SM_SBF_L, in the row or at the next row's first cell, and the
            // difference it enters is only multiplied by a zero weight when outside the row.
            // T[i + 1] comes from the next lane of the row; the difference formed with the next
            // row's lane at cell 2^lw - 1 >= xs - 1 is never read (the loop reads differences
            // at cells up to xs - 2).
            const int wm = max(xs, ys);
            const int lw = wm <= 4 ? 2 : (wm <= 8 ? 3 : (wm <= 16 ? 4 : 5));
            const int rstep = 64 >> lw;
            const int j = lane & ((1 << lw) - 1), rl = lane >> lw;
            const int jxc = min(j, xs - 1), jyc = min(j, ys - 1);
            static_assert(kRows + 16 < SM_SBF_N, "staged rows and the next row must exist in the basis table");
            const unsigned ox = (unsigned)((rl * SM_SBF_L + bx0 + jxc) * sizeof(T));
            const unsigned oy = (unsigned)((rl * SM_SBF_L + by0 + jyc) * sizeof(T));
            const char *sbf_b = reinterpret_cast<const char *>(sbf);
            constexpr int kBatch = 2;
            for (int i0 = 0; i0 < pr4.w; i0 += kBatch * rstep) {
                T tx[kBatch], ty[kBatch];
#pragma unroll
                for (int i = 0; i < kBatch; ++i) {
                    const int base = i0 + i * rstep;
                    const unsigned step = (unsigned)(base * SM_SBF_L * sizeof(T));
                    if (base < pr4.w) {
                        tx[i] = *reinterpret_cast<const T *>(sbf_b + (ox + step));
                        ty[i] = *reinterpret_cast<const T *>(sbf_b + (oy + step));
                    }
                }
#pragma unroll
                for (int i = 0; i < kBatch; ++i) {
                    const int base = i0 + i * rstep, r = base + rl;
                    if (base < pr4.w) {  // wave-uniform: every lane takes part in the shuffles
                        const T tx1 = __shfl_down(tx[i], 1), ty1 = __shfl_down(ty[i], 1);
                        if (r < pr4.w) {
                            if (j < xs) wx_lds[wv][r][j] = CellTD<T>{tx[i], tx1 - tx[i]};
                            if (j < ys) wy_lds[wv][r][j] = CellTD<T>{ty[i], ty1 - ty[i]};
                        }
                    }
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // Each channel's weights on (T[i0+1] - T[i0]) and (T[i0+2] - T[i0+1]).
            const int ix0l = min(ixa, ixb), iy0l = min(iya, iyb);
            const int jx = ix0l - bx0, jy = iy0l - by0;
            const bool xa_lo = ixa == ix0l, xb_lo = ixb == ix0l, ya_lo = iya == iy0l, yb_lo = iyb == iy0l;
            const T ax_a = xa_lo ? wxa : T(1), bx_a = xa_lo ? T(0) : wxa;
            const T ax_b = xb_lo ? wxb : T(1), bx_b = xb_lo ? T(0) : wxb;
            const T ay_a = ya_lo ? wya : T(1), by_a = ya_lo ? T(0) : wya;
            const T ay_b = yb_lo ? wyb : T(1), by_b = yb_lo ? T(0) : wyb;
#pragma unroll
            for (int pt = 0; pt < 2; ++pt) {
                const PackedCoeff *cp = packed + pr4.x + (pt ? pr4.y : 0);
                const int cnt = pt ? pr4.z : pr4.y;
                T acc_a = T(0), acc_b = T(0);
#pragma unroll 4
                for (int k = 0; k < cnt; ++k) {
                    double val;
                    int n1, n2;
                    load_packed_uniform(cp + k, val, n1, n2);
                    const CellTD<T> cx = wx_lds[wv][n1][jx], cy = wy_lds[wv][n2][jy];
                    const T dx12 = wx_lds[wv][n1][jx + 1].d, dy12 = wy_lds[wv][n2][jy + 1].d;
                    const T u_a = fma_t(dx12, bx_a, fma_t(cx.d, ax_a, cx.t));
                    const T u_b = fma_t(dx12, bx_b, fma_t(cx.d, ax_b, cx.t));
                    const T v_a = fma_t(dy12, by_a, fma_t(cy.d, ay_a, cy.t));
                    const T v_b = fma_t(dy12, by_b, fma_t(cy.d, ay_b, cy.t));
                    acc_a += ((T)val * u_a) * v_a;
                    acc_b += ((T)val * u_b) * v_b;
                }
                (pt ? ei : er)[0] = acc_a;
                (pt ? ei : er)[1] = acc_b;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        } else {
            // Rare: one coefficient at a time from the table in memory, per channel.
            const int ixs[2] = {ixa, ixb}, iys[2] = {iya, iyb};
            const T wxs[2] = {wxa, wxb}, wys[2] = {wya, wyb};
#pragma unroll
            for (int ch = 0; ch < 2; ++ch) {
#pragma unroll
                for (int pt = 0; pt < 2; ++pt) {
                    const PackedCoeff *cp = packed + pr4.x + (pt ? pr4.y : 0);
                    const int cnt = pt ? pr4.z : pr4.y;
                    T accp = T(0);
#pragma unroll 1
                    for (int k = 0; k < cnt; ++k) {
                        double val;
                        int n1, n2;
                        load_packed_uniform(cp + k, val, n1, n2);
                        const T *bxp = sbf + n1 * SM_SBF_L + ixs[ch], *byp = sbf + n2 * SM_SBF_L + iys[ch];
                        const T u_value = bxp[0] + (bxp[1] - bxp[0]) * wxs[ch];
                        const T v_value = byp[0] + (byp[1] - byp[0]) * wys[ch];
                        accp += ((T)val * u_value) * v_value;
                    }
                    (pt ? ei : er)[ch] = accp;
                }
            }
        }
#pragma unroll
        for (int ch = 0; ch < 2; ++ch) {
            const T pr = zr[ch] * er[ch] - zi[ch] * ei[ch];
            const T pi = zr[ch] * ei[ch] + zi[ch] * er[ch];
            acc[ch][0] = fma_t(cq[ch][0], pr, acc[ch][0]);
            acc[ch][1] = fma_t(cq[ch][0], pi, acc[ch][1]);
            acc[ch][2] = fma_t(cq[ch][1], pr, acc[ch][2]);
            acc[ch][3] = fma_t(cq[ch][1], pi, acc[ch][3]);
            acc[ch][4] = fma_t(cq[ch][2], pr, acc[ch][4]);
            acc[ch][5] = fma_t(cq[ch][2], pi, acc[ch][5]);
            acc[ch][6] = fma_t(cq[ch][3], pr, acc[ch][6]);
            acc[ch][7] = fma_t(cq[ch][3], pi, acc[ch][7]);
        }
        if ((c & 15) == 15) flush_stage(acc, tot);
    }
#pragma unroll
    for (int ch = 0; ch < 2; ++ch) {
        if (!(ch ? valid_b : valid_a)) continue;
        const size_t o = (size_t)(ch ? fb : fa) * a.nbl + bl;
        T s8[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) s8[e] = two_level_sum(acc[ch][e], tot[ch][e]);
        // XX, XY = A + i B, YX = A - i B, YY.
        const T v[8] = {s8[0], s8[1], s8[2] - s8[5], s8[3] + s8[4], s8[2] + s8[5], s8[3] - s8[4], s8[6], s8[7]};
        float *v32 = reinterpret_cast<float *>(&a.vis32[o]);
#pragma unroll
        for (int e = 0; e < 8; ++e) v32[e] += (float)v[e];
        if (a.vis64 != nullptr) {
            double *v64 = reinterpret_cast<double *>(&a.vis64[o]);
#pragma unroll
            for (int e = 0; e < 8; ++e) v64[e] += (double)v[e];
        }
    }
}

// Shared beam.
__global__ __launch_bounds__(256) void k_shapelet_chan(const ShapeletChanArgs a) { shapelet_pair<double>(a); }
__global__ __launch_bounds__(256) void k_shapelet_chan32(const ShapeletChanArgs a) { shapelet_pair<float>(a); }

// Per-tile beams: held to 128 VGPRs (4 waves per SIMD, a few spilled values) instead of the
// 164 the compiler picks (3 waves). Measured on the 256-tile, 384-channel case: 129 -> 111 ms
// per timestep.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_shapelet_chan_tiles(
    const ShapeletChanArgs a) {
    shapelet_chan<true, double>(a);
}
// FP32 per-tile beams: held to 96 VGPRs (5 waves per SIMD, a few spilled values) instead of 118
// (4 waves). Measured on C5: 66.2 -> 57.8 ms per timestep.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_shapelet_chan_tiles32(
    const ShapeletChanArgs a) {
    shapelet_chan<true, float>(a);
}

hipError_t launch_shapelet_chan(bool per_tile, const ShapeletChanArgs &a, hipStream_t stream) {
    if (a.ncomp <= 0 || a.nbl <= 0 || a.nchan <= 0) return hipSuccess;
    // Per-tile beams: a.ngroups groups of 64 channels; shared beam: groups of 128 channels.
    const int total = ((a.nbl + 3) / 4) * (per_tile ? a.ngroups : (a.nchan + 127) / 128);
    const bool f32 = a.coh32 != nullptr;
    if (f32 && (a.sbf32 == nullptr || (per_tile && a.beam.coh_cls != nullptr && a.coh_cls32 == nullptr)))
        return hipErrorInvalidValue;
    if (per_tile)
        hipLaunchKernelGGL(f32 ? k_shapelet_chan_tiles32 : k_shapelet_chan_tiles, dim3(total), dim3(256), 0, stream, a);
    else
        hipLaunchKernelGGL(f32 ? k_shapelet_chan32 : k_shapelet_chan, dim3(total), dim3(256), 0, stream, a);
    return hipGetLastError();
}

__global__ void k_to_f32(size_t n, const double *__restrict__ in, float *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

hipError_t launch_to_f32(size_t n, const double *d_in, float *d_out, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, n, d_in, d_out);
    return hipGetLastError();
}

// Tile UVWs from baseline UVWs: r_0 = 0, r_q = -uvw(0, q), so r_p - r_q = uvw(p, q) for
// UVWs that are differences of tile UVWs, as xyzs_to_cross_uvws produces.
__global__ void k_tile_uvw_from_baselines(int ntiles, const UVW *__restrict__ uvw, double *__restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    if (t == 0) {
        out[0] = out[1] = out[2] = 0.0;
        return;
    }
    const UVW b = uvw[t - 1];  // baseline (0, t)
    out[3 * t] = -b.u;
    out[3 * t + 1] = -b.v;
    out[3 * t + 2] = -b.w;
}

hipError_t launch_tile_uvw_from_baselines(int ntiles, const UVW *d_uvw, double *d_out, hipStream_t stream) {
    hipLaunchKernelGGL(k_tile_uvw_from_baselines, dim3((ntiles + 255) / 256), dim3(256), 0, stream, ntiles, d_uvw,
                       d_out);
    return hipGetLastError();
}

}  // namespace sm
