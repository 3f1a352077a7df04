// Split-precision MFMA path for point components (no counterpart in the reference).
//
// With one beam shared by every tile, V^e_pq(f) = sum_c K_pc C^e_c K*_qc for each Jones
// entry e, where K_pc = e^{i theta_pc f / c} (theta_pc = uvw_p . LMN_c, tile uvw) and
// C_c = J S_c J^H. Per channel that is a complex GEMM (T x Nc) . (Nc x T); in real form
//   [V_re | V_im] = [A_re A_im] . [[K_re, -K_im], [K_im, K_re]],  A = K (.) C^e,
// i.e. M = 4 Jones entries x T tiles, N = 2 x T, K = 2 Nc.
//
// A work-group (8 waves) owns one channel and one 64 x 64 block of tile pairs (p-block <=
// q-block; blocks below the diagonal are never computed). Per step of 16 components it
//   * evaluates the phases and their sincos in FP64 (theta f / c is reduced in FP64),
//   * forms A = K C^e s (s a power-of-two scale keeping |A| inside the fp16 range) and
//     the K operand, and splits each value x into hi = rn(x), lo = rn(x - hi) in the MFMA
//     input type (fp16 or bf16), written to LDS,
//   * runs hi*hi, hi*lo, lo*hi and lo*lo on v_mfma_f32_32x32x16_{f16,bf16} (2 x 2 output
//     tiles of 32 x 32 per wave), the hi*hi products into one FP32 accumulator and the three
//     correction products into a second, so the corrections' roundings stay small,
// and every `seg_steps` steps adds the FP32 partials into an FP64 buffer, so each FP32
// accumulation spans at most 512 components. A scatter kernel maps the FP64 block
// results to baselines and accumulates the visibilities.
#include <cmath>
#include <vector>

#include "sm_factored.h"
#include "sm_math.h"

namespace sm {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

constexpr int kFB = 64;             // tiles per block side
constexpr int kFC = 16;             // components per step
constexpr int kFK = 2 * kFC;        // real K per step
constexpr int kStride = kFK + 8;    // LDS row stride in elements (16-byte aligned, staggered banks)

template <class E> struct Mfma;
template <> struct Mfma<_Float16> {
    typedef h8 vec;
    static __device__ __forceinline__ f16v run(vec a, vec b, f16v c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
};
template <> struct Mfma<__bf16> {
    typedef b8 vec;
    static __device__ __forceinline__ f16v run(vec a, vec b, f16v c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
};

template <class E> __device__ __forceinline__ void split(float x, E &hi, E &lo) {
    hi = (E)x;
    lo = (E)(x - (float)hi);
}

template <class E> struct Vec2;
template <> struct Vec2<_Float16> { typedef _Float16 type __attribute__((ext_vector_type(2))); };
template <> struct Vec2<__bf16> { typedef __bf16 type __attribute__((ext_vector_type(2))); };

// Eight waves: wave w computes Jones entry w % 4 for output columns (w / 4) * 64 .. + 63,
// i.e. 2 x 2 tiles of 32 x 32. Generation: thread t handles tile t / 8 and components
// 2 (t % 8), 2 (t % 8) + 1 of the step, writing the (re, im) pair of each per operand row.
template <class E>
__global__ __launch_bounds__(512) void k_factored(const FactoredArgs a) {
    typedef typename Mfma<E>::vec V;
    typedef typename Vec2<E>::type V2;
    __shared__ __attribute__((aligned(16))) E s_ahi[4][kFB][kStride];
    __shared__ __attribute__((aligned(16))) E s_alo[4][kFB][kStride];
    __shared__ __attribute__((aligned(16))) E s_bhi[2 * kFB][kStride];
    __shared__ __attribute__((aligned(16))) E s_blo[2 * kFB][kStride];
    __shared__ double s_uvw[2 * kFB][3];
    // The step's component data, staged once: 4 LMN doubles and 8 coherency doubles each.
    __shared__ double s_cd[kFC][12];

    const int total = a.nchan * a.nblocks;
    int b = blockIdx.x;
    if ((total & 7) == 0) b = (b & 7) * (total >> 3) + (b >> 3);
    const int f = b / a.nblocks;
    const int blk = b - f * a.nblocks;
    const int2 pq = a.blocks[blk];
    const bool diag = pq.x == pq.y;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int pol = wave & 3;
    const int colbase = (wave >> 2) * kFB;

    if (tid < 2 * kFB) {
        const int g = tid < kFB ? pq.x : pq.y;
        int t = g * kFB + (tid & (kFB - 1));
        if (t >= a.ntiles) t = 0;
        s_uvw[tid][0] = a.tile_uvw[3 * t];
        s_uvw[tid][1] = a.tile_uvw[3 * t + 1];
        s_uvw[tid][2] = a.tile_uvw[3 * t + 2];
    }
    const double fc = a.freqs[f] / SM_VEL_C;
    const double scale = *a.scale;
    const int gp = tid >> 3;          // tile handled in generation
    const int gc = (tid & 7) * 2;     // first of this thread's 2 components in the step

    // acc: hi*hi products; cor: the hi*lo + lo*hi corrections, about 2^-11 smaller, kept
    // apart so their FP32 roundings do not land on the large partial sums.
    f16v acc[2][2], cor[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = cor[i][j][r] = 0.0f;

    // Thread t < 192 fetches element t % 12 of component t / 12 of the next step while the
    // current step's MFMAs run, and stores it to s_cd once they are issued.
    auto fetch = [&](int step) -> double {
        if (tid >= kFC * 12) return 0.0;
        const int c = step * kFC + tid / 12;
        const int e = tid % 12;
        if (c >= a.ncomp) return 0.0;
        if (e < 4) {
            const double4 l = a.lmn[c];
            return e == 0 ? l.x : (e == 1 ? l.y : (e == 2 ? l.z : 0.0));
        }
        return a.coh[((size_t)c * a.nchan_pad + f) * 8 + (e - 4)] * scale;
    };
    const int nsteps = (a.ncomp + kFC - 1) / kFC;
    if (tid < kFC * 12) s_cd[tid / 12][tid % 12] = fetch(0);
    for (int step = 0; step < nsteps; ++step) {
        const int c0 = step * kFC;
        __syncthreads();
#pragma unroll 1
        for (int j = 0; j < 2; ++j) {
            const int cl = gc + j;
            const int c = c0 + cl;
            double kpr = 0.0, kpi = 0.0, kqr = 0.0, kqi = 0.0;
            const double *sd = s_cd[cl];
            if (c < a.ncomp) {
                const double thp = s_uvw[gp][0] * sd[0] + s_uvw[gp][1] * sd[1] + s_uvw[gp][2] * sd[2];
                fast_sincos(thp * fc, kpi, kpr);
                if (diag) {
                    kqr = kpr;
                    kqi = kpi;
                } else {
                    const double thq = s_uvw[kFB + gp][0] * sd[0] + s_uvw[kFB + gp][1] * sd[1] +
                                       s_uvw[kFB + gp][2] * sd[2];
                    fast_sincos(thq * fc, kqi, kqr);
                }
            }
            const float pr = (float)kpr, pi = (float)kpi;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float cr = c < a.ncomp ? (float)sd[4 + 2 * e] : 0.0f;
                const float ci = c < a.ncomp ? (float)sd[5 + 2 * e] : 0.0f;
                E h0, l0, h1, l1;
                split<E>(pr * cr - pi * ci, h0, l0);
                split<E>(pr * ci + pi * cr, h1, l1);
                const V2 h = {h0, h1}, l2 = {l0, l1};
                *reinterpret_cast<V2 *>(&s_ahi[e][gp][2 * cl]) = h;
                *reinterpret_cast<V2 *>(&s_alo[e][gp][2 * cl]) = l2;
            }
            // Column gp carries K_q: rows (2c, 2c+1) hold (K_re, K_im) for V_re and
            // (-K_im, K_re) for V_im.
            const float qr = (float)kqr, qi = (float)kqi;
            E qrh, qrl, qih, qil, nqih, nqil;
            split<E>(qr, qrh, qrl);
            split<E>(qi, qih, qil);
            split<E>(-qi, nqih, nqil);
            const V2 h0 = {qrh, qih}, l0 = {qrl, qil}, h1 = {nqih, qrh}, l1 = {nqil, qrl};
            *reinterpret_cast<V2 *>(&s_bhi[gp][2 * cl]) = h0;
            *reinterpret_cast<V2 *>(&s_blo[gp][2 * cl]) = l0;
            *reinterpret_cast<V2 *>(&s_bhi[kFB + gp][2 * cl]) = h1;
            *reinterpret_cast<V2 *>(&s_blo[kFB + gp][2 * cl]) = l1;
        }
        __syncthreads();
        const double next = step + 1 < nsteps ? fetch(step + 1) : 0.0;
        const int row = lane & 31;
        const int kofs = (lane >> 5) * 8;
#pragma unroll 1
        for (int kb = 0; kb < kFK / 16; ++kb) {
            V ahi[2], alo[2], bhi[2], blo[2];
#pragma unroll
            for (int t2 = 0; t2 < 2; ++t2) {
                ahi[t2] = *reinterpret_cast<const V *>(&s_ahi[pol][t2 * 32 + row][kb * 16 + kofs]);
                alo[t2] = *reinterpret_cast<const V *>(&s_alo[pol][t2 * 32 + row][kb * 16 + kofs]);
                bhi[t2] = *reinterpret_cast<const V *>(&s_bhi[colbase + t2 * 32 + row][kb * 16 + kofs]);
                blo[t2] = *reinterpret_cast<const V *>(&s_blo[colbase + t2 * 32 + row][kb * 16 + kofs]);
            }
#pragma unroll
            for (int rt = 0; rt < 2; ++rt)
#pragma unroll
                for (int ct = 0; ct < 2; ++ct) {
                    acc[rt][ct] = Mfma<E>::run(ahi[rt], bhi[ct], acc[rt][ct]);
                    cor[rt][ct] = Mfma<E>::run(ahi[rt], blo[ct], cor[rt][ct]);
                    cor[rt][ct] = Mfma<E>::run(alo[rt], bhi[ct], cor[rt][ct]);
                    cor[rt][ct] = Mfma<E>::run(alo[rt], blo[ct], cor[rt][ct]);
                }
        }
        if (tid < kFC * 12) s_cd[tid / 12][tid % 12] = next;
        if ((step + 1) % a.seg_steps == 0 || step + 1 == nsteps) {
            // Output layout of v_mfma_f32_32x32x16: element r of lane l is row
            // 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32.
            double *out = a.out + ((size_t)(f * a.nblocks + blk) * 4 + pol) * (kFB * 2 * kFB) +
                          (4 * (lane >> 5)) * (2 * kFB) + colbase + (lane & 31);
#pragma unroll
            for (int rt = 0; rt < 2; ++rt)
#pragma unroll
                for (int ct = 0; ct < 2; ++ct)
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const int o = (rt * 32 + 8 * (r >> 2) + (r & 3)) * (2 * kFB) + ct * 32;
                        out[o] += (double)acc[rt][ct][r] + (double)cor[rt][ct][r];
                        acc[rt][ct][r] = 0.0f;
                        cor[rt][ct][r] = 0.0f;
                        if ((r & 3) == 3) __builtin_amdgcn_sched_barrier(0);
                    }
        }
    }
}

// FP64 block results -> visibilities [freq][baseline] (+=), undoing the scale.
__global__ void k_factored_scatter(const FactoredArgs a, const int2 *__restrict__ bl_tiles, int nbl,
                                   const int *__restrict__ block_of, JonesF32 *vis32, JonesF64 *vis64) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)a.nchan * nbl) return;
    const int f = (int)(idx / nbl);
    const int bl = (int)(idx % nbl);
    const int2 t = bl_tiles[bl];
    const int gb = (a.ntiles + kFB - 1) / kFB;
    const int blk = block_of[(t.x / kFB) * gb + (t.y / kFB)];
    const int pl = t.x % kFB, ql = t.y % kFB;
    const double inv = 1.0 / *a.scale;
    double v[8];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const double *o = a.out + ((size_t)(f * a.nblocks + blk) * 4 + e) * (kFB * 2 * kFB) + pl * (2 * kFB);
        v[2 * e] = o[ql] * inv;
        v[2 * e + 1] = o[kFB + ql] * inv;
    }
    float *v32 = reinterpret_cast<float *>(&vis32[idx]);
#pragma unroll
    for (int e = 0; e < 8; ++e) v32[e] += (float)v[e];
    if (vis64 != nullptr) {
        double *v64 = reinterpret_cast<double *>(&vis64[idx]);
#pragma unroll
        for (int e = 0; e < 8; ++e) v64[e] += v[e];
    }
}

// max |coh| over the table (positive doubles order like their bit patterns).
__global__ void k_absmax(size_t n, const double *__restrict__ x, unsigned long long *out) {
    double m = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        m = fmax(m, fabs(x[i]));
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

// scale = 2^k with max|A| * scale <= 2^14 (fp16 keeps headroom for the K products).
__global__ void k_scale_from_max(const unsigned long long *bits, double *scale) {
    const double m = __longlong_as_double((long long)*bits);
    double s = 1.0;
    if (m > 0.0 && isfinite(m)) s = exp2(floor(log2(16384.0 / m)));
    *scale = s;
}

__global__ void k_fill(size_t n, double *x) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) x[i] = 0.0;
}

__global__ void k_zero_u64(unsigned long long *x) { *x = 0ull; }

std::vector<int2> factored_blocks(int ntiles, std::vector<int> &block_of) {
    const int G = (ntiles + kFB - 1) / kFB;
    std::vector<int2> out;
    block_of.assign((size_t)G * G, 0);
    for (int p = 0; p < G; ++p)
        for (int q = p; q < G; ++q) {
            block_of[(size_t)p * G + q] = (int)out.size();
            out.push_back(make_int2(p, q));
        }
    return out;
}

size_t factored_out_doubles(int nchan, int nblocks) { return (size_t)nchan * nblocks * 4 * kFB * 2 * kFB; }

hipError_t launch_factored(int dtype, FactoredArgs a, const double *coh_for_scale, size_t coh_n,
                           unsigned long long *d_max_bits, double *d_scale, const int2 *bl_tiles, int nbl,
                           const int *block_of, JonesF32 *vis32, JonesF64 *vis64, hipStream_t stream,
                           hipEvent_t ev0, hipEvent_t ev1) {
    if (a.ncomp <= 0 || a.nchan <= 0 || nbl <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, stream, d_max_bits);
    hipLaunchKernelGGL(k_absmax, dim3(1024), dim3(256), 0, stream, coh_n, coh_for_scale, d_max_bits);
    hipLaunchKernelGGL(k_scale_from_max, dim3(1), dim3(1), 0, stream, d_max_bits, d_scale);
    a.scale = d_scale;
    const size_t nout = factored_out_doubles(a.nchan, a.nblocks);
    hipLaunchKernelGGL(k_fill, dim3(2048), dim3(256), 0, stream, nout, a.out);
    if (ev0 != nullptr) (void)hipEventRecord(ev0, stream);
    const int total = a.nchan * a.nblocks;
    if (dtype == SM_MFMA_BF16)
        hipLaunchKernelGGL(k_factored<__bf16>, dim3(total), dim3(512), 0, stream, a);
    else
        hipLaunchKernelGGL(k_factored<_Float16>, dim3(total), dim3(512), 0, stream, a);
    if (ev1 != nullptr) (void)hipEventRecord(ev1, stream);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const size_t n = (size_t)a.nchan * nbl;
    hipLaunchKernelGGL(k_factored_scatter, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, a, bl_tiles, nbl,
                       block_of, vis32, vis64);
    return hipGetLastError();
}

// Single-wave check of the MFMA operand and result layouts assumed above:
// D (32 x 32, f32, row-major) = A (32 x 16, f16, row-major) . B (16 x 32, f16, row-major).
__global__ void k_mfma_layout_probe(const _Float16 *A, const _Float16 *B, float *D) {
    const int lane = threadIdx.x;
    h8 av, bv;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        av[j] = A[(lane & 31) * 16 + (lane >> 5) * 8 + j];
        bv[j] = B[((lane >> 5) * 8 + j) * 32 + (lane & 31)];
    }
    f16v c;
#pragma unroll
    for (int r = 0; r < 16; ++r) c[r] = 0.0f;
    c = __builtin_amdgcn_mfma_f32_32x32x16_f16(av, bv, c, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 16; ++r) D[(8 * (r >> 2) + 4 * (lane >> 5) + (r & 3)) * 32 + (lane & 31)] = c[r];
}

}  // namespace sm

extern "C" const char *sm_selftest_mfma_layout(const void *d_a, const void *d_b, float *d_d) {
    if (d_a == nullptr || d_b == nullptr || d_d == nullptr) return "null argument";
    hipLaunchKernelGGL(sm::k_mfma_layout_probe, dim3(1), dim3(64), 0, 0, (const _Float16 *)d_a,
                       (const _Float16 *)d_b, d_d);
    SM_TRY(hipGetLastError());
    SM_TRY(hipStreamSynchronize(0));
    return nullptr;
}
