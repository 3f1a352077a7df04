// MLP surrogate of the beam Jones (no counterpart in the reference; it stands in for
// BeamGpu::calc_jones_pair, src/beam/mod.rs:143-177).
//
// Network: 53 inputs -> 4 hidden layers of 256 with SiLU -> 8 outputs (re, im of J00, J01,
// J10, J11); other widths and depths train and evaluate in FP32 (sm_surrogate_new_shaped). Inputs: sin az, cos az, sin za, cos za, (f - 182 MHz) / 15 MHz, the 16 dipole
// delays / 31 and the 32 dipole gains.
//
// Training runs on the device with the kernels below: GEMMs on MFMA in FP32
// (v_mfma_f32_32x32x2_f32, products exact in FP32), SiLU forward / backward, mean-squared
// error, column sums for the bias gradients and Adam. Inference runs the same GEMM kernel
// on FP16 MFMA (v_mfma_f32_32x32x16_f16) with FP32 accumulation and SiLU fused into the
// epilogue. The exact beam is the fallback: every unique tile is probed against it, and a
// tile whose surrogate error exceeds the caller's threshold takes the exact beam.
#include <cmath>
#include <cstring>
#include <vector>

#include "sm_internal.h"

namespace sm {

namespace {

constexpr int kNIn = 53;
constexpr int kHid = 256;
constexpr int kNOut = 8;
constexpr int kLayers = 5;
constexpr int kDimIn[kLayers] = {kNIn, kHid, kHid, kHid, kHid};
constexpr int kDimOut[kLayers] = {kHid, kHid, kHid, kHid, kNOut};
constexpr int kMaxLayers = 9;  // up to 8 hidden layers in the FP32 path

typedef float f16v __attribute__((ext_vector_type(16)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));

// ------------------------------------------------------------------ GEMM
// C (M x N, row-major, ldc) = op(A) op(B), op(A)(m, k) = TA ? A[k lda + m] : A[m lda + k],
// op(B)(k, n) = TB ? B[n ldb + k] : B[k ldb + n]; T = float (FP32 MFMA) or _Float16.
// A 256-thread block computes a 64 x 64 tile, each wave 32 x 32, staging 32-deep K slices
// of both operands in LDS. Out-of-range rows, columns and K are read as 0.
enum Epilogue { EPI_STORE = 0, EPI_ACCUMULATE = 1, EPI_BIAS_SILU = 2, EPI_BIAS = 3, EPI_BIAS_SILU_HALF = 4 };

constexpr int kGT = 64;   // tile edge
constexpr int kGK = 32;   // K slice

template <class T> struct GemmOp;
template <> struct GemmOp<float> {
    // One v_mfma_f32_32x32x2_f32 per 2 K: lane l holds A[l % 32][l / 32], B[l / 32][l % 32].
    static __device__ __forceinline__ f16v slice(const float (*sa)[kGK + 1], const float (*sb)[kGK + 1], int ar,
                                                 int bc, int lane, f16v c) {
#pragma unroll
        for (int k = 0; k < kGK; k += 2) {
            const float a = sa[ar + (lane & 31)][k + (lane >> 5)];
            const float b = sb[bc + (lane & 31)][k + (lane >> 5)];
            c = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
        }
        return c;
    }
};
template <> struct GemmOp<_Float16> {
    // One v_mfma_f32_32x32x16_f16 per 16 K: lane l holds A[l % 32][8 (l / 32) + j].
    static __device__ __forceinline__ f16v slice(const _Float16 (*sa)[kGK + 8], const _Float16 (*sb)[kGK + 8], int ar,
                                                 int bc, int lane, f16v c) {
#pragma unroll
        for (int k = 0; k < kGK; k += 16) {
            const h8 a = *reinterpret_cast<const h8 *>(&sa[ar + (lane & 31)][k + 8 * (lane >> 5)]);
            const h8 b = *reinterpret_cast<const h8 *>(&sb[bc + (lane & 31)][k + 8 * (lane >> 5)]);
            c = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
        }
        return c;
    }
};
template <class T> struct LdsRow { static constexpr int pad = 1; };
template <> struct LdsRow<_Float16> { static constexpr int pad = 8; };

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }

template <class T, class S, bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const S *__restrict__ A, int lda,
                                              const S *__restrict__ B, int ldb, float *C, int ldc, int epi,
                                              const float *__restrict__ bias, void *H) {
    constexpr int P = LdsRow<T>::pad;
    __shared__ __attribute__((aligned(16))) T sa[kGT][kGK + P];
    __shared__ __attribute__((aligned(16))) T sb[kGT][kGK + P];
    const int m0 = blockIdx.y * kGT, n0 = blockIdx.x * kGT;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ar = (wave >> 1) * 32, bc = (wave & 1) * 32;
    f16v acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    for (int k0 = 0; k0 < K; k0 += kGK) {
        // 64 x 32 elements per operand, 8 per thread, the global-contiguous index fastest.
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int e = tid + 256 * i;
            int m, k;
            if (TA) {
                k = e >> 6;
                m = e & 63;
            } else {
                m = e >> 5;
                k = e & 31;
            }
            const int gm = m0 + m, gk = k0 + k;
            float v = 0.0f;
            if (gm < M && gk < K) v = (float)(TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk]);
            sa[m][k] = (T)v;
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int e = tid + 256 * i;
            int n, k;
            if (TB) {
                n = e >> 5;
                k = e & 31;
            } else {
                k = e >> 6;
                n = e & 63;
            }
            const int gn = n0 + n, gk = k0 + k;
            float v = 0.0f;
            if (gn < N && gk < K) v = (float)(TB ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn]);
            sb[n][k] = (T)v;
        }
        __syncthreads();
        acc = GemmOp<T>::slice(sa, sb, ar, bc, lane, acc);
        __syncthreads();
    }
    // Element r of lane l: row 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32.
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int gm = m0 + ar + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int gn = n0 + bc + (lane & 31);
        if (gm >= M || gn >= N) continue;
        float v = acc[r];
        float *c = &C[(size_t)gm * ldc + gn];
        switch (epi) {
        case EPI_STORE: *c = v; break;
        case EPI_ACCUMULATE: *c += v; break;
        case EPI_BIAS: *c = v + bias[gn]; break;
        case EPI_BIAS_SILU:
            v += bias[gn];
            *c = v;
            reinterpret_cast<float *>(H)[(size_t)gm * ldc + gn] = silu(v);
            break;
        default:  // EPI_BIAS_SILU_HALF: only the activation, as FP16
            reinterpret_cast<_Float16 *>(H)[(size_t)gm * ldc + gn] = (_Float16)silu(v + bias[gn]);
            break;
        }
    }
}

template <class T, class S, bool TA, bool TB>
hipError_t gemm(int M, int N, int K, const S *A, int lda, const S *B, int ldb, float *C, int ldc, int epi,
                const float *bias, void *H, hipStream_t st) {
    if (M <= 0 || N <= 0) return hipSuccess;
    dim3 grid((N + kGT - 1) / kGT, (M + kGT - 1) / kGT);
    hipLaunchKernelGGL((k_gemm<T, S, TA, TB>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc, epi, bias, H);
    return hipGetLastError();
}

// ------------------------------------------------------------------ elementwise
__global__ void k_silu_back(size_t n, const float *__restrict__ dh, const float *__restrict__ z, float *dz) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float s = 1.0f / (1.0f + __expf(-z[i]));
    dz[i] = dh[i] * s * (1.0f + z[i] * (1.0f - s));
}

// dY = 2 (Y - T) / (n 8) and the batch's sum of squared errors.
__global__ void k_mse(size_t n, const float *__restrict__ y, const float *__restrict__ t, float *dy, double *sse) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double e2 = 0.0;
    if (i < n) {
        const float d = y[i] - t[i];
        dy[i] = 2.0f * d / (float)n;
        e2 = (double)d * d;
    }
    for (int o = 32; o > 0; o >>= 1) e2 += __shfl_xor(e2, o);
    if ((threadIdx.x & 63) == 0) atomicAdd(sse, e2);
}

// db[j] += sum_i dz[i][j] (db zeroed beforehand): a block takes 64 columns and one slice of
// rows, 4 threads per column, then one atomic add per column.
constexpr int kColRows = 1024;
__global__ void k_colsum(int n, int d, const float *__restrict__ dz, float *db) {
    __shared__ float part[4][64];
    const int c = threadIdx.x & 63, r = threadIdx.x >> 6;
    const int j = blockIdx.x * 64 + c;
    const int i0 = blockIdx.y * kColRows;
    const int i1 = min(n, i0 + kColRows);
    float s = 0.0f;
    if (j < d)
        for (int i = i0 + r; i < i1; i += 4) s += dz[(size_t)i * d + j];
    part[r][c] = s;
    __syncthreads();
    if (r == 0 && j < d) atomicAdd(&db[j], part[0][c] + part[1][c] + part[2][c] + part[3][c]);
}

__global__ void k_zero_f(int n, float *x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0f;
}

__global__ void k_adam(size_t n, float *w, const float *__restrict__ g, float *m, float *v, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_to_half(size_t n, const float *__restrict__ x, _Float16 *y) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (_Float16)x[i];
}

__global__ void k_zero_d(double *x) { *x = 0.0; }

// ------------------------------------------------------------------ samples and features
// SplitMix64 (Steele, Lea and Flood 2014): a counter-based stream per sample.
__device__ __forceinline__ unsigned long long splitmix(unsigned long long &s) {
    s += 0x9E3779B97F4A7C15ull;
    unsigned long long z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double uniform(unsigned long long &s) { return (double)(splitmix(s) >> 11) * 0x1.0p-53; }

__device__ void features(double az, double za, double freq, const unsigned *delays, const double *gains, float *x) {
    x[0] = (float)sin(az);
    x[1] = (float)cos(az);
    x[2] = (float)sin(za);
    x[3] = (float)cos(za);
    x[4] = (float)((freq - 182e6) / 15e6);
    for (int d = 0; d < 16; ++d) x[5 + d] = (float)((double)(delays[d] == 32 ? 0u : delays[d]) / 31.0);
    for (int d = 0; d < 32; ++d) x[21 + d] = (float)((delays[d & 15] == 32) ? 0.0 : gains[d]);
}

// The analytic beam for one (direction, frequency, tile): the per-direction form of
// k_analytic_beam (direct.hip), 16 dipole phasors summed directly.
__device__ void analytic_jones(double az, double za, double freq, const unsigned *delays, const double *gains,
                               double *out) {
    double s_za, c_za, s_az, c_az;
    sincos(za, &s_za, &c_za);
    sincos(az, &s_az, &c_az);
    const double r_e = s_za * s_az, r_n = s_za * c_az, r_u = c_za;
    const double k = 6.283185307179586 * freq / SM_VEL_C;
    double ax_r = 0, ax_i = 0, ay_r = 0, ay_i = 0;
    for (int d = 0; d < 16; ++d) {
        unsigned dl = delays[d];
        double gx = gains[d], gy = gains[16 + d];
        if (dl == 32) {
            gx = gy = 0.0;
            dl = 0;
        }
        const double x = ((double)(d % 4) - 1.5) * 1.1, y = (1.5 - (double)(d / 4)) * 1.1;
        double s, c;
        sincos(k * (x * r_e + y * r_n) - 6.283185307179586 * freq * ((double)dl * 435e-12), &s, &c);
        ax_r += gx * c;
        ax_i += gx * s;
        ay_r += gy * c;
        ay_i += gy * s;
    }
    const double g = sin(k * 0.278 * r_u) / (16.0 * sin(k * 0.278));
    double sx = sqrt(1.0 - r_e * r_e);
    if (sx < 1e-12) sx = 1e-12;
    const double px = g * sx, pyx = g * (-r_e * r_n / sx), pyy = g * (r_u / sx);
    out[0] = ax_r * px;
    out[1] = ax_i * px;
    out[2] = 0.0;
    out[3] = 0.0;
    out[4] = ay_r * pyx;
    out[5] = ay_i * pyx;
    out[6] = ay_r * pyy;
    out[7] = ay_i * pyy;
}

// Sample i (global index first + i): az ~ U[0, 2 pi), cos za ~ U[0, 1], f ~ U[167, 197] MHz,
// each of the 32 gains 0 with probability 0.02 (dead dipole), else 1, and delays
//   mode 0: each ~ U{0..31} independently;
//   mode 1: the delays that point the tile at a random direction within 60 deg of zenith
//           (cos za0 ~ U[cos 60 deg, 1], az0 ~ U[0, 2 pi)): round((x_d sin za0 sin az0 +
//           y_d sin za0 cos az0) / (c 435 ps)), shifted so the smallest is 0, capped at 31.
__device__ void draw(unsigned long long seed, long long gi, int mode, double &az, double &za, double &freq,
                     unsigned *delays, double *gains) {
    unsigned long long s = seed ^ (0xD1B54A32D192ED03ull * (unsigned long long)(gi + 1));
    az = 6.283185307179586 * uniform(s);
    za = acos(uniform(s));
    freq = 167e6 + 30e6 * uniform(s);
    if (mode == 0) {
        for (int d = 0; d < 16; ++d) delays[d] = (unsigned)(uniform(s) * 32.0) & 31u;
    } else {
        const double az0 = 6.283185307179586 * uniform(s);
        const double za0 = acos(0.5 + 0.5 * uniform(s));
        const double se = sin(za0) * sin(az0), sn = sin(za0) * cos(az0);
        double t[16], lo = 1e30;
        for (int d = 0; d < 16; ++d) {
            const double x = ((double)(d % 4) - 1.5) * 1.1, y = (1.5 - (double)(d / 4)) * 1.1;
            t[d] = rint((x * se + y * sn) / SM_VEL_C / 435e-12);
            lo = fmin(lo, t[d]);
        }
        for (int d = 0; d < 16; ++d) delays[d] = (unsigned)fmin(t[d] - lo, 31.0);
    }
    for (int d = 0; d < 32; ++d) gains[d] = uniform(s) < 0.02 ? 0.0 : 1.0;
}

__global__ void k_samples(unsigned long long seed, long long first, int n, int mode, float *x, float *y,
                          double *raw) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double az, za, freq, gains[32];
    unsigned delays[16];
    draw(seed, first + i, mode, az, za, freq, delays, gains);
    if (x != nullptr) features(az, za, freq, delays, gains, x + (size_t)i * kNIn);
    double j[8];
    analytic_jones(az, za, freq, delays, gains, j);
    if (y != nullptr)
        for (int k = 0; k < 8; ++k) y[(size_t)i * kNOut + k] = (float)j[k];
    if (raw != nullptr) {
        // az, za, freq, 16 delays, 32 gains (as doubles)
        double *r = raw + (size_t)i * 51;
        r[0] = az;
        r[1] = za;
        r[2] = freq;
        for (int d = 0; d < 16; ++d) r[3 + d] = (double)delays[d];
        for (int d = 0; d < 32; ++d) r[19 + d] = gains[d];
    }
}

// Features of every (freq, dir) of one tile, rows ordered [freq][dir].
__global__ void k_tile_features(int n_freqs, const double *__restrict__ freqs, int n_dirs, const double *__restrict__ az,
                                const double *__restrict__ za, const unsigned *__restrict__ delays,
                                const double *__restrict__ gains, float *x) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n_freqs * n_dirs) return;
    const int f = (int)(i / n_dirs), d = (int)(i % n_dirs);
    features(az[d], za[d], freqs[f], delays, gains, x + i * kNIn);
}

__global__ void k_out_to_jones(size_t n, const float *__restrict__ y, JonesF64 *out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double *o = reinterpret_cast<double *>(&out[i]);
    for (int k = 0; k < 8; ++k) o[k] = (double)y[i * kNOut + k];
}

__global__ void k_zero_bits(unsigned long long *x) { *x = 0ull; }

inline unsigned blocks(size_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

template <class T> hipError_t upload(DevBuf &d, const T *h, size_t n) {
    hipError_t e = d.reserve(sizeof(T) * (n > 0 ? n : 1));
    if (e != hipSuccess || n == 0) return e;
    return hipMemcpy(d.ptr, h, sizeof(T) * n, hipMemcpyHostToDevice);
}

// Probe of the surrogate against the exact beam: directions [0, n) sampled at a stride so
// that at most kProbeDirs are used, at up to kProbeFreqs frequencies spread over the band.
constexpr int kProbeDirs = 512;
constexpr int kProbeFreqs = 8;
__global__ void k_probe_err(int n_freqs, int n_dirs, int dstride, int fstride, int pf, int pd,
                            const double *__restrict__ freqs, const double *__restrict__ az,
                            const double *__restrict__ za, const unsigned *__restrict__ delays,
                            const double *__restrict__ gains, const JonesF64 *__restrict__ sur,
                            unsigned long long *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0.0;
    if (i < pf * pd) {
        const int f = (i / pd) * fstride, d = (i % pd) * dstride;
        double j[8];
        analytic_jones(az[d], za[d], freqs[f], delays, gains, j);
        const double *x = reinterpret_cast<const double *>(&sur[(size_t)f * n_dirs + d]);
        for (int k = 0; k < 4; ++k) m = fmax(m, hypot(x[2 * k] - j[2 * k], x[2 * k + 1] - j[2 * k + 1]));
    }
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

}  // namespace

}  // namespace sm

using namespace sm;

// ------------------------------------------------------------------ fused FP16 inference
// The whole network for 64 samples per 4-wave work-group, activations resident in LDS (two
// 64 x 256 FP16 buffers, so two work-groups share a CU), weights read as MFMA fragments
// straight from global memory (L2-resident, 16 bytes per lane). Each wave computes 64 neurons
// of every hidden layer for 64 samples (2 x 2 tiles of v_mfma_f32_32x32x16_f16), then bias,
// SiLU and the FP16 store into the other buffer; the last layer (8 outputs, padded to 32)
// gives each of the first waves one 32-sample tile. (Measured on 6.1 M samples: 7.5 ms, against
// 9.5 ms with 128 samples per 8-wave group and 9.1 ms with 128 per 4-wave group.) The products and their K order are those of the layer-by-layer
// GEMM; the SiLU uses a fast reciprocal.
constexpr int kK0 = 64;        // layer-0 K, 53 features padded
constexpr int kMR = 64;        // samples per work-group (two work-groups per CU)
constexpr int kMW = kMR / 16;  // waves per work-group: 4 neuron groups x (kMR / 64) sample halves
constexpr int kST = 2;         // sample tiles per wave
constexpr int kActPad = 8;     // FP16 row padding: 528-byte rows put 8 rows on distinct banks

struct MlpArgs {
    const _Float16 *w0;        // [256][64]
    const _Float16 *w[3];      // layers 1..3, [256][256]
    const _Float16 *w4;        // [32][256]
    const float *b[kLayers];
};

__global__ void k_pad_half(int rows, int cols, int rows_in, int cols_in, const float *__restrict__ src,
                           _Float16 *__restrict__ dst) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)rows * cols) return;
    const int r = (int)(i / cols), c = (int)(i % cols);
    dst[i] = (_Float16)(r < rows_in && c < cols_in ? src[(size_t)r * cols_in + c] : 0.0f);
}

// SiLU with a fast reciprocal: the network is evaluated in FP16, far coarser than this.
__device__ __forceinline__ float silu_fast(float z) { return __fdividef(z, 1.0f + __expf(-z)); }

// One hidden layer, computed transposed (out^T = W in^T): the weights are the MFMA A operand
// and the activations the B operand, so each lane's 4 consecutive accumulator elements are 4
// consecutive neurons of one sample and leave as one 8-byte LDS store.
template <int K>
__device__ __forceinline__ void mlp_hidden(const _Float16 (*in_all)[kHid + kActPad],
                                           _Float16 (*out_all)[kHid + kActPad], const _Float16 *__restrict__ w,
                                           const float *__restrict__ bias, int wave, int lane) {
    // Wave w: neurons 64 (w % 4) .. + 63 for samples 64 (w / 4) .. + 63.
    const _Float16 (*in)[kHid + kActPad] = in_all + 64 * (wave >> 2);
    _Float16 (*out)[kHid + kActPad] = out_all + 64 * (wave >> 2);
    wave &= 3;
    typedef _Float16 h4 __attribute__((ext_vector_type(4)));
    f16v acc[2][kST];  // [neuron tile][sample tile]
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int st = 0; st < kST; ++st)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[nt][st][r] = 0.0f;
    const int n0 = wave * 64;
    const int kl = 8 * (lane >> 5);
#pragma unroll 4
    for (int k0 = 0; k0 < K; k0 += 16) {
        h8 wa[2], xb[kST];
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
            wa[nt] = *reinterpret_cast<const h8 *>(&w[(size_t)(n0 + nt * 32 + (lane & 31)) * K + k0 + kl]);
#pragma unroll
        for (int st = 0; st < kST; ++st) xb[st] = *reinterpret_cast<const h8 *>(&in[st * 32 + (lane & 31)][k0 + kl]);
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int st = 0; st < kST; ++st)
                acc[nt][st] = __builtin_amdgcn_mfma_f32_32x32x16_f16(wa[nt], xb[st], acc[nt][st], 0, 0, 0);
    }
    // Element r of lane l: neuron 8 (r / 4) + 4 (l / 32) + r % 4 of the tile, sample l % 32.
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int nb = n0 + nt * 32 + 8 * g + 4 * (lane >> 5);
            const float4 bc = *reinterpret_cast<const float4 *>(&bias[nb]);
#pragma unroll
            for (int st = 0; st < kST; ++st) {
                h4 v;
                v[0] = (_Float16)silu_fast(acc[nt][st][4 * g] + bc.x);
                v[1] = (_Float16)silu_fast(acc[nt][st][4 * g + 1] + bc.y);
                v[2] = (_Float16)silu_fast(acc[nt][st][4 * g + 2] + bc.z);
                v[3] = (_Float16)silu_fast(acc[nt][st][4 * g + 3] + bc.w);
                *reinterpret_cast<h4 *>(&out[st * 32 + (lane & 31)][nb]) = v;
            }
        }
}

__global__ __launch_bounds__(kMW * 64) void k_mlp16(int n, const float *__restrict__ x, const MlpArgs a,
                                               float *__restrict__ y) {
    __shared__ __attribute__((aligned(16))) _Float16 act[2][kMR][kHid + kActPad];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const size_t row0 = (size_t)blockIdx.x * kMR;
    // Features in FP16, K padded with zeros to 64; rows past n are zeros.
    for (int e = tid; e < kMR * kK0; e += kMW * 64) {
        const int r = e / kK0, k = e % kK0;
        const size_t gr = row0 + r;
        act[0][r][k] = (_Float16)(gr < (size_t)n && k < kNIn ? x[gr * kNIn + k] : 0.0f);
    }
    __syncthreads();
    mlp_hidden<kK0>(act[0], act[1], a.w0, a.b[0], wave, lane);
    __syncthreads();
    mlp_hidden<kHid>(act[1], act[0], a.w[0], a.b[1], wave, lane);
    __syncthreads();
    mlp_hidden<kHid>(act[0], act[1], a.w[1], a.b[2], wave, lane);
    __syncthreads();
    mlp_hidden<kHid>(act[1], act[0], a.w[2], a.b[3], wave, lane);
    __syncthreads();
    // Output layer: wave w < 4 takes samples 32 w .. + 31; columns 0..7 of the padded 32 are
    // real.
    if (wave >= kMR / 32) return;
    f16v acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    const int kl = 8 * (lane >> 5);
#pragma unroll 4
    for (int k0 = 0; k0 < kHid; k0 += 16) {
        const h8 b = *reinterpret_cast<const h8 *>(&a.w4[(size_t)(lane & 31) * kHid + k0 + kl]);
        const h8 av = *reinterpret_cast<const h8 *>(&act[0][wave * 32 + (lane & 31)][k0 + kl]);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(av, b, acc, 0, 0, 0);
    }
    const int col = lane & 31;
    if (col < kNOut) {
        const float bc = a.b[kLayers - 1][col];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const size_t gr = row0 + wave * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            if (gr < (size_t)n) y[gr * kNOut + col] = acc[r] + bc;
        }
    }
}

struct SmSurrogate {
    int hid = kHid, nl = kLayers;          // hidden width; layers (hidden layers + output)
    int din[kMaxLayers] = {}, dout[kMaxLayers] = {};
    DevBuf w[kMaxLayers], b[kMaxLayers], mw[kMaxLayers], vw[kMaxLayers], mb[kMaxLayers], vb[kMaxLayers],
        gw[kMaxLayers], gb[kMaxLayers];
    DevBuf w16[kLayers];
    DevBuf w0p, w4p;  // FP16 layer 0 padded to K = 64 and layer 4 padded to 32 outputs
    bool w16_valid = false;
    DevBuf z[kMaxLayers], h[kMaxLayers];  // training activations: z pre-activation, h post (h[0] = input)
    DevBuf dz, dh, dz2, sse;
    DevBuf a16[2], zf;              // inference ping-pong FP16 activations and FP32 scratch
    DevBuf bits, feat, fout;
    int t = 0;                      // Adam step count
};

namespace {

size_t wsize(const SmSurrogate *s, int l) { return (size_t)s->dout[l] * s->din[l]; }

// The shape the fused FP16 inference kernel is built for.
bool fused_shape(const SmSurrogate *s) { return s->hid == kHid && s->nl == kLayers; }

const char *ensure_w16(SmSurrogate *s) {
    if (!fused_shape(s)) return "the FP16 network is 4 x 256; other shapes evaluate in FP32 (dtype 0)";
    if (s->w16_valid) return nullptr;
    for (int l = 0; l < kLayers; ++l) {
        SM_TRY(s->w16[l].reserve(sizeof(_Float16) * wsize(s, l)));
        hipLaunchKernelGGL(k_to_half, dim3(blocks(wsize(s, l))), dim3(256), 0, 0, wsize(s, l), s->w[l].as<float>(),
                           s->w16[l].as<_Float16>());
    }
    SM_TRY(s->w0p.reserve(sizeof(_Float16) * kHid * kK0));
    SM_TRY(s->w4p.reserve(sizeof(_Float16) * 32 * kHid));
    hipLaunchKernelGGL(k_pad_half, dim3(blocks((size_t)kHid * kK0)), dim3(256), 0, 0, kHid, kK0, kHid, kNIn,
                       s->w[0].as<float>(), s->w0p.as<_Float16>());
    hipLaunchKernelGGL(k_pad_half, dim3(blocks((size_t)32 * kHid)), dim3(256), 0, 0, 32, kHid, kNOut, kHid,
                       s->w[kLayers - 1].as<float>(), s->w4p.as<_Float16>());
    SM_TRY(hipGetLastError());
    s->w16_valid = true;
    return nullptr;
}

// FP32 forward over n rows, keeping z and h of every layer; output in z[nl - 1].
const char *forward32(SmSurrogate *s, int n, const float *d_x) {
    const float *hin = d_x;
    const int L = s->nl;
    for (int l = 0; l < L; ++l) {
        SM_TRY(s->z[l].reserve(sizeof(float) * (size_t)n * s->dout[l]));
        if (l < L - 1) SM_TRY(s->h[l + 1].reserve(sizeof(float) * (size_t)n * s->dout[l]));
        const int epi = l < L - 1 ? EPI_BIAS_SILU : EPI_BIAS;
        SM_TRY((gemm<float, float, false, true>(n, s->dout[l], s->din[l], hin, s->din[l], s->w[l].as<float>(),
                                                s->din[l], s->z[l].as<float>(), s->dout[l], epi, s->b[l].as<float>(),
                                                l < L - 1 ? s->h[l + 1].ptr : nullptr, 0)));
        if (l < L - 1) hin = s->h[l + 1].as<float>();
    }
    return nullptr;
}

// Loss and gradients into gw / gb for the batch.
const char *backward32(SmSurrogate *s, int n, const float *d_x, const float *d_y, double *loss) {
    if (const char *e = forward32(s, n, d_x)) return e;
    SM_TRY(s->sse.reserve(sizeof(double)));
    const int wmax = s->hid > kNOut ? s->hid : kNOut;
    SM_TRY(s->dz.reserve(sizeof(float) * (size_t)n * wmax));
    SM_TRY(s->dh.reserve(sizeof(float) * (size_t)n * wmax));
    SM_TRY(s->dz2.reserve(sizeof(float) * (size_t)n * wmax));
    hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(1), 0, 0, s->sse.as<double>());
    const size_t no = (size_t)n * kNOut;
    hipLaunchKernelGGL(k_mse, dim3(blocks(no)), dim3(256), 0, 0, no, s->z[s->nl - 1].as<float>(), d_y,
                       s->dz.as<float>(), s->sse.as<double>());
    float *dz = s->dz.as<float>(), *dzn = s->dz2.as<float>();
    for (int l = s->nl - 1; l >= 0; --l) {
        const float *hin = l == 0 ? d_x : s->h[l].as<float>();
        // dW = dZ^T H_in (out x in), db = column sums of dZ.
        SM_TRY((gemm<float, float, true, false>(s->dout[l], s->din[l], n, dz, s->dout[l], hin, s->din[l],
                                                s->gw[l].as<float>(), s->din[l], EPI_STORE, nullptr, nullptr, 0)));
        hipLaunchKernelGGL(k_zero_f, dim3(1), dim3(256), 0, 0, s->dout[l], s->gb[l].as<float>());
        hipLaunchKernelGGL(k_colsum, dim3((s->dout[l] + 63) / 64, (n + kColRows - 1) / kColRows), dim3(256), 0, 0, n,
                           s->dout[l], dz, s->gb[l].as<float>());
        if (l == 0) break;
        // dH_in = dZ W (n x in), then through SiLU of the layer below.
        SM_TRY((gemm<float, float, false, false>(n, s->din[l], s->dout[l], dz, s->dout[l], s->w[l].as<float>(),
                                                 s->din[l], s->dh.as<float>(), s->din[l], EPI_STORE, nullptr, nullptr,
                                                 0)));
        const size_t nh = (size_t)n * s->din[l];
        hipLaunchKernelGGL(k_silu_back, dim3(blocks(nh)), dim3(256), 0, 0, nh, s->dh.as<float>(),
                           s->z[l - 1].as<float>(), dzn);
        std::swap(dz, dzn);
    }
    SM_TRY(hipGetLastError());
    if (loss != nullptr) {
        double sse = 0.0;
        SM_TRY(hipMemcpy(&sse, s->sse.ptr, sizeof(double), hipMemcpyDeviceToHost));
        *loss = sse / (double)no;
    }
    return nullptr;
}

// FP16 MFMA forward over n rows of FP32 features; FP32 outputs [n][8].
const char *forward16(SmSurrogate *s, int n, const float *d_x, float *d_y) {
    if (const char *e = ensure_w16(s)) return e;
    MlpArgs a;
    a.w0 = s->w0p.as<_Float16>();
    for (int l = 1; l < kLayers - 1; ++l) a.w[l - 1] = s->w16[l].as<_Float16>();
    a.w4 = s->w4p.as<_Float16>();
    for (int l = 0; l < kLayers; ++l) a.b[l] = s->b[l].as<float>();
    const unsigned grid = (unsigned)(((size_t)n + kMR - 1) / kMR);
    hipLaunchKernelGGL(k_mlp16, dim3(grid), dim3(kMW * 64), 0, 0, n, d_x, a, d_y);
    SM_TRY(hipGetLastError());
    return nullptr;
}

}  // namespace

extern "C" {

int32_t sm_surrogate_num_params(void) {
    size_t n = 0;
    for (int l = 0; l < kLayers; ++l) n += (size_t)kDimOut[l] * kDimIn[l] + kDimOut[l];
    return (int32_t)n;
}

const char *sm_surrogate_new_shaped(uint64_t seed, int32_t hidden, int32_t depth, SmSurrogate **out) {
    if (out == nullptr) return "null argument";
    *out = nullptr;
    if (hidden < 1 || hidden > 4096) return "hidden width must be 1..4096";
    if (depth < 1 || depth > kMaxLayers - 1) return "depth must be 1..8 hidden layers";
    SmSurrogate *s = new SmSurrogate();
    s->hid = hidden;
    s->nl = depth + 1;
    for (int l = 0; l < s->nl; ++l) {
        s->din[l] = l == 0 ? kNIn : hidden;
        s->dout[l] = l == s->nl - 1 ? kNOut : hidden;
    }
    // Uniform(-sqrt(6 / fan_in), +) weights (He-style for SiLU), zero biases, from a host
    // SplitMix64 stream.
    unsigned long long st = seed;
    auto next = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        unsigned long long z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return (double)((z ^ (z >> 31)) >> 11) * 0x1.0p-53;
    };
    for (int l = 0; l < s->nl; ++l) {
        std::vector<float> w(wsize(s, l)), zb(s->dout[l], 0.0f), zw(wsize(s, l), 0.0f);
        const double a = std::sqrt(6.0 / s->din[l]) * (l == s->nl - 1 ? 0.25 : 1.0);
        for (auto &x : w) x = (float)((2.0 * next() - 1.0) * a);
        const char *e = nullptr;
        auto put = [&](DevBuf &d, const std::vector<float> &v) {
            if (e) return;
            if (d.reserve(sizeof(float) * v.size()) != hipSuccess ||
                hipMemcpy(d.ptr, v.data(), sizeof(float) * v.size(), hipMemcpyHostToDevice) != hipSuccess)
                e = "surrogate: device allocation or copy failed";
        };
        put(s->w[l], w);
        put(s->b[l], zb);
        put(s->mw[l], zw);
        put(s->vw[l], zw);
        put(s->gw[l], zw);
        put(s->mb[l], zb);
        put(s->vb[l], zb);
        put(s->gb[l], zb);
        if (e) {
            delete s;
            return e;
        }
    }
    *out = s;
    return nullptr;
}

const char *sm_surrogate_new(uint64_t seed, SmSurrogate **out) { return sm_surrogate_new_shaped(seed, kHid, 4, out); }

const char *sm_surrogate_shape(const SmSurrogate *s, int32_t *hidden, int32_t *depth, int64_t *num_params) {
    if (s == nullptr) return "null argument";
    size_t n = 0;
    for (int l = 0; l < s->nl; ++l) n += wsize(s, l) + s->dout[l];
    if (hidden != nullptr) *hidden = s->hid;
    if (depth != nullptr) *depth = s->nl - 1;
    if (num_params != nullptr) *num_params = (int64_t)n;
    return nullptr;
}

void sm_surrogate_free(SmSurrogate *s) {
    if (s == nullptr) return;
    (void)hipDeviceSynchronize();
    delete s;
}

const char *sm_surrogate_get_params(SmSurrogate *s, float *host) {
    if (s == nullptr || host == nullptr) return "null argument";
    size_t o = 0;
    for (int l = 0; l < s->nl; ++l) {
        SM_TRY(hipMemcpy(host + o, s->w[l].ptr, sizeof(float) * wsize(s, l), hipMemcpyDeviceToHost));
        o += wsize(s, l);
        SM_TRY(hipMemcpy(host + o, s->b[l].ptr, sizeof(float) * s->dout[l], hipMemcpyDeviceToHost));
        o += s->dout[l];
    }
    return nullptr;
}

const char *sm_surrogate_set_params(SmSurrogate *s, const float *host) {
    if (s == nullptr || host == nullptr) return "null argument";
    size_t o = 0;
    for (int l = 0; l < s->nl; ++l) {
        SM_TRY(hipMemcpy(s->w[l].ptr, host + o, sizeof(float) * wsize(s, l), hipMemcpyHostToDevice));
        o += wsize(s, l);
        SM_TRY(hipMemcpy(s->b[l].ptr, host + o, sizeof(float) * s->dout[l], hipMemcpyHostToDevice));
        o += s->dout[l];
    }
    s->w16_valid = false;
    return nullptr;
}

const char *sm_surrogate_samples(uint64_t seed, int64_t first, int32_t n, int32_t mode, float *d_x, float *d_y,
                                 double *d_raw) {
    if (n < 0) return "negative sample count";
    if (mode != 0 && mode != 1) return "mode must be 0 (independent delays) or 1 (pointing delays)";
    if (n == 0) return nullptr;
    hipLaunchKernelGGL(k_samples, dim3(blocks(n)), dim3(256), 0, 0, (unsigned long long)seed, (long long)first, n,
                       mode, d_x, d_y, d_raw);
    SM_TRY(hipGetLastError());
    return nullptr;
}

const char *sm_surrogate_gradients(SmSurrogate *s, int32_t n, const float *d_x, const float *d_y, double *loss,
                                   float *host_grads) {
    if (s == nullptr || d_x == nullptr || d_y == nullptr) return "null argument";
    if (n <= 0) return "batch must be positive";
    if (const char *e = backward32(s, n, d_x, d_y, loss)) return e;
    if (host_grads != nullptr) {
        size_t o = 0;
        for (int l = 0; l < s->nl; ++l) {
            SM_TRY(hipMemcpy(host_grads + o, s->gw[l].ptr, sizeof(float) * wsize(s, l), hipMemcpyDeviceToHost));
            o += wsize(s, l);
            SM_TRY(hipMemcpy(host_grads + o, s->gb[l].ptr, sizeof(float) * s->dout[l], hipMemcpyDeviceToHost));
            o += s->dout[l];
        }
    }
    return nullptr;
}

const char *sm_surrogate_train_step(SmSurrogate *s, int32_t n, const float *d_x, const float *d_y, float lr,
                                    double *loss) {
    if (s == nullptr || d_x == nullptr || d_y == nullptr) return "null argument";
    if (n <= 0) return "batch must be positive";
    if (const char *e = backward32(s, n, d_x, d_y, loss)) return e;
    s->t += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f - std::pow(b1, (float)s->t), c2 = 1.0f - std::pow(b2, (float)s->t);
    for (int l = 0; l < s->nl; ++l) {
        hipLaunchKernelGGL(k_adam, dim3(blocks(wsize(s, l))), dim3(256), 0, 0, wsize(s, l), s->w[l].as<float>(),
                           s->gw[l].as<float>(), s->mw[l].as<float>(), s->vw[l].as<float>(), lr, b1, b2, eps, c1, c2);
        hipLaunchKernelGGL(k_adam, dim3(blocks(s->dout[l])), dim3(256), 0, 0, (size_t)s->dout[l], s->b[l].as<float>(),
                           s->gb[l].as<float>(), s->mb[l].as<float>(), s->vb[l].as<float>(), lr, b1, b2, eps, c1, c2);
    }
    SM_TRY(hipGetLastError());
    s->w16_valid = false;
    return nullptr;
}

const char *sm_surrogate_forward(SmSurrogate *s, int32_t n, const float *d_x, int32_t dtype, float *d_y) {
    if (s == nullptr || d_x == nullptr || d_y == nullptr) return "null argument";
    if (n <= 0) return nullptr;
    if (dtype == 0) {
        if (const char *e = forward32(s, n, d_x)) return e;
        SM_TRY(hipMemcpyAsync(d_y, s->z[s->nl - 1].ptr, sizeof(float) * (size_t)n * kNOut, hipMemcpyDeviceToDevice,
                              0));
        return nullptr;
    }
    if (dtype != 1) return "dtype must be 0 (FP32) or 1 (FP16 MFMA)";
    return forward16(s, n, d_x, d_y);
}

const char *sm_surrogate_beam(SmSurrogate *s, int32_t n_tiles, const uint32_t *delays, const double *gains,
                              int32_t n_freqs, const double *freqs, int32_t n_dirs, const double *d_az,
                              const double *d_za, double threshold, JonesF64 *d_out, int32_t *n_fallback) {
    if (s == nullptr || delays == nullptr || gains == nullptr || freqs == nullptr || d_out == nullptr)
        return "null argument";
    if (n_tiles <= 0 || n_freqs <= 0 || n_dirs <= 0) return nullptr;
    DevBuf dd, dg, df;
    SM_TRY(upload(dd, delays, 16 * (size_t)n_tiles));
    SM_TRY(upload(dg, gains, 32 * (size_t)n_tiles));
    SM_TRY(upload(df, freqs, (size_t)n_freqs));
    const size_t per_tile = (size_t)n_freqs * n_dirs;
    SM_TRY(s->feat.reserve(sizeof(float) * per_tile * kNIn));
    SM_TRY(s->fout.reserve(sizeof(float) * per_tile * kNOut));
    SM_TRY(s->bits.reserve(sizeof(unsigned long long)));
    const int dstride = (n_dirs + kProbeDirs - 1) / kProbeDirs, pd = (n_dirs + dstride - 1) / dstride;
    const int fstride = (n_freqs + kProbeFreqs - 1) / kProbeFreqs, pf = (n_freqs + fstride - 1) / fstride;
    int fallbacks = 0;
    for (int t = 0; t < n_tiles; ++t) {
        JonesF64 *out = d_out + (size_t)t * per_tile;
        const unsigned *dl = dd.as<unsigned>() + 16 * (size_t)t;
        const double *gn = dg.as<double>() + 32 * (size_t)t;
        hipLaunchKernelGGL(k_tile_features, dim3(blocks(per_tile)), dim3(256), 0, 0, n_freqs, df.as<double>(), n_dirs,
                           d_az, d_za, dl, gn, s->feat.as<float>());
        if (fused_shape(s)) {
            if (const char *e = forward16(s, (int)per_tile, s->feat.as<float>(), s->fout.as<float>())) return e;
        } else {
            if (const char *e = forward32(s, (int)per_tile, s->feat.as<float>())) return e;
            SM_TRY(hipMemcpyAsync(s->fout.ptr, s->z[s->nl - 1].ptr, sizeof(float) * per_tile * kNOut,
                                  hipMemcpyDeviceToDevice, 0));
        }
        hipLaunchKernelGGL(k_out_to_jones, dim3(blocks(per_tile)), dim3(256), 0, 0, per_tile, s->fout.as<float>(), out);
        // Judge: the exact beam on the probe directions; the whole tile takes the exact beam
        // where the surrogate misses by more than the threshold.
        hipLaunchKernelGGL(k_zero_bits, dim3(1), dim3(1), 0, 0, s->bits.as<unsigned long long>());
        hipLaunchKernelGGL(k_probe_err, dim3(blocks((size_t)pf * pd)), dim3(256), 0, 0, n_freqs, n_dirs, dstride,
                           fstride, pf, pd, df.as<double>(), d_az, d_za, dl, gn, out,
                           s->bits.as<unsigned long long>());
        unsigned long long b = 0;
        SM_TRY(hipMemcpy(&b, s->bits.ptr, sizeof(b), hipMemcpyDeviceToHost));
        double err;
        std::memcpy(&err, &b, sizeof(err));
        if (!(err <= threshold)) {
            SM_TRY(launch_analytic_beam(1, dl, gn, n_freqs, df.as<double>(), n_dirs, d_az, d_za, out, 0));
            ++fallbacks;
        }
    }
    SM_TRY(hipGetLastError());
    if (n_fallback != nullptr) *n_fallback = fallbacks;
    return nullptr;
}

}  // extern "C"
