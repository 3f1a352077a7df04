// Split-precision MFMA path for point components (no counterpart in the reference).
//
// With one beam shared by every tile, V^e_pq(f) = sum_c K_pc C^e_c K*_qc for each Jones
// entry e, where K_pc = e^{i theta_pc f / c} (theta_pc = uvw_p . LMN_c, tile uvw) and
// C_c = J S_c J^H. Per channel that is a complex GEMM (T x Nc) . (Nc x T); in real form
//   [V_re | V_im] = [A_re A_im] . [[K_re, -K_im], [K_im, K_re]],  A = K (.) C^e,
// i.e. M = 4 Jones entries x T tiles, N = 2 x T, K = 2 Nc.
//
// Two kernels per chunk of components:
//   * k_factored_phasors evaluates every K_tc(f) once: the phase theta f / c and its range
//     reduction in FP64 at the first channel of each 8-channel group, an FP64 rotation
//     across the group's evenly spaced channels, stored as FP32 complex [f][tile][comp].
//   * k_factored: a work-group (4 waves) owns one channel, one 64 x 64 block of tile pairs
//     (p-block <= q-block; blocks below the diagonal are never computed) and two of the four
//     Jones entries. Per step of 16 components it forms A = K C^e s (s a power-of-two scale
//     keeping |A| inside the fp16 range) and the K operand, splits each value x into
//     hi = rn(x), lo = rn(x - hi) in the MFMA input type (fp16 or bf16) in LDS, and runs
//     hi*hi, hi*lo and lo*hi on v_mfma_f32_32x32x16_{f16,bf16} (2 x 2 output tiles of 32 x 32
//     per wave): hi*hi into one FP32 accumulator, the two corrections into a second, so their
//     roundings stay apart from the large partial sums. For fp16, lo*lo (about 2^-22 of the
//     product) is below the other errors and is not computed; bf16 adds it. Every `seg_steps`
//     steps the FP32 partials are added into an FP64 buffer. Two work-groups share a CU, so
//     one group's operand generation overlaps the other's MFMAs.
//   * k_factored_ws (the default) does the same work with the roles split between waves:
//     per work-group, 8 producer waves generate the operands of step s + 1 into LDS while 4
//     consumer waves (one per SIMD) run the MFMAs of step s. The products, their order and
//     the results are those of k_factored, bit for bit.
// A scatter kernel maps the FP64 block results to baselines and accumulates visibilities.
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "sm_factored.h"
#include "sm_jones.h"
#include "sm_math.h"

namespace sm {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

constexpr int kFB = 64;             // tiles per block side
constexpr int kFC = 16;             // components per step
constexpr int kFK = 2 * kFC;        // real K per step
// LDS operand rows are kFK = 32 elements (64 bytes, four 16-byte chunks) with no padding;
// chunk c of row r lives at chunk c ^ swz(r). A 16-byte access by 16 lanes then touches 16
// distinct 4-bank groups both for the fragment reads (16 rows, one chunk) and for the
// generation writes (4 rows x 4 chunks).
__device__ __forceinline__ int swz(int row) { return kFK == 32 ? (row >> 2) & 3 : row & 7; }
__device__ __forceinline__ int lds_at(int row, int k) { return row * kFK + ((((k >> 3) ^ swz(row)) << 3) | (k & 7)); }
constexpr size_t kKbufCap = size_t(4) << 30;  // phasor buffer bytes per chunk

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

// Splits (x0, x1) into hi = rn_E(x) and lo = rn_E(x - hi), each pair packed in one 32-bit
// word (x0 in the low half). The empty asm rounds x to FP32 first: it keeps the compiler from
// folding the producing FMA and the conversion into one mixed-precision FMA that rounds once
// to 16 bits, which would make the split depend on instruction selection.
template <class E> __device__ __forceinline__ void split_words(float x0, float x1, unsigned &hi, unsigned &lo);

// fp16: one packed conversion for hi, then lo = rn16(x - hi) as one mixed-precision FMA per
// element (x * 1 - hi, hi read as fp16 from its half of the packed word). x - hi is exact in
// FP32, so the FMA's single rounding gives the same lo as rounding twice.
template <> __device__ __forceinline__ void split_words<_Float16>(float x0, float x1, unsigned &hi, unsigned &lo) {
    asm("" : "+v"(x0));
    asm("" : "+v"(x1));
    typedef _Float16 h2 __attribute__((ext_vector_type(2)));
    hi = __builtin_bit_cast(unsigned, h2{(_Float16)x0, (_Float16)x1});
    asm("v_fma_mixlo_f16 %0, %1, 1.0, -%2 op_sel_hi:[0,0,1]" : "=v"(lo) : "v"(x0), "v"(hi));
    asm("v_fma_mixhi_f16 %0, %1, 1.0, -%2 op_sel:[0,0,1] op_sel_hi:[0,0,1]" : "+v"(lo) : "v"(x1), "v"(hi));
}

// bf16: packed round-to-nearest-even conversions; hi widens back to FP32 by a shift (low
// half) and a mask (high half).
template <> __device__ __forceinline__ void split_words<__bf16>(float x0, float x1, unsigned &hi, unsigned &lo) {
    asm("" : "+v"(x0));
    asm("" : "+v"(x1));
    asm("v_cvt_pk_bf16_f32 %0, %1, %2" : "=v"(hi) : "v"(x0), "v"(x1));
    const float d0 = x0 - __builtin_bit_cast(float, hi << 16);
    const float d1 = x1 - __builtin_bit_cast(float, hi & 0xffff0000u);
    asm("v_cvt_pk_bf16_f32 %0, %1, %2" : "=v"(lo) : "v"(d0), "v"(d1));
}

// (re, im) -> (-im, re) on a packed pair of 16-bit floats.
__device__ __forceinline__ unsigned swap_neg(unsigned w) { return ((w >> 16) | (w << 16)) ^ 0x8000u; }

// K[f][t][c] for the chunk's components; tiles >= ntiles and components past the chunk
// hold 0, so the MFMA kernel needs no bounds checks.
__global__ __launch_bounds__(256) void k_factored_phasors(const FactoredArgs a) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)a.tpad * a.cpad) return;
    const int cl = (int)(i % a.cpad);
    const int t = (int)(i / a.cpad);
    const int c = a.c_begin + cl;
    const size_t fstride = (size_t)a.tpad * a.cpad;
    float2 *dst = a.kbuf + i;
    if (t >= a.ntiles || cl >= a.c_count || c >= a.ncomp) {
        for (int f = 0; f < a.nchan; ++f) {
            dst[f * fstride] = make_float2(0.0f, 0.0f);
            if (a.khi != nullptr) a.khi[i + f * fstride] = a.klo[i + f * fstride] = 0u;
        }
        return;
    }
    const double4 l = a.lmn[c];
    const double th = a.tile_uvw[3 * t] * l.x + a.tile_uvw[3 * t + 1] * l.y + a.tile_uvw[3 * t + 2] * l.z;
    // Stores K at channel f, and with planes its hi / lo split in the MFMA type.
    auto put = [&](int f, double co, double s) {
        const float re = (float)co, im = (float)s;
        dst[f * fstride] = make_float2(re, im);
        if (a.khi != nullptr) {
            unsigned hi, lo;
            if (a.planes_dtype == SM_MFMA_BF16)
                split_words<__bf16>(re, im, hi, lo);
            else
                split_words<_Float16>(re, im, hi, lo);
            a.khi[i + f * fstride] = hi;
            a.klo[i + f * fstride] = lo;
        }
    };
    for (int f0 = 0; f0 < a.nchan; f0 += SM_NCH) {
        double s, co;
        table_sincos(th * (a.freqs[f0] / SM_VEL_C), kSinCos32, s, co);
        put(f0, co, s);
        if (a.even) {
            double ds, dc;
            table_sincos(th * ((a.freqs[f0 + 1] - a.freqs[f0]) / SM_VEL_C), kSinCos32, ds, dc);
            for (int k = 1; k < SM_NCH && f0 + k < a.nchan; ++k) {
                const double ns = s * dc + co * ds;
                co = co * dc - s * ds;
                s = ns;
                put(f0 + k, co, s);
            }
        } else {
            for (int k = 1; k < SM_NCH && f0 + k < a.nchan; ++k) {
                table_sincos(th * (a.freqs[f0 + k] / SM_VEL_C), kSinCos32, s, co);
                put(f0 + k, co, s);
            }
        }
    }
}

// cohf[f][c][8] = J S J^H * scale in FP32 for the chunk's components (0 past the chunk).
__global__ __launch_bounds__(256) void k_factored_coh32(const FactoredArgs a) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)a.nchan * a.cpad * 8) return;
    const int k = (int)(i & 7);
    const size_t fc = i >> 3;
    const int cl = (int)(fc % a.cpad);
    const int f = (int)(fc / a.cpad);
    float v = 0.0f;
    if (cl < a.c_count) v = (float)(a.coh[((size_t)(a.c_begin + cl) * a.nchan_pad + f) * 8 + k] * *a.scale);
    reinterpret_cast<float *>(a.cohf)[i] = v;
}

// Per-tile beams: cohf[f][r][c][8] = J_r S J_0^H * scale for class r (beam row class_row[r])
// and every component of the chunk; coh holds S.
__device__ __forceinline__ void class_coherency(const FactoredArgs &a, int r, int c, int f, double o[8]) {
    double sv[8], ja[8], jb[8];
    load_jones(reinterpret_cast<const JonesF64 *>(a.coh + ((size_t)c * a.nchan_pad + f) * 8), sv);
    const BeamDesc &b = a.beam;
    const size_t col = (size_t)b.freq_map[f] * b.n_dir + b.dir_offset + c;
    load_jones(&b.jones[(size_t)a.class_row[r] * b.n_ufreq * b.n_dir + col], ja);
    load_jones(&b.jones[(size_t)a.class_row[0] * b.n_ufreq * b.n_dir + col], jb);
    sandwich(ja, sv, jb, o);
}

__global__ __launch_bounds__(256) void k_factored_coh32_classes(const FactoredArgs a) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)a.nchan * a.nclass * a.cpad) return;
    const int cl = (int)(i % a.cpad);
    const size_t fr = i / a.cpad;
    const int r = (int)(fr % a.nclass);
    const int f = (int)(fr / a.nclass);
    float *dst = reinterpret_cast<float *>(a.cohf) + i * 8;
    if (cl >= a.c_count) {
#pragma unroll
        for (int k = 0; k < 8; ++k) dst[k] = 0.0f;
        return;
    }
    double o[8];
    class_coherency(a, r, a.c_begin + cl, f, o);
    const double sc = *a.scale;
#pragma unroll
    for (int k = 0; k < 8; ++k) dst[k] = (float)(o[k] * sc);
}

// max |J_r S J_0^H| over classes, components and channels, for the scale. A thread takes one
// (channel, component) with components fastest, so the Jones reads of a wave are contiguous,
// and loops over the classes with S and J_0 loaded once.
__global__ void k_absmax_classes(const FactoredArgs a, unsigned long long *out) {
    double m = 0.0;
    const size_t n = (size_t)a.ncomp * a.nchan;
    const BeamDesc &b = a.beam;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % a.ncomp);
        const int f = (int)(i / a.ncomp);
        double sv[8], j0[8];
        load_jones(reinterpret_cast<const JonesF64 *>(a.coh + ((size_t)c * a.nchan_pad + f) * 8), sv);
        const size_t col = (size_t)b.freq_map[f] * b.n_dir + b.dir_offset + c;
        load_jones(&b.jones[(size_t)a.class_row[0] * b.n_ufreq * b.n_dir + col], j0);
        for (int r = 0; r < a.nclass; ++r) {
            double jr[8], o[8];
            load_jones(&b.jones[(size_t)a.class_row[r] * b.n_ufreq * b.n_dir + col], jr);
            sandwich(jr, sv, j0, o);
#pragma unroll
            for (int k = 0; k < 8; ++k) m = fmax(m, fabs(o[k]));
        }
    }
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

// FP32 complex product K C with the rounding fixed by explicit fused multiply-adds, so that
// every kernel organisation produces the same operands bit for bit.
__device__ __forceinline__ float cmul_re(float kr, float ki, float cr, float ci) {
    return __builtin_fmaf(kr, cr, -(ki * ci));
}
__device__ __forceinline__ float cmul_im(float kr, float ki, float cr, float ci) {
    return __builtin_fmaf(kr, ci, ki * cr);
}

// A thread's operands for one step: K_p and K_q of its 4 components (FP32 complex pairs)
// and their scaled coherencies for the work-group's two Jones entries.
struct Ops {
    float4 p[2], q[2], c[4];
};

// LDS operands of one step: A (hi, lo) for the two Jones entries, B (hi, lo) with the V_re
// columns 0..63 and V_im columns 64..127. Two such buffers alternate between steps.
template <class E> struct StepLds {
    E a[2][2][kFB * kFK];   // [hi / lo][entry]
    E b[2][2 * kFB * kFK];  // [hi / lo]
};

// Operand generation for one step from registers into one LDS buffer. Part 0 writes entry 0
// of A, part 1 entry 1 of A and the B columns, so each half can ride along one K slice of
// MFMAs. Thread t handles tile t / 4 of each block side and components 4 (t % 4) .. + 3.
template <class E>
__device__ __forceinline__ void generate(const Ops &g, StepLds<E> &L, int part, int gr, int gq) {
    const float kr[4] = {g.p[0].x, g.p[0].z, g.p[1].x, g.p[1].z};
    const float ki[4] = {g.p[0].y, g.p[0].w, g.p[1].y, g.p[1].w};
    {
        const int e = part;
        uint4 h, l;
        unsigned *hw = &h.x, *lw = &l.x;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float cr = e ? g.c[j].z : g.c[j].x;
            const float ci = e ? g.c[j].w : g.c[j].y;
            split_words<E>(cmul_re(kr[j], ki[j], cr, ci), cmul_im(kr[j], ki[j], cr, ci), hw[j], lw[j]);
        }
        *reinterpret_cast<uint4 *>(&L.a[0][e][lds_at(gr, 2 * gq)]) = h;
        *reinterpret_cast<uint4 *>(&L.a[1][e][lds_at(gr, 2 * gq)]) = l;
    }
    if (part == 0) return;
    // Column gr carries K_q: rows (2c, 2c+1) hold (K_re, K_im) for V_re and (-K_im, K_re)
    // for V_im.
    const float qr[4] = {g.q[0].x, g.q[0].z, g.q[1].x, g.q[1].z};
    const float qi[4] = {g.q[0].y, g.q[0].w, g.q[1].y, g.q[1].w};
    uint4 h0, l0, h1, l1;
    unsigned *h0w = &h0.x, *l0w = &l0.x, *h1w = &h1.x, *l1w = &l1.x;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        split_words<E>(qr[j], qi[j], h0w[j], l0w[j]);
        h1w[j] = swap_neg(h0w[j]);
        l1w[j] = swap_neg(l0w[j]);
    }
    *reinterpret_cast<uint4 *>(&L.b[0][lds_at(gr, 2 * gq)]) = h0;
    *reinterpret_cast<uint4 *>(&L.b[1][lds_at(gr, 2 * gq)]) = l0;
    *reinterpret_cast<uint4 *>(&L.b[0][lds_at(kFB + gr, 2 * gq)]) = h1;
    *reinterpret_cast<uint4 *>(&L.b[1][lds_at(kFB + gr, 2 * gq)]) = l1;
}

// Four waves: wave w computes local Jones entry w % 2 (entry 2 * half + w % 2) for output
// columns (w / 2) * 64 .. + 63 (V_re columns, then V_im columns), 2 x 2 tiles of 32 x 32.
// Each step s runs the MFMAs of step s from one LDS buffer while generating step s + 1 into
// the other, in one basic block with the MFMAs and the generation's ALU work interleaved,
// and a single barrier per step. DIAG: a diagonal block, whose output tile (rows 32..63,
// columns 0..31 of each half) holds only pairs p > q that no baseline reads; it is skipped.
template <class E, bool DIAG>
__device__ __forceinline__ void factored_block(const FactoredArgs &a, StepLds<E> (&buf)[2], int f, int blk,
                                               int half, int2 pq) {
    typedef typename Mfma<E>::vec V;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int le = wave & 1;
    const int colbase = (wave >> 1) * kFB;
    const int gr = tid >> 2;
    const int gq = (tid & 3) * 4;
    // Work-group-uniform bases with 32-bit per-thread offsets keep the addresses out of the
    // vector registers. Diagonal blocks load K_q from the same rows as K_p.
    const float2 *kf = a.kbuf + (size_t)f * a.tpad * a.cpad;
    const unsigned offp = (unsigned)((pq.x * kFB + gr) * a.cpad + gq);
    const unsigned offq = (unsigned)((pq.y * kFB + gr) * a.cpad + gq);
    // Scaled FP32 coherencies of the thread's 4 components: (re, im) of the work-group's first
    // Jones entry, then of its second, one float4 per component.
    const int rc = a.rclass != nullptr ? a.rclass[pq.x * kFB + gr] : 0;
    const float4 *cf = a.cohf + ((size_t)f * a.nclass + rc) * a.cpad * 2 + gq * 2 + half;
    auto load = [&](Ops &g, int step) {
        const unsigned o = (unsigned)step * kFC;
        g.p[0] = *reinterpret_cast<const float4 *>(kf + offp + o);
        g.p[1] = *reinterpret_cast<const float4 *>(kf + offp + o + 2);
        g.q[0] = *reinterpret_cast<const float4 *>(kf + offq + o);
        g.q[1] = *reinterpret_cast<const float4 *>(kf + offq + o + 2);
#pragma unroll
        for (int j = 0; j < 4; ++j) g.c[j] = cf[2 * (o + j)];
    };

    f16v acc[2][2], cor[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = cor[i][j][r] = 0.0f;

    const int nsteps = a.cpad / kFC;
    const int row = lane & 31;
    const int kofs = (lane >> 5) * 8;
    constexpr bool kBfLoLo = std::is_same<E, __bf16>::value;

    Ops g;
    load(g, 0);
    generate<E>(g, buf[0], 0, gr, gq);
    generate<E>(g, buf[0], 1, gr, gq);
    load(g, nsteps > 1 ? 1 : 0);
    for (int step = 0; step < nsteps; ++step) {
        const int cur = step & 1;
        __syncthreads();
        auto kslice = [&](auto kbc) {
            constexpr int kb = decltype(kbc)::value;
            V ahi[2], alo[2], bhi[2], blo[2];
#pragma unroll
            for (int t2 = 0; t2 < 2; ++t2) {
                const int ia = lds_at(t2 * 32 + row, kb * 16 + kofs);
                const int ib = lds_at(colbase + t2 * 32 + row, kb * 16 + kofs);
                ahi[t2] = *reinterpret_cast<const V *>(&buf[cur].a[0][le][ia]);
                alo[t2] = *reinterpret_cast<const V *>(&buf[cur].a[1][le][ia]);
                bhi[t2] = *reinterpret_cast<const V *>(&buf[cur].b[0][ib]);
                blo[t2] = *reinterpret_cast<const V *>(&buf[cur].b[1][ib]);
            }
            // Product rounds over the tiles, so that MFMAs writing the same accumulator are
            // never back to back: hi*hi, hi*lo, lo*hi (and, for bf16, whose halves carry 8 bits
            // each, lo*lo at 2^-16 of the product).
#pragma unroll
            for (int round = 0; round < (kBfLoLo ? 4 : 3); ++round)
#pragma unroll
                for (int rt = 0; rt < 2; ++rt)
#pragma unroll
                    for (int ct = 0; ct < 2; ++ct) {
                        if (DIAG && rt == 1 && ct == 0) continue;
                        if (round == 0) acc[rt][ct] = Mfma<E>::run(ahi[rt], bhi[ct], acc[rt][ct]);
                        if (round == 1) cor[rt][ct] = Mfma<E>::run(ahi[rt], blo[ct], cor[rt][ct]);
                        if (round == 2) cor[rt][ct] = Mfma<E>::run(alo[rt], bhi[ct], cor[rt][ct]);
                        if (round == 3) cor[rt][ct] = Mfma<E>::run(alo[rt], blo[ct], cor[rt][ct]);
                    }
            // The next step's operands into the other buffer while these MFMAs run (after the
            // last step this fills a buffer nobody reads; no branch splits the block).
            generate<E>(g, buf[cur ^ 1], kb, gr, gq);
            // Issue order: the fragment reads, then each MFMA followed by a share of the
            // generation's ALU work, then the generation's LDS writes.
            constexpr int kM = (DIAG ? 3 : 4) * (kBfLoLo ? 4 : 3);
            constexpr int kV = kb == 0 ? 4 : 8;
            __builtin_amdgcn_sched_group_barrier(0x100, 8, 0);
#pragma unroll
            for (int i = 0; i < kM; ++i) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x002, kV, 0);
            }
            __builtin_amdgcn_sched_group_barrier(0x200, kb == 0 ? 2 : 6, 0);
        };
        kslice(std::integral_constant<int, 0>{});
        kslice(std::integral_constant<int, 1>{});
        load(g, step + 2 < nsteps ? step + 2 : nsteps - 1);
        if ((step + 1) % a.seg_steps == 0 || step + 1 == nsteps) {
            // Output layout of v_mfma_f32_32x32x16: element r of lane l is row
            // 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32. The work-group is the only
            // writer of its region, so the FP64 adds go out as atomics that return nothing
            // and need no wait. The empty asm keeps the compiler from hoisting 64 addresses
            // out of the step loop.
            double *out = a.out + ((size_t)(f * a.nblocks + blk) * 4 + 2 * half) * (kFB * 2 * kFB);
            unsigned lo = (unsigned)(le * (kFB * 2 * kFB) + (4 * (lane >> 5)) * (2 * kFB) + colbase + (lane & 31));
            asm volatile("" : "+v"(lo));
#pragma unroll
            for (int rt = 0; rt < 2; ++rt)
#pragma unroll
                for (int ct = 0; ct < 2; ++ct) {
                    if (DIAG && rt == 1 && ct == 0) continue;
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const int o = (rt * 32 + 8 * (r >> 2) + (r & 3)) * (2 * kFB) + ct * 32;
                        unsafeAtomicAdd(&out[lo + o], (double)acc[rt][ct][r] + (double)cor[rt][ct][r]);
                        acc[rt][ct][r] = 0.0f;
                        cor[rt][ct][r] = 0.0f;
                        if ((r & 3) == 3) __builtin_amdgcn_sched_barrier(0);
                    }
                }
        }
    }
}

template <class E>
__global__ __launch_bounds__(256, 2) void k_factored(const FactoredArgs a) {
    __shared__ __attribute__((aligned(16))) StepLds<E> buf[2];
    const int total = a.nchan * a.nblocks * 2;
    int b = blockIdx.x;
    if ((total & 7) == 0) b = (b & 7) * (total >> 3) + (b >> 3);
    const int f = b / (2 * a.nblocks);
    const int rem = b - f * 2 * a.nblocks;
    const int blk = rem >> 1;
    const int half = rem & 1;
    const int2 pq = a.blocks[blk];
    if (pq.x == pq.y)
        factored_block<E, true>(a, buf, f, blk, half, pq);
    else
        factored_block<E, false>(a, buf, f, blk, half, pq);
}

// ---------------------------------------------------------------- wave-specialised variant
// Four consumer waves (one per SIMD) run only the MFMAs of step s from one LDS buffer while
// 8 producer waves generate step s + 1 into the other; one barrier per step. Each SIMD thus
// always holds a wave with MFMAs to issue and two with ALU work, instead of relying on two
// general waves to drift out of phase. Register use is the larger of the two roles.

// A producer thread's operands for one step: 2 components.
struct Ops2 {
    float4 p, q, c[2];
};

template <class E>
__device__ __forceinline__ void generate2(const Ops2 &g, StepLds<E> &L, int gr, int g2) {
    const float kr[2] = {g.p.x, g.p.z};
    const float ki[2] = {g.p.y, g.p.w};
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        uint2 h, l;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const float cr = e ? g.c[j].z : g.c[j].x;
            const float ci = e ? g.c[j].w : g.c[j].y;
            split_words<E>(cmul_re(kr[j], ki[j], cr, ci), cmul_im(kr[j], ki[j], cr, ci), j ? h.y : h.x,
                           j ? l.y : l.x);
        }
        *reinterpret_cast<uint2 *>(&L.a[0][e][lds_at(gr, 2 * g2)]) = h;
        *reinterpret_cast<uint2 *>(&L.a[1][e][lds_at(gr, 2 * g2)]) = l;
    }
    uint2 h0, l0;
    split_words<E>(g.q.x, g.q.y, h0.x, l0.x);
    split_words<E>(g.q.z, g.q.w, h0.y, l0.y);
    const uint2 h1{swap_neg(h0.x), swap_neg(h0.y)};
    const uint2 l1{swap_neg(l0.x), swap_neg(l0.y)};
    *reinterpret_cast<uint2 *>(&L.b[0][lds_at(gr, 2 * g2)]) = h0;
    *reinterpret_cast<uint2 *>(&L.b[1][lds_at(gr, 2 * g2)]) = l0;
    *reinterpret_cast<uint2 *>(&L.b[0][lds_at(kFB + gr, 2 * g2)]) = h1;
    *reinterpret_cast<uint2 *>(&L.b[1][lds_at(kFB + gr, 2 * g2)]) = l1;
}

// Fragments of one 16-deep K slice for a consumer wave: A rows 0..31 and 32..63 of its Jones
// entry and B columns colbase + 0..31 and + 32..63, hi and lo halves.
template <class E> struct Frags {
    typename Mfma<E>::vec ahi[2], alo[2], bhi[2], blo[2];
};

template <class E>
__device__ __forceinline__ void read_frags(Frags<E> &F, const StepLds<E> &L, int kb, int le, int colbase, int row,
                                           int kofs) {
    typedef typename Mfma<E>::vec V;
#pragma unroll
    for (int t2 = 0; t2 < 2; ++t2) {
        const int ia = lds_at(t2 * 32 + row, kb * 16 + kofs);
        const int ib = lds_at(colbase + t2 * 32 + row, kb * 16 + kofs);
        F.ahi[t2] = *reinterpret_cast<const V *>(&L.a[0][le][ia]);
        F.alo[t2] = *reinterpret_cast<const V *>(&L.a[1][le][ia]);
        F.bhi[t2] = *reinterpret_cast<const V *>(&L.b[0][ib]);
        F.blo[t2] = *reinterpret_cast<const V *>(&L.b[1][ib]);
    }
}

// Product rounds over the 2 x 2 tiles (hi*hi into acc; hi*lo, lo*hi and, for bf16, lo*lo
// into cor), so that MFMAs writing the same accumulator are never back to back.
template <class E, bool DIAG>
__device__ __forceinline__ void mfma_rounds(const Frags<E> &F, f16v (&acc)[2][2], f16v (&cor)[2][2]) {
    constexpr bool kBfLoLo = std::is_same<E, __bf16>::value;
#pragma unroll
    for (int round = 0; round < (kBfLoLo ? 4 : 3); ++round)
#pragma unroll
        for (int rt = 0; rt < 2; ++rt)
#pragma unroll
            for (int ct = 0; ct < 2; ++ct) {
                if (DIAG && rt == 1 && ct == 0) continue;
                if (round == 0) acc[rt][ct] = Mfma<E>::run(F.ahi[rt], F.bhi[ct], acc[rt][ct]);
                if (round == 1) cor[rt][ct] = Mfma<E>::run(F.ahi[rt], F.blo[ct], cor[rt][ct]);
                if (round == 2) cor[rt][ct] = Mfma<E>::run(F.alo[rt], F.bhi[ct], cor[rt][ct]);
                if (round == 3) cor[rt][ct] = Mfma<E>::run(F.alo[rt], F.blo[ct], cor[rt][ct]);
            }
}

// Wave-specialised block: 4 consumer waves run the MFMAs of step s from one LDS buffer while 8
// producer waves generate step s + 1 into the other. Producer thread t handles tile t / 8 of
// each block side and components 2 (t % 8), 2 (t % 8) + 1 of a step. Its phasor and coherency
// loads are issued D steps ahead, in D register sets used in turn, since each step's loads
// miss in the caches.
template <class E, bool DIAG, int D>
__device__ __forceinline__ void factored_ws_block(const FactoredArgs &a, StepLds<E> (&buf)[2], int f, int blk,
                                                  int half, int2 pq) {
    static_assert(D >= 2 && D % 2 == 0, "an even number of register sets");
    static_assert(kFC == 16 && kFB == 64, "512 producer threads cover 64 rows x 16 components");
    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    const int nsteps = a.cpad / kFC;
    if (wave >= 4) {
        const int pt = tid - 256;
        const int gr = pt >> 3;
        const int gc = (pt & 7) * 2;
        const float2 *kf = a.kbuf + (size_t)f * a.tpad * a.cpad;
        const unsigned offp = (unsigned)((pq.x * kFB + gr) * a.cpad + gc);
        const unsigned offq = (unsigned)((pq.y * kFB + gr) * a.cpad + gc);
        const int rc = a.rclass != nullptr ? a.rclass[pq.x * kFB + gr] : 0;
        const float4 *cf = a.cohf + ((size_t)f * a.nclass + rc) * a.cpad * 2 + gc * 2 + half;
        // Loads past the last step are clamped to it; they fill buffers no consumer reads.
        auto load = [&](Ops2 &g, int step) {
            step = step < nsteps ? step : nsteps - 1;
            const unsigned o = (unsigned)step * kFC;
            g.p = *reinterpret_cast<const float4 *>(kf + offp + o);
            g.q = *reinterpret_cast<const float4 *>(kf + offq + o);
            g.c[0] = cf[2 * o];
            g.c[1] = cf[2 * (o + 1)];
        };
        // Step 0 goes in before the first barrier; after barrier s the producers write step
        // s + 1 into buffer (s + 1) % 2 from register set (s + 1) % D, then reload that set
        // with step s + 1 + D.
        Ops2 g[D];
#pragma unroll
        for (int j = 0; j < D; ++j) load(g[j], j);
        generate2<E>(g[0], buf[0], gr, gc);
        load(g[0], D);
        for (int s = 0; s < nsteps; s += D) {
#pragma unroll
            for (int d = 0; d < D; ++d) {
                if (s + d >= nsteps) break;
                __syncthreads();
                const int j = (d + 1) % D;
                generate2<E>(g[j], buf[(d + 1) & 1], gr, gc);
                load(g[j], s + d + 1 + D);
            }
        }
        return;
    }
    // Consumers: wave w computes local Jones entry w % 2 for output columns (w / 2) * 64 ..
    // + 63 (V_re columns, then V_im columns), 2 x 2 tiles of 32 x 32.
    const int lane = tid & 63;
    const int le = wave & 1;
    const int colbase = (wave >> 1) * kFB;
    const int row = lane & 31;
    const int kofs = (lane >> 5) * 8;
    f16v acc[2][2], cor[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = cor[i][j][r] = 0.0f;
    Frags<E> F;
    for (int step = 0; step < nsteps; ++step) {
        __syncthreads();
#pragma unroll
        for (int kb = 0; kb < kFK / 16; ++kb) {
            read_frags<E>(F, buf[step & 1], kb, le, colbase, row, kofs);
            mfma_rounds<E, DIAG>(F, acc, cor);
        }
        if ((step + 1) % a.seg_steps == 0 || step + 1 == nsteps) {
            // Output layout of v_mfma_f32_32x32x16: element r of lane l is row
            // 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32; no-return FP64 atomics, the
            // work-group being the only writer of its region.
            double *out = a.out + ((size_t)(f * a.nblocks + blk) * 4 + 2 * half) * (kFB * 2 * kFB);
            unsigned lo = (unsigned)(le * (kFB * 2 * kFB) + (4 * (lane >> 5)) * (2 * kFB) + colbase + (lane & 31));
            asm volatile("" : "+v"(lo));
#pragma unroll
            for (int rt = 0; rt < 2; ++rt)
#pragma unroll
                for (int ct = 0; ct < 2; ++ct) {
                    if (DIAG && rt == 1 && ct == 0) continue;
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const int o = (rt * 32 + 8 * (r >> 2) + (r & 3)) * (2 * kFB) + ct * 32;
                        unsafeAtomicAdd(&out[lo + o], (double)acc[rt][ct][r] + (double)cor[rt][ct][r]);
                        acc[rt][ct][r] = 0.0f;
                        cor[rt][ct][r] = 0.0f;
                        if ((r & 3) == 3) __builtin_amdgcn_sched_barrier(0);
                    }
                }
        }
    }
}

template <class E, int D>
__global__ __launch_bounds__(768) void k_factored_ws(const FactoredArgs a) {
    __shared__ __attribute__((aligned(16))) StepLds<E> buf[2];
    const int total = a.nchan * a.nblocks * 2;
    int b = blockIdx.x;
    if ((total & 7) == 0) b = (b & 7) * (total >> 3) + (b >> 3);
    const int f = b / (2 * a.nblocks);
    const int rem = b - f * 2 * a.nblocks;
    const int blk = rem >> 1;
    const int half = rem & 1;
    const int2 pq = a.blocks[blk];
    if (pq.x == pq.y)
        factored_ws_block<E, true, D>(a, buf, f, blk, half, pq);
    else
        factored_ws_block<E, false, D>(a, buf, f, blk, half, pq);
}

// ---------------------------------------------------------------- K operand from planes
// Kernel 2: the phasor kernel stores K already split into hi and lo halves of the MFMA type
// (khi, klo), so the K operand is the same for all six work-groups of a channel and is not
// regenerated by each. Four MFMA waves read their K fragments straight from those planes
// (one 16-byte load per fragment, one step ahead), and four producer waves generate only the
// A operand into LDS: half the LDS traffic of kernel 1 per step. Operands, products and their
// order are those of kernels 0 and 1, so the results agree bit for bit.

template <class E> struct StepLdsA {
    E a[2][2][kFB * kFK];  // [hi / lo][entry]
};

// A producer thread's operands for one step: K_p of 4 components and their coherencies.
struct OpsA {
    float4 p[2], c[4];
};

template <class E>
__device__ __forceinline__ void generate_a(const OpsA &g, StepLdsA<E> &L, int gr, int gq) {
    const float kr[4] = {g.p[0].x, g.p[0].z, g.p[1].x, g.p[1].z};
    const float ki[4] = {g.p[0].y, g.p[0].w, g.p[1].y, g.p[1].w};
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        uint4 h, l;
        unsigned *hw = &h.x, *lw = &l.x;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float cr = e ? g.c[j].z : g.c[j].x;
            const float ci = e ? g.c[j].w : g.c[j].y;
            split_words<E>(cmul_re(kr[j], ki[j], cr, ci), cmul_im(kr[j], ki[j], cr, ci), hw[j], lw[j]);
        }
        *reinterpret_cast<uint4 *>(&L.a[0][e][lds_at(gr, 2 * gq)]) = h;
        *reinterpret_cast<uint4 *>(&L.a[1][e][lds_at(gr, 2 * gq)]) = l;
    }
}

// A step's K fragments of one MFMA wave: [K slice][column tile] hi and lo (8 values each).
struct KFrags {
    uint4 hi[2][2], lo[2][2];
};

template <class E, bool DIAG, int D>
__device__ __forceinline__ void factored_gb_block(const FactoredArgs &a, StepLdsA<E> (&buf)[2], int f, int blk,
                                                  int half, int2 pq) {
    typedef typename Mfma<E>::vec V;
    static_assert(D >= 2 && D % 2 == 0, "an even number of register sets");
    static_assert(kFC == 16 && kFB == 64, "256 producer threads cover 64 rows x 16 components");
    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    const int nsteps = a.cpad / kFC;
    if (wave >= 4) {
        // Producers: thread pt handles row pt / 4 and components 4 (pt % 4) .. + 3 of a step.
        const int pt = tid - 256;
        const int gr = pt >> 2;
        const int gq = (pt & 3) * 4;
        const float2 *kf = a.kbuf + (size_t)f * a.tpad * a.cpad;
        const unsigned offp = (unsigned)((pq.x * kFB + gr) * a.cpad + gq);
        const int rc = a.rclass != nullptr ? a.rclass[pq.x * kFB + gr] : 0;
        const float4 *cf = a.cohf + ((size_t)f * a.nclass + rc) * a.cpad * 2 + gq * 2 + half;
        auto load = [&](OpsA &g, int step) {
            step = step < nsteps ? step : nsteps - 1;
            const unsigned o = (unsigned)step * kFC;
            g.p[0] = *reinterpret_cast<const float4 *>(kf + offp + o);
            g.p[1] = *reinterpret_cast<const float4 *>(kf + offp + o + 2);
#pragma unroll
            for (int j = 0; j < 4; ++j) g.c[j] = cf[2 * (o + j)];
        };
        OpsA g[D];
#pragma unroll
        for (int j = 0; j < D; ++j) load(g[j], j);
        generate_a<E>(g[0], buf[0], gr, gq);
        load(g[0], D);
        for (int s = 0; s < nsteps; s += D) {
#pragma unroll
            for (int d = 0; d < D; ++d) {
                if (s + d >= nsteps) break;
                __syncthreads();
                const int j = (d + 1) % D;
                generate_a<E>(g[j], buf[(d + 1) & 1], gr, gq);
                load(g[j], s + d + 1 + D);
            }
        }
        return;
    }
    // MFMA waves: wave w computes local Jones entry w % 2 for output columns (w / 2) * 64 ..
    // + 63 (V_re columns carry K, V_im columns (-K_im, K_re)), 2 x 2 tiles of 32 x 32.
    const int lane = tid & 63;
    const int le = wave & 1;
    const bool vim = (wave >> 1) != 0;
    const int colbase = (wave >> 1) * kFB;
    const int row = lane & 31;
    const int kofs = (lane >> 5) * 8;
    // K fragment of lane l: column tile t2 holds tile q = pq.y * 64 + t2 * 32 + l % 32; the
    // 8 values are components 4 (l / 32) .. + 3 of the K slice, (re, im) each.
    const size_t pf = (size_t)f * a.tpad * a.cpad;
    const uint4 *khi = reinterpret_cast<const uint4 *>(a.khi + pf);
    const uint4 *klo = reinterpret_cast<const uint4 *>(a.klo + pf);
    unsigned qoff[2];
#pragma unroll
    for (int t2 = 0; t2 < 2; ++t2)
        qoff[t2] = (unsigned)(((pq.y * kFB + t2 * 32 + row) * a.cpad + (lane >> 5) * 4) / 4);
    auto load_k = [&](KFrags &K, int step) {
        step = step < nsteps ? step : nsteps - 1;
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
#pragma unroll
            for (int t2 = 0; t2 < 2; ++t2) {
                const unsigned o = qoff[t2] + (unsigned)(step * kFC + kb * 8) / 4;
                K.hi[kb][t2] = khi[o];
                K.lo[kb][t2] = klo[o];
            }
    };
    auto as_v = [&](uint4 w) {
        if (vim) {
            w.x = swap_neg(w.x);
            w.y = swap_neg(w.y);
            w.z = swap_neg(w.z);
            w.w = swap_neg(w.w);
        }
        return __builtin_bit_cast(V, w);
    };
    f16v acc[2][2], cor[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = cor[i][j][r] = 0.0f;
    // One step: the MFMA rounds on the A operand in LDS and the K fragments in Kc, then the
    // FP64 flush at the end of a segment.
    auto run_step = [&](const KFrags &Kc, int step) {
#pragma unroll
        for (int kb = 0; kb < kFK / 16; ++kb) {
            Frags<E> F;
#pragma unroll
            for (int t2 = 0; t2 < 2; ++t2) {
                const int ia = lds_at(t2 * 32 + row, kb * 16 + kofs);
                F.ahi[t2] = *reinterpret_cast<const V *>(&buf[step & 1].a[0][le][ia]);
                F.alo[t2] = *reinterpret_cast<const V *>(&buf[step & 1].a[1][le][ia]);
                F.bhi[t2] = as_v(Kc.hi[kb][t2]);
                F.blo[t2] = as_v(Kc.lo[kb][t2]);
            }
            mfma_rounds<E, DIAG>(F, acc, cor);
        }
        if ((step + 1) % a.seg_steps == 0 || step + 1 == nsteps) {
            double *out = a.out + ((size_t)(f * a.nblocks + blk) * 4 + 2 * half) * (kFB * 2 * kFB);
            unsigned lo = (unsigned)(le * (kFB * 2 * kFB) + (4 * (lane >> 5)) * (2 * kFB) + colbase + (lane & 31));
            asm volatile("" : "+v"(lo));
#pragma unroll
            for (int rt = 0; rt < 2; ++rt)
#pragma unroll
                for (int ct = 0; ct < 2; ++ct) {
                    if (DIAG && rt == 1 && ct == 0) continue;
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const int o = (rt * 32 + 8 * (r >> 2) + (r & 3)) * (2 * kFB) + ct * 32;
                        unsafeAtomicAdd(&out[lo + o], (double)acc[rt][ct][r] + (double)cor[rt][ct][r]);
                        acc[rt][ct][r] = 0.0f;
                        cor[rt][ct][r] = 0.0f;
                        if ((r & 3) == 3) __builtin_amdgcn_sched_barrier(0);
                    }
                }
        }
    };
    // Two named register sets, the loop unrolled by two: a set indexed by the step's parity
    // would be a dynamically indexed private array, which lives in scratch memory. One barrier
    // per step, as the producer waves expect.
    KFrags K0, K1;
    load_k(K0, 0);
    for (int step = 0; step < nsteps; step += 2) {
        load_k(K1, step + 1);
        __syncthreads();
        run_step(K0, step);
        if (step + 1 >= nsteps) break;
        load_k(K0, step + 2);
        __syncthreads();
        run_step(K1, step + 1);
    }
}

template <class E, int D>
__global__ __launch_bounds__(512) void k_factored_gb(const FactoredArgs a) {
    __shared__ __attribute__((aligned(16))) StepLdsA<E> buf[2];
    const int total = a.nchan * a.nblocks * 2;
    int b = blockIdx.x;
    if ((total & 7) == 0) b = (b & 7) * (total >> 3) + (b >> 3);
    const int f = b / (2 * a.nblocks);
    const int rem = b - f * 2 * a.nblocks;
    const int blk = rem >> 1;
    const int half = rem & 1;
    const int2 pq = a.blocks[blk];
    if (pq.x == pq.y)
        factored_gb_block<E, true, D>(a, buf, f, blk, half, pq);
    else
        factored_gb_block<E, false, D>(a, buf, f, blk, half, pq);
}

// FP64 block results -> visibilities [freq][baseline] (+=), undoing the scale. map[i] =
// (baseline, p, q): the output baseline and its two tiles in the factored tile numbering.
__global__ void k_factored_scatter(const FactoredArgs a, const int4 *__restrict__ map, int nmap, int nbl,
                                   const int *__restrict__ block_of, JonesF32 *vis32, JonesF64 *vis64) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)a.nchan * nmap) return;
    const int f = (int)(idx / nmap);
    const int4 e4 = map[idx % nmap];
    const int gb = (a.ntiles + kFB - 1) / kFB;
    const int blk = block_of[(e4.y / kFB) * gb + (e4.z / kFB)];
    const int pl = e4.y % kFB, ql = e4.z % kFB;
    const double inv = 1.0 / *a.scale;
    // e4.w = 1: the baseline is the conjugate transpose of block result (p, q): entry (i, j)
    // comes from entry (j, i), conjugated.
    const bool herm = e4.w != 0;
    double v[8];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int es = herm ? ((e & 1) << 1) | (e >> 1) : e;
        const double *o = a.out + ((size_t)(f * a.nblocks + blk) * 4 + es) * (kFB * 2 * kFB) + pl * (2 * kFB);
        v[2 * e] = o[ql] * inv;
        v[2 * e + 1] = (herm ? -o[kFB + ql] : o[kFB + ql]) * inv;
    }
    const size_t vo = (size_t)f * nbl + e4.x;
    float *v32 = reinterpret_cast<float *>(&vis32[vo]);
#pragma unroll
    for (int e = 0; e < 8; ++e) v32[e] += (float)v[e];
    if (vis64 != nullptr) {
        double *v64 = reinterpret_cast<double *>(&vis64[vo]);
#pragma unroll
        for (int e = 0; e < 8; ++e) v64[e] += v[e];
    }
}

// dst[i] = src[idx[i]] for rows of 3 doubles (tile UVWs of a tile subset); idx[i] < 0 gives
// zeros (padding tiles).
__global__ void k_gather_rows3(int n, const int *__restrict__ idx, const double *__restrict__ src, double *dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int t = idx[i];
    dst[3 * i] = t >= 0 ? src[3 * t] : 0.0;
    dst[3 * i + 1] = t >= 0 ? src[3 * t + 1] : 0.0;
    dst[3 * i + 2] = t >= 0 ? src[3 * t + 2] : 0.0;
}

hipError_t launch_gather_rows3(int n, const int *d_idx, const double *d_src, double *d_dst, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_rows3, dim3((n + 255) / 256), dim3(256), 0, stream, n, d_idx, d_src, d_dst);
    return hipGetLastError();
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

int g_factored_kernel = 1;

int factored_chunk(int nchan, int ntiles, int ncomp, int seg_steps) {
    const size_t tpad = (size_t)((ntiles + kFB - 1) / kFB) * kFB;
    const size_t per_comp = (size_t)nchan * tpad * sizeof(float2);
    const int unit = kFC * std::max(seg_steps, 1);
    size_t fit = kKbufCap / std::max<size_t>(per_comp, 1);
    fit = std::max<size_t>(unit, fit / unit * unit);
    const int all = (ncomp + kFC - 1) / kFC * kFC;
    return (int)std::min<size_t>(fit, (size_t)std::max(all, kFC));
}

size_t factored_kbuf_bytes(int nchan, int ntiles, int chunk) {
    const size_t tpad = (size_t)((ntiles + kFB - 1) / kFB) * kFB;
    return (size_t)nchan * tpad * chunk * sizeof(float2);
}

size_t factored_cohf_bytes(int nchan, int chunk, int nclass) {
    return (size_t)nchan * nclass * chunk * 8 * sizeof(float);
}

hipError_t launch_factored(int dtype, FactoredArgs a, int chunk, const double *coh_for_scale, size_t coh_n,
                           unsigned long long *d_max_bits, double *d_scale, const int4 *map, int nmap, int nbl,
                           const int *block_of, JonesF32 *vis32, JonesF64 *vis64, hipStream_t stream,
                           hipEvent_t ev0, hipEvent_t ev1) {
    if (a.ncomp <= 0 || a.nchan <= 0 || nmap <= 0) return hipSuccess;
    if (chunk <= 0 || chunk % kFC != 0) return hipErrorInvalidValue;
    a.tpad = (a.ntiles + kFB - 1) / kFB * kFB;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, stream, d_max_bits);
    if (a.nclass > 1)
        hipLaunchKernelGGL(k_absmax_classes, dim3(1024), dim3(256), 0, stream, a, d_max_bits);
    else
        hipLaunchKernelGGL(k_absmax, dim3(1024), dim3(256), 0, stream, coh_n, coh_for_scale, d_max_bits);
    hipLaunchKernelGGL(k_scale_from_max, dim3(1), dim3(1), 0, stream, d_max_bits, d_scale);
    a.scale = d_scale;
    const size_t nout = factored_out_doubles(a.nchan, a.nblocks);
    hipLaunchKernelGGL(k_fill, dim3(2048), dim3(256), 0, stream, nout, a.out);
    if (ev0 != nullptr) (void)hipEventRecord(ev0, stream);
    const int total = a.nchan * a.nblocks * 2;
    for (int c0 = 0; c0 < a.ncomp; c0 += chunk) {
        a.c_begin = c0;
        a.c_count = std::min(chunk, a.ncomp - c0);
        a.cpad = (a.c_count + kFC - 1) / kFC * kFC;
        const size_t nph = (size_t)a.tpad * a.cpad;
        hipLaunchKernelGGL(k_factored_phasors, dim3((unsigned)((nph + 255) / 256)), dim3(256), 0, stream, a);
        const size_t ncf = (size_t)a.nchan * a.cpad * 8;
        if (a.nclass > 1) {
            const size_t ncl = (size_t)a.nchan * a.nclass * a.cpad;
            hipLaunchKernelGGL(k_factored_coh32_classes, dim3((unsigned)((ncl + 255) / 256)), dim3(256), 0, stream, a);
        } else {
            hipLaunchKernelGGL(k_factored_coh32, dim3((unsigned)((ncf + 255) / 256)), dim3(256), 0, stream, a);
        }
        if (g_factored_kernel == 2) {
            if (a.khi == nullptr || a.klo == nullptr || a.planes_dtype != dtype) return hipErrorInvalidValue;
            if (dtype == SM_MFMA_BF16)
                hipLaunchKernelGGL((k_factored_gb<__bf16, 4>), dim3(total), dim3(512), 0, stream, a);
            else
                hipLaunchKernelGGL((k_factored_gb<_Float16, 4>), dim3(total), dim3(512), 0, stream, a);
        } else if (g_factored_kernel == 1) {
            if (dtype == SM_MFMA_BF16)
                hipLaunchKernelGGL((k_factored_ws<__bf16, 4>), dim3(total), dim3(768), 0, stream, a);
            else
                hipLaunchKernelGGL((k_factored_ws<_Float16, 4>), dim3(total), dim3(768), 0, stream, a);
        } else if (dtype == SM_MFMA_BF16) {
            hipLaunchKernelGGL(k_factored<__bf16>, dim3(total), dim3(256), 0, stream, a);
        } else {
            hipLaunchKernelGGL(k_factored<_Float16>, dim3(total), dim3(256), 0, stream, a);
        }
    }
    if (ev1 != nullptr) (void)hipEventRecord(ev1, stream);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const size_t n = (size_t)a.nchan * nmap;
    hipLaunchKernelGGL(k_factored_scatter, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, a, map, nmap, nbl,
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
