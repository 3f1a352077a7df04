// FP64 direct measurement-equation kernels for gfx950.
//
// V_pq(f) += sum_c (J_p S_c(f) J_q^H) e^{i (u L + v M + w N) f/c} [x envelope]
// (src/model/cpu.rs:337-345,472-487,702-717; src/gpu/model.cu:99-366).
//
// Mapping: one lane per baseline, NCH consecutive channels per lane (kDirectNch), the component
// loop serial per lane. Everything that depends only on (component, channel) is
// wave-uniform: the coherency J S J^H (hoisted out of the baseline loop when every tile
// shares one beam), LMN, envelope parameters and shapelet coefficients arrive through
// scalar loads and SGPR operands, so the vector ALUs do only per-baseline work. Across a
// chunk of evenly spaced channels the phase e^{i theta f_k / c} advances by a fixed
// rotation, so two FP64 sincos per (baseline, component, chunk) replace NCH of them;
// the Gaussian envelope exp(q f_k^2) advances by a geometric recurrence in the same way.
// Blocks are remapped so that all baseline blocks of a channel chunk land on the same XCD
// and share that XCD's L2 copy of the chunk's coherency rows.
#include <algorithm>

#include "sm_internal.h"
#include "sm_jones.h"
#include "sm_math.h"

namespace sm {

// ------------------------------------------------------------------ coherency table
// coh[c][f] = J S J^H (one shared beam row) or S; S from the power-law / curved-power-law
// parameters at 150 MHz or the list table (src/gpu/model.cu:16-28,137-167). With `hermitian`,
// 4 values per entry instead of 8: Re XX, Re XY, Im XY, Re YY (the tile-pair kernels' form;
// J S J^H with the same J on both sides is Hermitian, as S is).
__global__ void k_coherency(int ncomp, int nchan, int nchan_pad, const double *__restrict__ freqs, FluxBuckets fb,
                            BeamDesc beam, int apply_beam, int hermitian, double *__restrict__ coh) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)ncomp * nchan_pad) return;
    const int c = (int)(idx / nchan_pad);
    const int f = (int)(idx % nchan_pad);
    double *out = coh + idx * (hermitian ? 4 : 8);
    if (f >= nchan) {
        for (int e = 0; e < (hermitian ? 4 : 8); ++e) out[e] = 0.0;
        return;
    }
    const double freq = freqs[f];
    double s[8];
    if (c < fb.n_pl) {
        double fd[8];
        load_jones(&fb.pl_fds[c], fd);
        const double ratio = pow(freq / SM_POWER_LAW_FD_REF_FREQ, fb.pl_sis[c]);
#pragma unroll
        for (int e = 0; e < 8; ++e) s[e] = fd[e] * ratio;
    } else if (c < fb.n_pl + fb.n_cpl) {
        const int i = c - fb.n_pl;
        double fd[8];
        load_jones(&fb.cpl_fds[i], fd);
        const double ratio = pow(freq / SM_POWER_LAW_FD_REF_FREQ, fb.cpl_sis[i]);
        const double lt = log(freq / SM_POWER_LAW_FD_REF_FREQ);
        const double curved = exp(fb.cpl_qs[i] * (lt * lt));
#pragma unroll
        for (int e = 0; e < 8; ++e) s[e] = (fd[e] * ratio) * curved;
    } else {
        const int i = c - fb.n_pl - fb.n_cpl;
        load_jones(&fb.list_fds[(size_t)f * fb.n_list + i], s);
    }
    double o[8];
    if (apply_beam && beam.jones != nullptr) {
        double j[8];
        load_jones(&beam.jones[((size_t)beam.uniform_row * beam.n_ufreq + beam.freq_map[f]) * beam.n_dir + beam.dir_offset + c], j);
        sandwich(j, s, j, o);
    } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) o[e] = s[e];
    }
    if (hermitian) {
        out[0] = o[0];
        out[1] = o[2];
        out[2] = o[3];
        out[3] = o[6];
    } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) out[e] = o[e];
    }
}

__global__ void k_class_coherency(int nclass, int ncomp, int nchan, int nchan_pad, const double *__restrict__ s_tab,
                                  BeamDesc beam, const int *__restrict__ class_row, double *__restrict__ out) {
    // One thread per (component, channel), looping over the classes: S and J_0 are read once
    // per (component, channel) instead of once per class.
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)ncomp * nchan_pad) return;
    const int f = (int)(idx % nchan_pad);
    const int c = (int)(idx / nchan_pad);
    const size_t cstride = (size_t)ncomp * nchan_pad * 8;
    if (f >= nchan) {
        for (int r = 0; r < nclass; ++r) {
            double *o = out + r * cstride + idx * 8;
#pragma unroll
            for (int e = 0; e < 8; ++e) o[e] = 0.0;
        }
        return;
    }
    double sv[8], jb[8];
    load_jones(reinterpret_cast<const JonesF64 *>(s_tab + ((size_t)c * nchan_pad + f) * 8), sv);
    const size_t col = (size_t)beam.freq_map[f] * beam.n_dir + beam.dir_offset + c;
    load_jones(&beam.jones[(size_t)class_row[0] * beam.n_ufreq * beam.n_dir + col], jb);
    for (int r = 0; r < nclass; ++r) {
        double ja[8], v[8];
        load_jones(&beam.jones[(size_t)class_row[r] * beam.n_ufreq * beam.n_dir + col], ja);
        sandwich(ja, sv, jb, v);
        double *o = out + r * cstride + idx * 8;
#pragma unroll
        for (int e = 0; e < 8; ++e) o[e] = v[e];
    }
}

__global__ void k_pair_coherency(int ncomp, int nchan, int nchan_pad, const double *__restrict__ s_tab, BeamDesc beam,
                                 int row_a, int row_b, double *__restrict__ out) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)ncomp * nchan_pad) return;
    const int f = (int)(idx % nchan_pad);
    const int c = (int)(idx / nchan_pad);
    double *o = out + idx * 8;
    if (f >= nchan) {
#pragma unroll
        for (int e = 0; e < 8; ++e) o[e] = 0.0;
        return;
    }
    double sv[8], ja[8], jb[8], v[8];
    load_jones(reinterpret_cast<const JonesF64 *>(s_tab + idx * 8), sv);
    const size_t col = (size_t)beam.freq_map[f] * beam.n_dir + beam.dir_offset + c;
    load_jones(&beam.jones[(size_t)row_a * beam.n_ufreq * beam.n_dir + col], ja);
    load_jones(&beam.jones[(size_t)row_b * beam.n_ufreq * beam.n_dir + col], jb);
    sandwich(ja, sv, jb, v);
#pragma unroll
    for (int e = 0; e < 8; ++e) o[e] = v[e];
}

hipError_t launch_pair_coherency(int ncomp, int nchan, int nchan_pad, const double *d_s, const BeamDesc &beam,
                                 int row_a, int row_b, double *d_out, hipStream_t stream) {
    const size_t n = (size_t)ncomp * nchan_pad;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pair_coherency, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, ncomp, nchan,
                       nchan_pad, d_s, beam, row_a, row_b, d_out);
    return hipGetLastError();
}

hipError_t launch_class_coherency(int nclass, int ncomp, int nchan, int nchan_pad, const double *d_s,
                                  const BeamDesc &beam, const int *d_class_row, double *d_coh_cls,
                                  hipStream_t stream) {
    const size_t n = (size_t)ncomp * nchan_pad;
    if (n == 0 || nclass <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_class_coherency, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, nclass, ncomp,
                       nchan, nchan_pad, d_s, beam, d_class_row, d_coh_cls);
    return hipGetLastError();
}

hipError_t launch_coherency(int ncomp, int nchan, int nchan_pad, const double *d_freqs, FluxBuckets fb,
                            BeamDesc beam, bool apply_beam, double *d_coh, hipStream_t stream, bool hermitian) {
    if (ncomp <= 0) return hipSuccess;
    const size_t n = (size_t)ncomp * nchan_pad;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_coherency, dim3(blocks), dim3(256), 0, stream, ncomp, nchan, nchan_pad, d_freqs, fb, beam,
                       apply_beam ? 1 : 0, hermitian ? 1 : 0, d_coh);
    return hipGetLastError();
}

// ------------------------------------------------------------------ shapelet envelope
// sum_k i^{(n1+n2) mod 4} c_k B_{n1}(x_pos) B_{n2}(y_pos) with linear interpolation in the
// basis table (src/model/cpu.rs:633-695, src/gpu/model.cu:39-97). Positions below the
// table clamp to index 0 with the unclamped fractional weight, as both reference paths do;
// positions above it clamp to the last interval, whose values are exact zeros.
__device__ __forceinline__ void shapelet_envelope(double x_pos, double y_pos, int4 range,
                                                  const PackedCoeff *__restrict__ packed,
                                                  const double *__restrict__ sbf, double &er, double &ei) {
    const double fx = floor(x_pos), fy = floor(y_pos);
    const int ix = (int)fmin(fmax(fx, 0.0), (double)(SM_SBF_L - 2));
    const int iy = (int)fmin(fmax(fy, 0.0), (double)(SM_SBF_L - 2));
    const double wx = x_pos - fx, wy = y_pos - fy;
    shapelet_sum(packed, range, sbf, ix, iy, wx, wy, er, ei);
}

// One thread per component: its coefficients with i^((n1 + n2) mod 4) resolved. Terms with
// an even power (real part) first, then odd (imaginary part), each in source order, the sign
// of -1 folded into the value (exact). Each part of the envelope then adds exactly the terms
// it added before, in the same order. prange = (offset, real terms, imaginary terms, number
// of table rows the component uses).
__global__ void k_pack_coeffs(int ncomp, const int2 *__restrict__ range, const ShapeletCoeff *__restrict__ coeffs,
                              int4 *prange, PackedCoeff *packed) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncomp) return;
    const int2 r = range[c];
    int o = r.x, rows = 0;
    for (int part = 0; part < 2; ++part) {
        for (int j = 0; j < r.y; ++j) {
            const ShapeletCoeff cf = coeffs[r.x + j];
            const int q = (cf.n1 + cf.n2) & 3;
            if ((q & 1) != part) continue;
            const int n1 = cf.n1 < SM_SBF_N ? cf.n1 : SM_SBF_N - 1;
            const int n2 = cf.n2 < SM_SBF_N ? cf.n2 : SM_SBF_N - 1;
            PackedCoeff p;
            p.value = (q & 2) ? -cf.value : cf.value;
            p.n1 = n1;
            p.n2 = n2;
            packed[o++] = p;
            rows = max(rows, max(n1, n2) + 1);
        }
    }
    const int nre = [&] {
        int k = 0;
        for (int j = 0; j < r.y; ++j) k += ((coeffs[r.x + j].n1 + coeffs[r.x + j].n2) & 1) == 0;
        return k;
    }();
    prange[c] = make_int4(r.x, nre, r.y - nre, rows);
}

hipError_t launch_pack_coeffs(int ncomp, const int2 *d_range, const ShapeletCoeff *d_coeffs, int4 *d_prange,
                              PackedCoeff *d_packed, hipStream_t stream) {
    if (ncomp <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack_coeffs, dim3((ncomp + 255) / 256), dim3(256), 0, stream, ncomp, d_range, d_coeffs,
                       d_prange, d_packed);
    return hipGetLastError();
}

// ------------------------------------------------------------------ direct kernel
// Logical block (baseline block, channel chunk) from the hardware block id, placing all
// baseline blocks of a chunk on one XCD (blocks are dealt round-robin to the 8 XCDs).
__device__ __forceinline__ void logical_block(int gx, int total, int &bx, int &chunk) {
    int b = blockIdx.x;
    if ((total & 7) == 0) b = (b & 7) * (total >> 3) + (b >> 3);
    chunk = b / gx;
    bx = b - chunk * gx;
}

template <int KIND, bool PER_TILE, bool EVEN, bool SPLIT, int NCH>
__global__ __launch_bounds__(256) void k_direct(const DirectArgs a, int gx, int total) {
    int bx, chunk;
    logical_block(gx, total, bx, chunk);
    // Component range: all, or one of csplit ranges (subset mode).
    int c_lo = 0, c_hi = a.ncomp;
    if (SPLIT) {
        const int gy = chunk_count(a.chunks, NCH, a.nchan);
        const int split = chunk / gy;
        chunk -= split * gy;
        const int per = (a.ncomp + a.csplit - 1) / a.csplit;
        c_lo = split * per;
        c_hi = min(a.ncomp, c_lo + per);
    }
    // Lane -> baseline: all baselines, or the listed subset (outputs keep the full layout).
    const int nlanes = a.bl_subset != nullptr ? a.nsub : a.nbl;
    const int bl_raw = bx * 256 + (int)threadIdx.x;
    const int li = bl_raw < nlanes ? bl_raw : nlanes - 1;
    const int bl = a.bl_subset != nullptr ? a.bl_subset[li] : li;
    int f0, cnt;
    chunk_range<NCH>(a.chunks, chunk, a.nchan, f0, cnt);

    const double *__restrict__ coh = a.coh;
    const double4 *__restrict__ lmn4 = a.lmn;
    const double4 *__restrict__ shape = a.shape;
    const double *__restrict__ freqs = a.freqs;

    const UVW uvw = a.uvw[bl];
    const double F0 = freqs[f0];
    const double DF = freqs[f0 + 1] - F0;
    const double f0c = F0 / SM_VEL_C;
    const double dfc = DF / SM_VEL_C;

    int2 rows = make_int2(0, 0);
    if (PER_TILE) rows = a.beam.bl_rows[bl];

    double acc[NCH][8];
#pragma unroll
    for (int k = 0; k < NCH; ++k)
#pragma unroll
        for (int e = 0; e < 8; ++e) acc[k][e] = 0.0;

    for (int c = c_lo; c < c_hi; ++c) {
        const double4 lmn = lmn4[c];
        const double theta = uvw.u * lmn.x + uvw.v * lmn.y + uvw.w * lmn.z;
        double zr = 1.0, zi = 0.0, rr = 1.0, ri = 0.0;
        if (EVEN) {
            table_sincos(theta * f0c, kSinCos32, zi, zr);
            table_sincos(theta * dfc, kSinCos32, ri, rr);
        }
        double q = 0.0, env = 1.0, er = 1.0, eg = 1.0;
        double xk = 0.0, yk = 0.0;
        int4 crange = make_int4(0, 0, 0, 0);
        if (KIND == KIND_GAUSSIAN) {
            const double4 g = shape[c];
            q = g.x * (uvw.u * uvw.u) + g.y * (uvw.u * uvw.v) + g.z * (uvw.v * uvw.v);
            if (EVEN) {
                env = exp(q * (F0 * F0));
                er = exp(q * (DF * (2.0 * F0 + DF)));
                eg = exp(q * (2.0 * DF * DF));
            }
        } else if (KIND == KIND_SHAPELET) {
            const double4 sh = shape[c];
            const double2 suv = a.bl_uv[(size_t)c * a.nbl + bl];
            xk = (suv.x * sh.x + suv.y * sh.y) * sh.z;
            yk = (suv.x * sh.y - suv.y * sh.x) * sh.w;
            crange = a.prange[c];
        }
        const double *cc = coh + ((size_t)c * a.nchan_pad + f0) * 8;
#pragma unroll
        for (int k = 0; k < NCH; ++k) {
            double pr, pi;
            if (EVEN) {
                pr = zr;
                pi = zi;
            } else {
                table_sincos(theta * (freqs[f0 + k] / SM_VEL_C), kSinCos32, pi, pr);
            }
            if (KIND == KIND_GAUSSIAN) {
                const double e = EVEN ? env : exp(q * (freqs[f0 + k] * freqs[f0 + k]));
                pr *= e;
                pi *= e;
            } else if (KIND == KIND_SHAPELET) {
                const double fk = freqs[f0 + k];
                double sr, si;
                shapelet_envelope(xk * fk + SM_SBF_C, yk * fk + SM_SBF_C, crange, a.packed, a.sbf, sr, si);
                const double tr = pr * sr - pi * si;
                const double ti = pr * si + pi * sr;
                pr = tr;
                pi = ti;
            }
            double cj[8];
            if (PER_TILE) {
                double s[8], j1[8], j2[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) s[e] = cc[k * 8 + e];
                const int uf = a.beam.freq_map[f0 + k < a.nchan ? f0 + k : a.nchan - 1];
                load_jones(&a.beam.jones[((size_t)rows.x * a.beam.n_ufreq + uf) * a.beam.n_dir + a.beam.dir_offset + c], j1);
                load_jones(&a.beam.jones[((size_t)rows.y * a.beam.n_ufreq + uf) * a.beam.n_dir + a.beam.dir_offset + c], j2);
                sandwich(j1, s, j2, cj);
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) cj[e] = cc[k * 8 + e];
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                acc[k][2 * e] = fma(cj[2 * e], pr, acc[k][2 * e]);
                acc[k][2 * e] = fma(-cj[2 * e + 1], pi, acc[k][2 * e]);
                acc[k][2 * e + 1] = fma(cj[2 * e], pi, acc[k][2 * e + 1]);
                acc[k][2 * e + 1] = fma(cj[2 * e + 1], pr, acc[k][2 * e + 1]);
            }
            if (EVEN) {
                const double nr = zr * rr - zi * ri;
                zi = zr * ri + zi * rr;
                zr = nr;
                if (KIND == KIND_GAUSSIAN) {
                    env *= er;
                    er *= eg;
                }
            }
        }
    }
    if (bl_raw >= nlanes) return;
    if (SPLIT) {
#pragma unroll
        for (int k = 0; k < NCH; ++k) {
            const int f = f0 + k;
            if (k >= cnt) break;
            double *dst = a.accum + ((size_t)f * a.nsub + li) * 8;
#pragma unroll
            for (int e = 0; e < 8; ++e) unsafeAtomicAdd(&dst[e], acc[k][e]);
        }
        return;
    }
#pragma unroll
    for (int k = 0; k < NCH; ++k) {
        const int f = f0 + k;
        if (k >= cnt) break;
        const size_t o = (size_t)f * a.nbl + bl;
        float *v32 = reinterpret_cast<float *>(&a.vis32[o]);
#pragma unroll
        for (int e = 0; e < 8; ++e) v32[e] += (float)acc[k][e];
        if (a.vis64 != nullptr) {
            double *v64 = reinterpret_cast<double *>(&a.vis64[o]);
#pragma unroll
            for (int e = 0; e < 8; ++e) v64[e] += acc[k][e];
        }
    }
}

// Channels per lane of the baseline kernel. Four channels hold half the accumulators of
// SM_NCH = 8, one more wave per SIMD for points and Gaussians. Measured on the 256-tile,
// 384-channel per-tile-beam case: points 929 -> 853 ms, Gaussians 336 -> 300 ms per timestep.
// Shapelets take this kernel only for uneven channels or the drop-in path and keep SM_NCH.
template <int KIND, bool PT>
constexpr int kDirectNch = KIND != KIND_SHAPELET ? 4 : SM_NCH;
static_assert(SM_NCH % 4 == 0, "chunks of 4 channels tile the chunks of SM_NCH that the padding covers");

template <int KIND, bool PT>
static hipError_t launch_direct_kind(const DirectArgs &a, bool even, hipStream_t stream) {
    constexpr int NCH = kDirectNch<KIND, PT>;
    const int gx = ((a.bl_subset != nullptr ? a.nsub : a.nbl) + 255) / 256;
    const int gy = chunk_count(a.chunks, NCH, a.nchan);
    const int total = gx * gy * a.csplit;
    // The split form is a separate instantiation so that the common kernels compile as before.
    if (a.csplit > 1 && even)
        hipLaunchKernelGGL((k_direct<KIND, PT, true, true, NCH>), dim3(total), dim3(256), 0, stream, a, gx, total);
    else if (a.csplit > 1)
        hipLaunchKernelGGL((k_direct<KIND, PT, false, true, NCH>), dim3(total), dim3(256), 0, stream, a, gx, total);
    else if (even)
        hipLaunchKernelGGL((k_direct<KIND, PT, true, false, NCH>), dim3(total), dim3(256), 0, stream, a, gx, total);
    else
        hipLaunchKernelGGL((k_direct<KIND, PT, false, false, NCH>), dim3(total), dim3(256), 0, stream, a, gx, total);
    return hipGetLastError();
}

std::vector<int2> even_chunks(const double *freqs, int nchan, int width) {
    std::vector<int2> out;
    for (int f = 0; f < nchan;) {
        int cnt = nchan - f >= 2 ? 2 : 1;
        if (cnt == 2) {
            const double d = freqs[f + 1] - freqs[f];
            while (cnt < width && f + cnt < nchan && freqs[f + cnt] == freqs[f] + (double)cnt * d) ++cnt;
        }
        out.push_back(make_int2(f, cnt));
        f += cnt;
    }
    return out;
}

// Whether every fixed chunk's channels are exactly f0 + k (f1 - f0), for the chunk widths
// 8 (and with it 4) and 6; otherwise each channel gets its own sincos/exp. Evaluated on the
// host copy of the padded frequencies.
bool chunks_even(const double *freqs, int nchan) {
    for (const int w : {SM_NCH, 6}) {
        for (int f0 = 0; f0 < nchan; f0 += w) {
            const double d = freqs[f0 + 1] - freqs[f0];
            for (int k = 2; k < w && f0 + k < nchan; ++k)
                if (freqs[f0 + k] != freqs[f0] + (double)k * d) return false;
        }
    }
    return true;
}

static hipError_t launch_direct_switch(Kind kind, bool per_tile, bool even, const DirectArgs &a, hipStream_t stream);

__global__ void k_zero_f64(size_t n, double *x) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        x[i] = 0.0;
}

// vis[f][subset[i]] += accum[f][i], FP64 and its f32 demotion.
__global__ void k_subset_finish(int nchan, int nsub, int nbl, const int *__restrict__ subset,
                                const double *__restrict__ accum, JonesF32 *vis32, JonesF64 *vis64) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nchan * nsub) return;
    const int f = (int)(i / nsub);
    const int s = (int)(i % nsub);
    const size_t o = (size_t)f * nbl + subset[s];
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

int direct_subset_split(int nsub, int nchan, int ncomp) {
    const long lanes = (long)nsub * ((nchan + SM_NCH - 1) / SM_NCH);
    const long want = 1L << 18;  // about four waves per SIMD on 256 CUs
    long split = (want + lanes - 1) / std::max(lanes, 1L);
    split = std::min<long>(split, std::max(1, ncomp / 32));
    return (int)std::max(1L, std::min(split, 64L));
}

hipError_t launch_direct(Kind kind, bool per_tile, bool even, const DirectArgs &a_in, hipStream_t stream) {
    DirectArgs a = a_in;
    if (a.ncomp <= 0 || a.nbl <= 0 || a.nchan <= 0) return hipSuccess;
    if (a.bl_subset != nullptr && a.nsub <= 0) return hipSuccess;
    if (a.bl_subset == nullptr || a.accum == nullptr) a.csplit = 1;
    if (a.csplit > 1) {
        const size_t n = (size_t)a.nchan * a.nsub * 8;
        hipLaunchKernelGGL(k_zero_f64, dim3(1024), dim3(256), 0, stream, n, a.accum);
    }
    hipError_t e = launch_direct_switch(kind, per_tile, even, a, stream);
    if (e != hipSuccess || a.csplit <= 1) return e;
    const size_t n = (size_t)a.nchan * a.nsub;
    hipLaunchKernelGGL(k_subset_finish, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, a.nchan, a.nsub,
                       a.nbl, a.bl_subset, a.accum, a.vis32, a.vis64);
    return hipGetLastError();
}

static hipError_t launch_direct_switch(Kind kind, bool per_tile, bool even, const DirectArgs &a, hipStream_t stream) {
    switch (kind) {
    case KIND_POINT:
        return per_tile ? launch_direct_kind<KIND_POINT, true>(a, even, stream)
                        : launch_direct_kind<KIND_POINT, false>(a, even, stream);
    case KIND_GAUSSIAN:
        return per_tile ? launch_direct_kind<KIND_GAUSSIAN, true>(a, even, stream)
                        : launch_direct_kind<KIND_GAUSSIAN, false>(a, even, stream);
    default:
        return per_tile ? launch_direct_kind<KIND_SHAPELET, true>(a, even, stream)
                        : launch_direct_kind<KIND_SHAPELET, false>(a, even, stream);
    }
}

// ------------------------------------------------------------------ autos
// V_tt(f) += sum_c J_t S_c J_t^H (src/model/cpu.rs:869-904, src/gpu/model.cu:368-485).
__global__ void k_autos(int ntiles, int nchan, int ncomp, const double *__restrict__ coh, int nchan_pad,
                        BeamDesc beam, int coh_has_beam, JonesF32 *vis32, JonesF64 *vis64, double *accum64,
                        int demote_now) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntiles * nchan) return;
    const int f = i % nchan;
    const int t = i / nchan;
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int c = 0; c < ncomp; ++c) {
        const double *s = coh + ((size_t)c * nchan_pad + f) * 8;
        if (coh_has_beam || beam.jones == nullptr) {
#pragma unroll
            for (int e = 0; e < 8; ++e) acc[e] += s[e];
        } else {
            double j[8], sv[8], o[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) sv[e] = s[e];
            load_jones(&beam.jones[((size_t)beam.tile_rows[t] * beam.n_ufreq + beam.freq_map[f]) * beam.n_dir + beam.dir_offset + c], j);
            sandwich(j, sv, j, o);
#pragma unroll
            for (int e = 0; e < 8; ++e) acc[e] += o[e];
        }
    }
    const size_t o = (size_t)f * ntiles + t;
    if (demote_now) {
        float *v32 = reinterpret_cast<float *>(&vis32[o]);
#pragma unroll
        for (int e = 0; e < 8; ++e) v32[e] += (float)acc[e];
        if (vis64 != nullptr) {
            double *v64 = reinterpret_cast<double *>(&vis64[o]);
#pragma unroll
            for (int e = 0; e < 8; ++e) v64[e] += acc[e];
        }
    } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) accum64[o * 8 + e] += acc[e];
    }
}

hipError_t launch_autos(int ntiles, int nchan, int ncomp, const double *d_coh, int nchan_pad, BeamDesc beam,
                        bool coh_has_beam, JonesF32 *vis32, JonesF64 *vis64, double *accum64, bool demote_now,
                        hipStream_t stream) {
    if (ncomp <= 0 || ntiles <= 0 || nchan <= 0) return hipSuccess;
    const int n = ntiles * nchan;
    hipLaunchKernelGGL(k_autos, dim3((n + 255) / 256), dim3(256), 0, stream, ntiles, nchan, ncomp, d_coh, nchan_pad,
                       beam, coh_has_beam ? 1 : 0, vis32, vis64, accum64, demote_now ? 1 : 0);
    return hipGetLastError();
}

__global__ void k_autos_demote(int n, const double *__restrict__ accum64, JonesF32 *vis32, JonesF64 *vis64) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float *v32 = reinterpret_cast<float *>(&vis32[i]);
#pragma unroll
    for (int e = 0; e < 8; ++e) v32[e] += (float)accum64[(size_t)i * 8 + e];
    if (vis64 != nullptr) {
        double *v64 = reinterpret_cast<double *>(&vis64[i]);
#pragma unroll
        for (int e = 0; e < 8; ++e) v64[e] += accum64[(size_t)i * 8 + e];
    }
}

hipError_t launch_autos_demote(int n, const double *accum64, JonesF32 *vis32, JonesF64 *vis64, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_autos_demote, dim3((n + 255) / 256), dim3(256), 0, stream, n, accum64, vis32, vis64);
    return hipGetLastError();
}

// ------------------------------------------------------------------ shapelet UVs
// Baseline (u, v) with each shapelet component as phase centre: per-tile UVW from the
// component's (sin ha, cos ha, sin dec, cos dec), then tile1 - tile2
// (src/srclist/types/components/mod.rs:299-315; coords.md "UVWs"). Contraction is off so
// the values equal the CPU restatement bit for bit.
__global__ void k_shapelet_bl_uv(int ncomp, int ntiles, int nbl, const double *__restrict__ hadec,
                                 const double *__restrict__ xyz, const int2 *__restrict__ bl_tiles,
                                 double2 *__restrict__ out) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)ncomp * nbl) return;
    const int c = (int)(idx / nbl);
    const int b = (int)(idx % nbl);
    const double4 hd = make_double4(hadec[4 * c], hadec[4 * c + 1], hadec[4 * c + 2], hadec[4 * c + 3]);
    const int2 t = bl_tiles[b];
    (void)ntiles;
    out[idx] = shapelet_uv(hd, xyz[3 * t.x], xyz[3 * t.x + 1], xyz[3 * t.x + 2], xyz[3 * t.y], xyz[3 * t.y + 1],
                           xyz[3 * t.y + 2]);
}

hipError_t launch_shapelet_bl_uv(int ncomp, int ntiles, int nbl, const double *d_comp_hadec, const double *d_xyz,
                                 const int2 *d_bl_tiles, double2 *d_bl_uv, hipStream_t stream) {
    if (ncomp <= 0 || nbl <= 0) return hipSuccess;
    const size_t n = (size_t)ncomp * nbl;
    hipLaunchKernelGGL(k_shapelet_bl_uv, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, ncomp, ntiles, nbl,
                       d_comp_hadec, d_xyz, d_bl_tiles, d_bl_uv);
    return hipGetLastError();
}

// [baseline][bucket component] (three buckets) -> [component][baseline].
__global__ void k_repack_uv(int nbl, int n_pl, int n_cpl, int n_list, const ShapeletUV *__restrict__ pl,
                            const ShapeletUV *__restrict__ cpl, const ShapeletUV *__restrict__ list,
                            double2 *__restrict__ out) {
    const int n = n_pl + n_cpl + n_list;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)n * nbl) return;
    const int c = (int)(idx / nbl);
    const int b = (int)(idx % nbl);
    ShapeletUV s;
    if (c < n_pl)
        s = pl[(size_t)b * n_pl + c];
    else if (c < n_pl + n_cpl)
        s = cpl[(size_t)b * n_cpl + (c - n_pl)];
    else
        s = list[(size_t)b * n_list + (c - n_pl - n_cpl)];
    out[idx] = make_double2(s.u, s.v);
}

hipError_t launch_repack_shapelet_uv(int nbl, int n_pl, int n_cpl, int n_list, const ShapeletUV *pl,
                                     const ShapeletUV *cpl, const ShapeletUV *list, double2 *d_bl_uv,
                                     hipStream_t stream) {
    const size_t n = (size_t)(n_pl + n_cpl + n_list) * nbl;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_repack_uv, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, nbl, n_pl, n_cpl, n_list,
                       pl, cpl, list, d_bl_uv);
    return hipGetLastError();
}

// ------------------------------------------------------------------ analytic beam
// The product's MWA-like beam (see the header comment of sm_analytic_beam): 4x4 crossed
// short dipoles, 1.1 m pitch, 0.278 m above a ground plane, 435 ps delay steps, delay 32 =
// dead dipole. A block owns 256 directions of one tile and kBeamFB frequencies: each thread
// forms its direction's geometry once and reuses it for every frequency. The dipole phase
// factorises into column x row terms, exp(i k (x r_e + y r_n)) = cx(x) cy(y), with
// x, y in {+-0.55, +-1.65} m; the 1.65 m terms follow from the 0.55 m ones by the triple-angle
// identities, so a (direction, frequency) takes two sincos for geometry and one for the
// ground-plane factor. The array factor sums row by row,
//   AF = sum_row cy(row) sum_col w(row, col) cx(col),
// with w = gain x delay phasor of each dipole, built per (tile, frequency) in LDS.
constexpr int kBeamFB = 8;

__global__ __launch_bounds__(256) void k_analytic_beam(int n_freqs, const double *__restrict__ freqs, int n_dirs,
                                                       const double *__restrict__ az, const double *__restrict__ za,
                                                       const unsigned *__restrict__ delays,
                                                       const double *__restrict__ gains, JonesF64 *__restrict__ out) {
    __shared__ double s_w[kBeamFB][16][4];  // per dipole: X re, X im, Y re, Y im
    __shared__ double s_k[kBeamFB], s_norm[kBeamFB];
    const int nfb = (n_freqs + kBeamFB - 1) / kBeamFB;
    const int tile = blockIdx.y / nfb;
    const int f0 = (blockIdx.y - tile * nfb) * kBeamFB;
    const int nf = min(kBeamFB, n_freqs - f0);
    if (threadIdx.x < 16 * kBeamFB) {
        const int j = threadIdx.x >> 4, d = threadIdx.x & 15;
        if (j < nf) {
            const double freq = freqs[f0 + j];
            unsigned dl = delays[16 * tile + d];
            double gx = gains[32 * tile + d], gy = gains[32 * tile + 16 + d];
            if (dl == 32) {
                gx = 0.0;
                gy = 0.0;
                dl = 0;
            }
            double s, c;
            sincos(-6.283185307179586 * freq * ((double)dl * 435e-12), &s, &c);
            s_w[j][d][0] = gx * c;
            s_w[j][d][1] = gx * s;
            s_w[j][d][2] = gy * c;
            s_w[j][d][3] = gy * s;
            if (d == 0) {
                const double k = 6.283185307179586 * freq / SM_VEL_C;
                s_k[j] = k;
                s_norm[j] = 1.0 / (16.0 * sin(k * 0.278));
            }
        }
    }
    __syncthreads();
    const int dir = blockIdx.x * blockDim.x + threadIdx.x;
    if (dir >= n_dirs) return;
    double s_za, c_za, s_az, c_az;
    sincos(za[dir], &s_za, &c_za);
    sincos(az[dir], &s_az, &c_az);
    const double r_e = s_za * s_az, r_n = s_za * c_az, r_u = c_za;
    double sx = sqrt(1.0 - r_e * r_e);
    if (sx < 1e-12) sx = 1e-12;
    const double qyx = -r_e * r_n / sx, qyy = r_u / sx;
    for (int j = 0; j < nf; ++j) {
        const double k = s_k[j];
        // Column phases for x = -1.65, -0.55, 0.55, 1.65 m; row phases for y = 1.65, 0.55,
        // -0.55, -1.65 m (rows 0..3). sin 3a = s (3 - 4 s^2), cos 3a = c (4 c^2 - 3).
        double cx[4][2], cy[4][2];
        {
            double s1, c1;
            table_sincos(k * 0.55 * r_e, kSinCos32, s1, c1);
            const double s3 = s1 * (3.0 - 4.0 * s1 * s1), c3 = c1 * (4.0 * c1 * c1 - 3.0);
            cx[0][0] = c3; cx[0][1] = -s3;
            cx[1][0] = c1; cx[1][1] = -s1;
            cx[2][0] = c1; cx[2][1] = s1;
            cx[3][0] = c3; cx[3][1] = s3;
        }
        {
            double s1, c1;
            table_sincos(k * 0.55 * r_n, kSinCos32, s1, c1);
            const double s3 = s1 * (3.0 - 4.0 * s1 * s1), c3 = c1 * (4.0 * c1 * c1 - 3.0);
            cy[0][0] = c3; cy[0][1] = s3;
            cy[1][0] = c1; cy[1][1] = s1;
            cy[2][0] = c1; cy[2][1] = -s1;
            cy[3][0] = c3; cy[3][1] = -s3;
        }
        double ax_r = 0, ax_i = 0, ay_r = 0, ay_i = 0;
#pragma unroll
        for (int row = 0; row < 4; ++row) {
            double rx_r = 0, rx_i = 0, ry_r = 0, ry_i = 0;
#pragma unroll
            for (int col = 0; col < 4; ++col) {
                const double *w = s_w[j][4 * row + col];
                rx_r += w[0] * cx[col][0] - w[1] * cx[col][1];
                rx_i += w[0] * cx[col][1] + w[1] * cx[col][0];
                ry_r += w[2] * cx[col][0] - w[3] * cx[col][1];
                ry_i += w[2] * cx[col][1] + w[3] * cx[col][0];
            }
            ax_r += rx_r * cy[row][0] - rx_i * cy[row][1];
            ax_i += rx_r * cy[row][1] + rx_i * cy[row][0];
            ay_r += ry_r * cy[row][0] - ry_i * cy[row][1];
            ay_i += ry_r * cy[row][1] + ry_i * cy[row][0];
        }
        double sg, cg;
        table_sincos(k * 0.278 * r_u, kSinCos32, sg, cg);
        const double g = sg * s_norm[j];
        const double px = g * sx, pyx = g * qyx, pyy = g * qyy;
        JonesF64 jo;
        jo.j00_re = ax_r * px;
        jo.j00_im = ax_i * px;
        jo.j01_re = 0.0;
        jo.j01_im = 0.0;
        jo.j10_re = ay_r * pyx;
        jo.j10_im = ay_i * pyx;
        jo.j11_re = ay_r * pyy;
        jo.j11_im = ay_i * pyy;
        out[((size_t)tile * n_freqs + f0 + j) * n_dirs + dir] = jo;
    }
}

hipError_t launch_analytic_beam(int n_tiles, const unsigned *d_delays, const double *d_gains, int n_freqs,
                                const double *d_freqs, int n_dirs, const double *d_az, const double *d_za,
                                JonesF64 *d_out, hipStream_t stream) {
    if (n_tiles <= 0 || n_freqs <= 0 || n_dirs <= 0) return hipSuccess;
    // Tiles in batches that keep the grid's y extent (tiles x frequency blocks) within 65535.
    const int nfb = (n_freqs + kBeamFB - 1) / kBeamFB;
    if (nfb > 65535) return hipErrorInvalidValue;
    const int per = 65535 / nfb;
    for (int t0 = 0; t0 < n_tiles; t0 += per) {
        const int nt = n_tiles - t0 < per ? n_tiles - t0 : per;
        dim3 grid((n_dirs + 255) / 256, (unsigned)(nt * nfb));
        hipLaunchKernelGGL(k_analytic_beam, grid, dim3(256), 0, stream, n_freqs, d_freqs, n_dirs, d_az, d_za,
                           d_delays + 16 * (size_t)t0, d_gains + 32 * (size_t)t0,
                           d_out + (size_t)t0 * n_freqs * n_dirs);
    }
    return hipGetLastError();
}

}  // namespace sm
