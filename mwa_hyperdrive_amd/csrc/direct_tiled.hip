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
constexpr int kCT = 32;   // components per LDS stage
// Channels per lane in k_direct_tiled. Fewer channels hold fewer accumulators (4 waves per
// SIMD instead of 3 for points) but share each component's phasor products over fewer
// channels. Measured on C2 (768 channels): points 35.1 ms with 8, 33.6 ms with 4; Gaussians
// 19.7 ms with 8, 21.7 ms with 4.
constexpr int kTiledNchPoint = 4;
constexpr int kTiledNchGaussian = 8;
static_assert(kTiledNchPoint <= SM_NCH && kTiledNchGaussian <= SM_NCH,
              "the padded channel count covers whole chunks of SM_NCH channels");

template <int KIND, int NCH>
__global__ __launch_bounds__(256) void k_direct_tiled(const TiledArgs a) {
    __shared__ double4 s_ph[kCT][32];
    __shared__ double s_uvw[32][3];
    const int total = a.npairs * a.nchunks;
    int b = blockIdx.x;
    if ((total & 7) == 0) b = (b & 7) * (total >> 3) + (b >> 3);
    const int chunk = b / a.npairs;
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

    const int f0 = chunk * NCH;
    const double *__restrict__ freqs = a.freqs;
    const double F0 = freqs[f0];
    const double DF = freqs[f0 + 1] - F0;
    const double f0c = F0 / SM_VEL_C;
    const double dfc = DF / SM_VEL_C;
    const double *__restrict__ coh = a.coh;
    const double4 *__restrict__ lmn4 = a.lmn;
    const double4 *__restrict__ shape = a.shape;

    double acc[NCH][8];
#pragma unroll
    for (int k = 0; k < NCH; ++k)
#pragma unroll
        for (int e = 0; e < 8; ++e) acc[k][e] = 0.0;

    for (int c0 = 0; c0 < a.ncomp; c0 += kCT) {
        const int nc = a.ncomp - c0 < kCT ? a.ncomp - c0 : kCT;
        // Per-tile phasors (start, step) of this stage's components.
        for (int e = tid; e < kCT * 32; e += 256) {
            const int slot = e & 31, cc = e >> 5;
            double4 ph = make_double4(1.0, 0.0, 1.0, 0.0);
            if (cc < nc) {
                const double4 l = lmn4[c0 + cc];
                const double th = s_uvw[slot][0] * l.x + s_uvw[slot][1] * l.y + s_uvw[slot][2] * l.z;
                fast_sincos(th * f0c, ph.y, ph.x);
                fast_sincos(th * dfc, ph.w, ph.z);
            }
            s_ph[cc][slot] = ph;
        }
        __syncthreads();
        for (int cc = 0; cc < nc; ++cc) {
            const int c = c0 + cc;
            const double4 A = s_ph[cc][sp];
            const double4 B = s_ph[cc][sq];
            double zr = A.x * B.x + A.y * B.y;
            double zi = A.y * B.x - A.x * B.y;
            const double rr = A.z * B.z + A.w * B.w;
            const double ri = A.w * B.z - A.z * B.w;
            double env = 1.0, er = 1.0, eg = 1.0;
            if (KIND == KIND_GAUSSIAN) {
                const double4 g = shape[c];
                const double qf = g.x * (bu * bu) + g.y * (bu * bv) + g.z * (bv * bv);
                env = exp(qf * (F0 * F0));
                er = exp(qf * (DF * (2.0 * F0 + DF)));
                eg = exp(qf * (2.0 * DF * DF));
            }
            const double *cc8 = coh + ((size_t)c * a.nchan_pad + f0) * 8;
#pragma unroll
            for (int k = 0; k < NCH; ++k) {
                double pr_ = zr, pi_ = zi;
                if (KIND == KIND_GAUSSIAN) {
                    pr_ *= env;
                    pi_ *= env;
                }
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const double cr = cc8[k * 8 + 2 * e], ci = cc8[k * 8 + 2 * e + 1];
                    acc[k][2 * e] = fma(cr, pr_, acc[k][2 * e]);
                    acc[k][2 * e] = fma(-ci, pi_, acc[k][2 * e]);
                    acc[k][2 * e + 1] = fma(cr, pi_, acc[k][2 * e + 1]);
                    acc[k][2 * e + 1] = fma(ci, pr_, acc[k][2 * e + 1]);
                }
                const double nr = zr * rr - zi * ri;
                zi = zr * ri + zi * rr;
                zr = nr;
                if (KIND == KIND_GAUSSIAN) {
                    env *= er;
                    er *= eg;
                }
            }
        }
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
        if (f >= a.nchan) break;
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
    const int nch = kind == KIND_GAUSSIAN ? kTiledNchGaussian : kTiledNchPoint;
    a.nchunks = (a.nchan + nch - 1) / nch;
    const int total = a.npairs * a.nchunks;
    if (kind == KIND_GAUSSIAN)
        hipLaunchKernelGGL((k_direct_tiled<KIND_GAUSSIAN, kTiledNchGaussian>), dim3(total), dim3(256), 0, stream, a);
    else
        hipLaunchKernelGGL((k_direct_tiled<KIND_POINT, kTiledNchPoint>), dim3(total), dim3(256), 0, stream, a);
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
// Gaussians here take 4 channels per lane (2 -> 3 waves per SIMD), unlike the tile-pair
// kernel: on the 256-tile, 384-channel case 236 -> 227 ms per timestep.
constexpr int kRectNchGaussian = 4;

template <int KIND, int NCH>
__global__ __launch_bounds__(256) void k_direct_rect(const RectArgs a) {
    __shared__ double4 s_ph[kRC][kRT];
    __shared__ double s_uvw[kRT][3];
    const int nrg = (a.n_row + 3) / 4, ncg = (a.n_col + 63) / 64;
    const int b = blockIdx.x;
    const int chunk = b / (nrg * ncg);
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

    const int f0 = chunk * NCH;
    const double *__restrict__ freqs = a.freqs;
    const double F0 = freqs[f0];
    const double DF = freqs[f0 + 1] - F0;
    const double f0c = F0 / SM_VEL_C;
    const double dfc = DF / SM_VEL_C;
    const double *__restrict__ coh = a.coh_cls + (size_t)cls * a.ncomp * a.nchan_pad * 8;
    const double4 *__restrict__ lmn4 = a.lmn;
    const double4 *__restrict__ shape = a.shape;

    double acc[NCH][8];
#pragma unroll
    for (int k = 0; k < NCH; ++k)
#pragma unroll
        for (int e = 0; e < 8; ++e) acc[k][e] = 0.0;

    for (int c0 = 0; c0 < a.ncomp; c0 += kRC) {
        const int nc = a.ncomp - c0 < kRC ? a.ncomp - c0 : kRC;
        for (int e = tid; e < kRC * kRT; e += 256) {
            const int cc = e / kRT, slot = e - cc * kRT;
            double4 ph = make_double4(1.0, 0.0, 1.0, 0.0);
            if (cc < nc) {
                const double4 l = lmn4[c0 + cc];
                const double th = s_uvw[slot][0] * l.x + s_uvw[slot][1] * l.y + s_uvw[slot][2] * l.z;
                fast_sincos(th * f0c, ph.y, ph.x);
                fast_sincos(th * dfc, ph.w, ph.z);
            }
            s_ph[cc][slot] = ph;
        }
        __syncthreads();
        if (row_ok) {
            for (int cc = 0; cc < nc; ++cc) {
                const int c = c0 + cc;
                const double4 A = s_ph[cc][sp];
                const double4 B = s_ph[cc][sq];
                double zr = A.x * B.x + A.y * B.y;
                double zi = A.y * B.x - A.x * B.y;
                const double rr = A.z * B.z + A.w * B.w;
                const double ri_ = A.w * B.z - A.z * B.w;
                double env = 1.0, er = 1.0, eg = 1.0;
                if (KIND == KIND_GAUSSIAN) {
                    const double4 g = shape[c];
                    const double qf = g.x * (bu * bu) + g.y * (bu * bv) + g.z * (bv * bv);
                    env = exp(qf * (F0 * F0));
                    er = exp(qf * (DF * (2.0 * F0 + DF)));
                    eg = exp(qf * (2.0 * DF * DF));
                }
                const double *cc8 = coh + ((size_t)c * a.nchan_pad + f0) * 8;
#pragma unroll
                for (int k = 0; k < NCH; ++k) {
                    double pr_ = zr, pi_ = zi;
                    if (KIND == KIND_GAUSSIAN) {
                        pr_ *= env;
                        pi_ *= env;
                    }
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const double cr = cc8[k * 8 + 2 * e], ci = cc8[k * 8 + 2 * e + 1];
                        acc[k][2 * e] = fma(cr, pr_, acc[k][2 * e]);
                        acc[k][2 * e] = fma(-ci, pi_, acc[k][2 * e]);
                        acc[k][2 * e + 1] = fma(cr, pi_, acc[k][2 * e + 1]);
                        acc[k][2 * e + 1] = fma(ci, pr_, acc[k][2 * e + 1]);
                    }
                    const double nr = zr * rr - zi * ri_;
                    zi = zr * ri_ + zi * rr;
                    zr = nr;
                    if (KIND == KIND_GAUSSIAN) {
                        env *= er;
                        er *= eg;
                    }
                }
            }
        }
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
        if (f >= a.nchan) break;
        double v[8];
        if (flip) {
            // (V^H)_ij = conj(V_ji): XX, YX, XY, YY with the imaginary parts negated.
            v[0] = acc[k][0];
            v[1] = -acc[k][1];
            v[2] = acc[k][4];
            v[3] = -acc[k][5];
            v[4] = acc[k][2];
            v[5] = -acc[k][3];
            v[6] = acc[k][6];
            v[7] = -acc[k][7];
        } else {
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = acc[k][e];
        }
        const size_t o = (size_t)f * a.nbl + bl;
        float *v32 = reinterpret_cast<float *>(&a.vis32[o]);
#pragma unroll
        for (int e = 0; e < 8; ++e) v32[e] += (float)v[e];
        if (a.vis64 != nullptr) {
            double *v64 = reinterpret_cast<double *>(&a.vis64[o]);
#pragma unroll
            for (int e = 0; e < 8; ++e) v64[e] += v[e];
        }
    }
}

hipError_t launch_direct_rect(Kind kind, const RectArgs &a, hipStream_t stream) {
    if (a.ncomp <= 0 || a.n_row <= 0 || a.n_col <= 0 || a.nchan <= 0) return hipSuccess;
    const int nch = kind == KIND_GAUSSIAN ? kRectNchGaussian : kTiledNchPoint;
    const long total = (long)((a.n_row + 3) / 4) * ((a.n_col + 63) / 64) * ((a.nchan + nch - 1) / nch);
    if (kind == KIND_GAUSSIAN)
        hipLaunchKernelGGL((k_direct_rect<KIND_GAUSSIAN, kRectNchGaussian>), dim3((unsigned)total), dim3(256), 0,
                           stream, a);
    else
        hipLaunchKernelGGL((k_direct_rect<KIND_POINT, kTiledNchPoint>), dim3((unsigned)total), dim3(256), 0, stream,
                           a);
    return hipGetLastError();
}

// ------------------------------------------------------------------ shapelets
template <bool PER_TILE>
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
    const double *__restrict__ sbf = a.sbf;
    // Per-wave windows of the basis table (kRows rows x kWin cells per axis).
    __shared__ double wx_lds[4][kRows][kWin], wy_lds[4][kRows][kWin];
    const int wv = threadIdx.x >> 6;
    // The wave's lowest and highest channel frequency (invalid lanes repeat the last channel).
    double fmin_w = freq, fmax_w = freq;
    for (int o = 32; o > 0; o >>= 1) {
        fmin_w = fmin(fmin_w, __shfl_xor(fmin_w, o));
        fmax_w = fmax(fmax_w, __shfl_xor(fmax_w, o));
    }
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // The wave's baseline is uniform; the next component's parameters are loaded one
    // iteration ahead.
    const int blu = __builtin_amdgcn_readfirstlane(bl);
    double4 l_next = a.lmn[0], sh_next = a.shape[0];
    double2 suv_next = a.bl_uv[blu];
    int4 pr_next = a.prange[0];
    for (int c = 0; c < a.ncomp; ++c) {
        const double4 l = l_next, sh = sh_next;
        const double2 suv = suv_next;
        const int4 pr4 = pr_next;
        {
            const int cn = c + 1 < a.ncomp ? c + 1 : c;
            l_next = a.lmn[cn];
            sh_next = a.shape[cn];
            suv_next = a.bl_uv[(size_t)cn * a.nbl + blu];
            pr_next = a.prange[cn];
        }
        const double theta = uvw.u * l.x + uvw.v * l.y + uvw.w * l.z;
        double zr, zi;
        fast_sincos(theta * fc, zi, zr);
        const double xk = (suv.x * sh.x + suv.y * sh.y) * sh.z;
        const double yk = (suv.x * sh.y - suv.y * sh.x) * sh.w;
        const double x_pos = xk * freq + SM_SBF_C;
        const double y_pos = yk * freq + SM_SBF_C;
        const double fx = floor(x_pos), fy = floor(y_pos);
        const int ix = (int)fmin(fmax(fx, 0.0), (double)(SM_SBF_L - 2));
        const int iy = (int)fmin(fmax(fy, 0.0), (double)(SM_SBF_L - 2));
        const double wx = x_pos - fx, wy = y_pos - fy;
        double er, ei;
        // The cells the wave touches: positions are monotone in frequency, so the extreme
        // channels bound them (the same expressions, so the same roundings).
        const double xa = xk * fmin_w + SM_SBF_C, xb = xk * fmax_w + SM_SBF_C;
        const double ya = yk * fmin_w + SM_SBF_C, yb = yk * fmax_w + SM_SBF_C;
        const int ix0 = (int)fmin(fmax(floor(fmin(xa, xb)), 0.0), (double)(SM_SBF_L - 2));
        const int ix1 = (int)fmin(fmax(floor(fmax(xa, xb)), 0.0), (double)(SM_SBF_L - 2));
        const int iy0 = (int)fmin(fmax(floor(fmin(ya, yb)), 0.0), (double)(SM_SBF_L - 2));
        const int iy1 = (int)fmin(fmax(floor(fmax(ya, yb)), 0.0), (double)(SM_SBF_L - 2));
        const int xw = __builtin_amdgcn_readfirstlane(ix1 - ix0 + 2);
        const int yw = __builtin_amdgcn_readfirstlane(iy1 - iy0 + 2);
        double cpre[8];
        const double *s_c = a.coh + ((size_t)c * a.nchan_pad + f) * 8;
        if (xw <= kWin && yw <= kWin && pr4.w <= kRows) {
            // Stage rows 0 .. pr4.w - 1 of both windows in this wave's LDS, then interpolate
            // from LDS: one 16-byte read per axis per coefficient instead of two gathers.
            const int bx0 = __builtin_amdgcn_readfirstlane(ix0), by0 = __builtin_amdgcn_readfirstlane(iy0);
            // All loads first, then all LDS writes: the loads are in flight together instead of
            // one round trip per row. A lane past the window or past row pr4.w - 1 reads an
            // in-bounds cell (the window's last cell; rows below kRows < SM_SBF_N) that is never
            // written. Byte offsets are 32-bit (the table is 8 MB), from the uniform base.
            const int j = lane & (kWin - 1), half = lane / kWin;
            const int jxc = min(j, xw - 1), jyc = min(j, yw - 1);
            static_assert(kRows <= SM_SBF_N, "staged rows must exist in the basis table");
            const unsigned ox = (unsigned)((half * SM_SBF_L + bx0 + jxc) * 8);
            const unsigned oy = (unsigned)((half * SM_SBF_L + by0 + jyc) * 8);
            const char *sbf_b = reinterpret_cast<const char *>(sbf);
            constexpr int kBatch = 2;  // row steps in flight at once (register budget)
            for (int i0 = 0; i0 * (64 / kWin) < pr4.w; i0 += kBatch) {
                double tx[kBatch], ty[kBatch];
#pragma unroll
                for (int i = 0; i < kBatch; ++i) {
                    const unsigned step = (unsigned)((i0 + i) * (64 / kWin) * SM_SBF_L * 8);
                    if ((i0 + i) * (64 / kWin) < pr4.w) {
                        tx[i] = *reinterpret_cast<const double *>(sbf_b + (ox + step));
                        ty[i] = *reinterpret_cast<const double *>(sbf_b + (oy + step));
                    }
                }
#pragma unroll
                for (int i = 0; i < kBatch; ++i) {
                    const int r = half + (i0 + i) * (64 / kWin);
                    if ((i0 + i) * (64 / kWin) < pr4.w && r < pr4.w) {
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
            double part[2];
#pragma unroll
            for (int pt = 0; pt < 2; ++pt) {
                const PackedCoeff *cp = packed + pr4.x + (pt ? pr4.y : 0);
                const int cnt = pt ? pr4.z : pr4.y;
                double accp = 0.0;
#pragma unroll 2
                for (int k = 0; k < cnt; ++k) {
                    double val;
                    int n1, n2;
                    load_packed_uniform(cp + k, val, n1, n2);
                    const double xl = wx_lds[wv][n1][jx], xh = wx_lds[wv][n1][jx + 1];
                    const double yl = wy_lds[wv][n2][jy], yh = wy_lds[wv][n2][jy + 1];
                    const double u_value = xl + (xh - xl) * wx;
                    const double v_value = yl + (yh - yl) * wy;
                    accp += (val * u_value) * v_value;
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
            double part[2];
#pragma unroll
            for (int pt = 0; pt < 2; ++pt) {
                const PackedCoeff *cp = packed + pr4.x + (pt ? pr4.y : 0);
                const int cnt = pt ? pr4.z : pr4.y;
                double accp = 0.0;
#pragma unroll 1
                for (int k = 0; k < cnt; ++k) {
                    double val;
                    int n1, n2;
                    load_packed_uniform(cp + k, val, n1, n2);
                    const double *bx = sbf + n1 * SM_SBF_L + ix, *by = sbf + n2 * SM_SBF_L + iy;
                    const double u_value = bx[0] + (bx[1] - bx[0]) * wx;
                    const double v_value = by[0] + (by[1] - by[0]) * wy;
                    accp += (val * u_value) * v_value;
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
        const double pr = zr * er - zi * ei;
        const double pi = zr * ei + zi * er;
        double cj[8];
        const double *s = a.coh + ((size_t)c * a.nchan_pad + f) * 8;
        if (PER_TILE && cls >= 0) {
            class_coh(a.beam.coh_cls + (((size_t)cls * a.ncomp + c) * a.nchan_pad + f) * 8, herm, cj);
        } else if (PER_TILE) {
            double sv[8], j1[8], j2[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) sv[e] = s[e];
            load_jones_t(&a.beam.jones[((size_t)rows.x * a.beam.n_ufreq + uf) * a.beam.n_dir + a.beam.dir_offset + c], j1);
            load_jones_t(&a.beam.jones[((size_t)rows.y * a.beam.n_ufreq + uf) * a.beam.n_dir + a.beam.dir_offset + c], j2);
            sandwich_t(j1, sv, j2, cj);
        } else {
#pragma unroll
            for (int e = 0; e < 8; ++e) cj[e] = cpre[e];
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            acc[2 * e] = fma(cj[2 * e], pr, acc[2 * e]);
            acc[2 * e] = fma(-cj[2 * e + 1], pi, acc[2 * e]);
            acc[2 * e + 1] = fma(cj[2 * e], pi, acc[2 * e + 1]);
            acc[2 * e + 1] = fma(cj[2 * e + 1], pr, acc[2 * e + 1]);
        }
    }
    if (!valid) return;
    const size_t o = (size_t)f * a.nbl + bl;
    float *v32 = reinterpret_cast<float *>(&a.vis32[o]);
#pragma unroll
    for (int e = 0; e < 8; ++e) v32[e] += (float)acc[e];
    if (a.vis64 != nullptr) {
        double *v64 = reinterpret_cast<double *>(&a.vis64[o]);
#pragma unroll
        for (int e = 0; e < 8; ++e) v64[e] += acc[e];
    }
}

__global__ __launch_bounds__(256) void k_shapelet_chan(const ShapeletChanArgs a) { shapelet_chan<false>(a); }

// Per-tile beams: held to 128 VGPRs (4 waves per SIMD, a few spilled values) instead of the
// 164 the compiler picks (3 waves). Measured on the 256-tile, 384-channel case: 129 -> 111 ms
// per timestep. The shared-beam kernel keeps the compiler's choice: with the same cap it ran
// 3.7% slower.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_shapelet_chan_tiles(
    const ShapeletChanArgs a) {
    shapelet_chan<true>(a);
}

hipError_t launch_shapelet_chan(bool per_tile, const ShapeletChanArgs &a, hipStream_t stream) {
    if (a.ncomp <= 0 || a.nbl <= 0 || a.nchan <= 0) return hipSuccess;
    const int total = ((a.nbl + 3) / 4) * a.ngroups;
    if (per_tile)
        hipLaunchKernelGGL(k_shapelet_chan_tiles, dim3(total), dim3(256), 0, stream, a);
    else
        hipLaunchKernelGGL(k_shapelet_chan, dim3(total), dim3(256), 0, stream, a);
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
