// Device-resident neighbours of the modeller in the reference's peel loop
// (src/params/peel/gpu.rs:441-470): tile XYZs -> baseline UVWs at a pointing centre, the phase
// rotation of visibilities between two sets of UVWs, and the addition of an
// ionospherically offset, scaled model to residuals. Same C signatures and argument meanings
// as src/gpu/peel.h:13-35 (FLOAT = double), so the reference's Rust bindings
// (src/gpu/peel_double.rs) link them unchanged.
//
// Layouts: visibilities [time][freq][baseline] JonesF32; UVWs [time][baseline]; XYZs
// [time][tile]; baselines in the upper-triangle order of src/math/mod.rs:96-109.
#include <cmath>

#include "sm_internal.h"

namespace sm {
namespace {

constexpr double kTauPeel = 6.283185307179586;

// xyzs_to_uvws_kernel (src/gpu/peel.cu:17-49): per baseline, the tiles' XYZ difference
// rotated to the pointing centre's (HA, Dec) at each timestep's LMST.
__global__ void k_xyzs_to_uvws(const XYZ *__restrict__ xyzs, const double *__restrict__ lmsts, UVW *__restrict__ uvws,
                               RADec pc, int ntiles, int nbl, int ntimes) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbl) return;
    int t1, t2;
    triangle_pair(b, ntiles, t1, t2);
    double s_dec, c_dec;
    sincos(pc.dec, &s_dec, &c_dec);
    for (int t = 0; t < ntimes; ++t) {
        const XYZ a = xyzs[(size_t)t * ntiles + t1], c = xyzs[(size_t)t * ntiles + t2];
        const double x = a.x - c.x, y = a.y - c.y, z = a.z - c.z;
        double s_ha, c_ha;
        sincos(lmsts[t] - pc.ra, &s_ha, &c_ha);
        UVW o;
        o.u = s_ha * x + c_ha * y;
        o.v = -s_dec * c_ha * x + s_dec * s_ha * y + c_dec * z;
        o.w = c_dec * c_ha * x - c_dec * s_ha * y + s_dec * z;
        uvws[(size_t)t * nbl + b] = o;
    }
}

// rotate_kernel (src/gpu/peel.cu:56-103): vis *= exp(-i tau (w_to - w_from) / lambda), the
// product formed in FP64 from the FP32 visibilities and rounded back to FP32.
__global__ void k_rotate(JonesF32 *__restrict__ vis, int ntimes, int nbl, int nfreqs, const UVW *__restrict__ from,
                         const UVW *__restrict__ to, const double *__restrict__ lambdas) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)ntimes * nfreqs * nbl) return;
    const int b = (int)(i % nbl);
    const size_t tf = i / nbl;
    const int f = (int)(tf % nfreqs);
    const int t = (int)(tf / nfreqs);
    const double arg = -kTauPeel * (to[(size_t)t * nbl + b].w - from[(size_t)t * nbl + b].w);
    double s, c;
    sincos(arg / lambdas[f], &s, &c);
    float *v = reinterpret_cast<float *>(&vis[i]);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const double re = v[2 * e], im = v[2 * e + 1];
        v[2 * e] = (float)(re * c - im * s);
        v[2 * e + 1] = (float)(re * s + im * c);
    }
}

// add_model_kernel (src/gpu/peel.cu:559-592): residual += model * (exp(-i tau (u alpha +
// v beta) lambda) gain). The factor is formed in FP64 and rounded to FP32; the complex product
// and the sum are FP32.
__global__ void k_add_model(JonesF32 *__restrict__ resid, const JonesF32 *__restrict__ model, IonoConsts ic,
                            const double *__restrict__ lambdas, const UVW *__restrict__ uvws, int ntimes, int nfreqs,
                            int nbl) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)ntimes * nfreqs * nbl) return;
    const int b = (int)(i % nbl);
    const size_t tf = i / nbl;
    const int f = (int)(tf % nfreqs);
    const int t = (int)(tf / nfreqs);
    double c = 1.0, s = 0.0;
    if (ic.alpha != 0.0 || ic.beta != 0.0) {
        const UVW uvw = uvws[(size_t)t * nbl + b];
        const double arg = -kTauPeel * (uvw.u * ic.alpha + uvw.v * ic.beta);
        sincos(arg * lambdas[f], &s, &c);
    }
    const float cr = (float)(c * ic.gain), ci = (float)(s * ic.gain);
    const float *m = reinterpret_cast<const float *>(&model[i]);
    float *r = reinterpret_cast<float *>(&resid[i]);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const float mr = m[2 * e], mi = m[2 * e + 1];
        r[2 * e] += mr * cr - mi * ci;
        r[2 * e + 1] += mr * ci + mi * cr;
    }
}

unsigned blocks_for(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace
}  // namespace sm

using namespace sm;

extern "C" const char *xyzs_to_uvws(const XYZ *d_xyzs, const double *d_lmsts, UVW *d_uvws, RADec pointing_centre,
                                    int num_tiles, int num_baselines, int num_timesteps) {
    if (num_baselines <= 0 || num_timesteps <= 0) return nullptr;
    if (num_tiles < 2 || (long long)num_tiles * (num_tiles - 1) / 2 != num_baselines)
        return "num_baselines must be num_tiles (num_tiles - 1) / 2";
    if (d_xyzs == nullptr || d_lmsts == nullptr || d_uvws == nullptr) return "null argument";
    hipLaunchKernelGGL(k_xyzs_to_uvws, dim3(blocks_for((size_t)num_baselines)), dim3(256), 0, 0, d_xyzs, d_lmsts,
                       d_uvws, pointing_centre, num_tiles, num_baselines, num_timesteps);
    SM_TRY(hipGetLastError());
    return nullptr;
}

extern "C" const char *rotate(JonesF32 *d_vis_fb, const int num_timesteps, const int num_baselines,
                              const int num_freqs, const UVW *d_uvws_from, const UVW *d_uvws_to,
                              const double *d_lambdas) {
    const size_t n = (size_t)(num_timesteps > 0 ? num_timesteps : 0) * (num_freqs > 0 ? num_freqs : 0) *
                     (num_baselines > 0 ? num_baselines : 0);
    if (n == 0) return nullptr;
    if (d_vis_fb == nullptr || d_uvws_from == nullptr || d_uvws_to == nullptr || d_lambdas == nullptr)
        return "null argument";
    hipLaunchKernelGGL(k_rotate, dim3(blocks_for(n)), dim3(256), 0, 0, d_vis_fb, num_timesteps, num_baselines,
                       num_freqs, d_uvws_from, d_uvws_to, d_lambdas);
    SM_TRY(hipGetLastError());
    return nullptr;
}

extern "C" const char *add_model(JonesF32 *d_vis_residual, const JonesF32 *d_vis_model, const IonoConsts iono_consts,
                                 const double *d_lambdas_m, const UVW *d_uvws, const int num_timesteps,
                                 const int num_baselines, const int num_freqs) {
    const size_t n = (size_t)(num_timesteps > 0 ? num_timesteps : 0) * (num_freqs > 0 ? num_freqs : 0) *
                     (num_baselines > 0 ? num_baselines : 0);
    if (n == 0) return nullptr;
    if (d_vis_residual == nullptr || d_vis_model == nullptr || d_lambdas_m == nullptr || d_uvws == nullptr)
        return "null argument";
    hipLaunchKernelGGL(k_add_model, dim3(blocks_for(n)), dim3(256), 0, 0, d_vis_residual, d_vis_model, iono_consts,
                       d_lambdas_m, d_uvws, num_timesteps, num_freqs, num_baselines);
    SM_TRY(hipGetLastError());
    return nullptr;
}
