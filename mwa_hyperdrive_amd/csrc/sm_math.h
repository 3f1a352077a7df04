// FP64 sin/cos for the measurement-equation phases.
//
// The phases reach |x| ~ 1e4 rad, far inside the range where a Cody-Waite reduction
// with a two-part pi/2 and fused multiply-adds is exact to within one rounding:
//   k = rint(x * 2/pi),  r = fma(-k, P1, x) - k P2,  |r| <= pi/4 (1 + eps),
// P1 = pi/2 rounded to double, P2 = (pi/2 - P1) rounded to double (the third part,
// ~1.5e-33, changes r by less than 1e-28 for |k| < 1e5 and is dropped). Both constants
// and 2/pi were computed from a 78-digit pi (Machin's formula). On |r| <= pi/4 the Taylor
// series truncated after r^19 (sin) and r^20 (cos) leave remainders below 1e-19, so the
// result is limited by the Horner rounding (about 1 ulp). Arguments up to |x| = 1e9 keep
// the reduction error below 1e-16 absolute; the kernels' phases stay far below that.
#pragma once

#ifdef __HIPCC__
#define SM_HD __host__ __device__
#else
#define SM_HD
#endif

#include <cmath>

namespace sm {

SM_HD inline void fast_sincos(double x, double &s, double &c) {
    const double k = rint(x * 0.6366197723675814);
    double r = fma(-k, 1.5707963267948966, x);
    r = fma(-k, 6.123233995736766e-17, r);
    const double r2 = r * r;
    // sin r = r + r^3 (-1/3! + r^2 (1/5! - ...)), coefficients (-1)^n / (2n+1)!
    double ps = -8.22063524662433e-18;
    ps = fma(ps, r2, 2.8114572543455206e-15);
    ps = fma(ps, r2, -7.647163731819816e-13);
    ps = fma(ps, r2, 1.6059043836821613e-10);
    ps = fma(ps, r2, -2.505210838544172e-08);
    ps = fma(ps, r2, 2.7557319223985893e-06);
    ps = fma(ps, r2, -0.0001984126984126984);
    ps = fma(ps, r2, 0.008333333333333333);
    ps = fma(ps, r2, -0.16666666666666666);
    const double sn = fma(r * r2, ps, r);
    // cos r = 1 + r^2 (-1/2! + r^2 (1/4! - ...)), coefficients (-1)^n / (2n)!
    double pc = 4.110317623312165e-19;
    pc = fma(pc, r2, -1.5619206968586225e-16);
    pc = fma(pc, r2, 4.779477332387385e-14);
    pc = fma(pc, r2, -1.1470745597729725e-11);
    pc = fma(pc, r2, 2.08767569878681e-09);
    pc = fma(pc, r2, -2.755731922398589e-07);
    pc = fma(pc, r2, 2.48015873015873e-05);
    pc = fma(pc, r2, -0.001388888888888889);
    pc = fma(pc, r2, 0.041666666666666664);
    pc = fma(pc, r2, -0.5);
    const double cn = fma(r2, pc, 1.0);
    const int q = (int)(long long)k & 3;
    const double a = (q & 1) ? cn : sn;  // sin x = sin r, cos r, -sin r, -cos r for q = 0..3
    const double b = (q & 1) ? sn : cn;  // cos x = cos r, -sin r, -cos r, sin r
    s = (q & 2) ? -a : a;
    c = ((q + 1) & 2) ? -b : b;
}

// (sin, cos) of k pi / 32 for k = 0 .. 63, each rounded to nearest from a 300-bit evaluation.
#define SM_SINCOS32_TABLE \
    0.0, 1.0, 0.0980171403295606, 0.9951847266721969, \
    0.19509032201612828, 0.9807852804032304, 0.2902846772544624, 0.9569403357322088, \
    0.3826834323650898, 0.9238795325112867, 0.47139673682599764, 0.881921264348355, \
    0.5555702330196022, 0.8314696123025452, 0.6343932841636455, 0.773010453362737, \
    0.7071067811865476, 0.7071067811865476, 0.773010453362737, 0.6343932841636455, \
    0.8314696123025452, 0.5555702330196022, 0.881921264348355, 0.47139673682599764, \
    0.9238795325112867, 0.3826834323650898, 0.9569403357322088, 0.2902846772544624, \
    0.9807852804032304, 0.19509032201612828, 0.9951847266721969, 0.0980171403295606, \
    1.0, 0.0, 0.9951847266721969, -0.0980171403295606, \
    0.9807852804032304, -0.19509032201612828, 0.9569403357322088, -0.2902846772544624, \
    0.9238795325112867, -0.3826834323650898, 0.881921264348355, -0.47139673682599764, \
    0.8314696123025452, -0.5555702330196022, 0.773010453362737, -0.6343932841636455, \
    0.7071067811865476, -0.7071067811865476, 0.6343932841636455, -0.773010453362737, \
    0.5555702330196022, -0.8314696123025452, 0.47139673682599764, -0.881921264348355, \
    0.3826834323650898, -0.9238795325112867, 0.2902846772544624, -0.9569403357322088, \
    0.19509032201612828, -0.9807852804032304, 0.0980171403295606, -0.9951847266721969, \
    0.0, -1.0, -0.0980171403295606, -0.9951847266721969, \
    -0.19509032201612828, -0.9807852804032304, -0.2902846772544624, -0.9569403357322088, \
    -0.3826834323650898, -0.9238795325112867, -0.47139673682599764, -0.881921264348355, \
    -0.5555702330196022, -0.8314696123025452, -0.6343932841636455, -0.773010453362737, \
    -0.7071067811865476, -0.7071067811865476, -0.773010453362737, -0.6343932841636455, \
    -0.8314696123025452, -0.5555702330196022, -0.881921264348355, -0.47139673682599764, \
    -0.9238795325112867, -0.3826834323650898, -0.9569403357322088, -0.2902846772544624, \
    -0.9807852804032304, -0.19509032201612828, -0.9951847266721969, -0.0980171403295606, \
    -1.0, 0.0, -0.9951847266721969, 0.0980171403295606, \
    -0.9807852804032304, 0.19509032201612828, -0.9569403357322088, 0.2902846772544624, \
    -0.9238795325112867, 0.3826834323650898, -0.881921264348355, 0.47139673682599764, \
    -0.8314696123025452, 0.5555702330196022, -0.773010453362737, 0.6343932841636455, \
    -0.7071067811865476, 0.7071067811865476, -0.6343932841636455, 0.773010453362737, \
    -0.5555702330196022, 0.8314696123025452, -0.47139673682599764, 0.881921264348355, \
    -0.3826834323650898, 0.9238795325112867, -0.2902846772544624, 0.9569403357322088, \
    -0.19509032201612828, 0.9807852804032304, -0.0980171403295606, 0.9951847266721969

// FP64 sin/cos with a table-driven reduction (P. T. P. Tang, "Table-lookup algorithms for
// elementary functions and their error analysis", 10th IEEE Symposium on Computer Arithmetic,
// 1991): x = k pi/32 + r with |r| <= pi/64 (1 + eps), then
//   sin x = S_k cos r + C_k sin r,   cos x = C_k cos r - S_k sin r,
// where (S_k, C_k) = (sin, cos)(k pi/32) come from `tab` (SM_SINCOS32_TABLE, index k mod 64)
// and the Taylor series of sin r to r^9 and of cos r to r^8 leave remainders below 3e-20 at
// |r| = pi/64. The reduction is fast_sincos's with a two-part pi/32 (P1 = pi/32 rounded to
// double, P2 = the rest rounded to double, 32/pi rounded to double, all from a 300-bit pi).
// Nine polynomial terms instead of twenty and no quadrant selection, for one table read.
SM_HD inline void table_sincos(double x, const double *tab, double &s, double &c) {
    const double k = rint(x * 10.185916357881302);
    double r = fma(-k, 0.09817477042468103, x);
    r = fma(-k, 3.827021247335479e-18, r);
    const double r2 = r * r;
    // coefficients (-1)^n / (2n+1)! and (-1)^n / (2n)!
    double ps = 2.7557319223985893e-06;
    ps = fma(ps, r2, -0.0001984126984126984);
    ps = fma(ps, r2, 0.008333333333333333);
    ps = fma(ps, r2, -0.16666666666666666);
    const double sr = fma(r * r2, ps, r);
    double pc = 2.48015873015873e-05;
    pc = fma(pc, r2, -0.001388888888888889);
    pc = fma(pc, r2, 0.041666666666666664);
    pc = fma(pc, r2, -0.5);
    const double cr = fma(r2, pc, 1.0);
    const int i = (int)k & 63;  // k fits an int for |x| < 2e8, far beyond any phase here
    const double S = tab[2 * i], C = tab[2 * i + 1];
    s = fma(S, cr, C * sr);
    c = fma(C, cr, -(S * sr));
}

#ifdef __HIPCC__
// The device copy of SM_SINCOS32_TABLE for table_sincos, one per translation unit; 16-byte
// aligned, so a lane reads its (sin, cos) pair in one load.
[[maybe_unused]] static __device__ const double kSinCos32[128] __attribute__((aligned(16))) = {SM_SINCOS32_TABLE};
#endif

}  // namespace sm
