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

}  // namespace sm
