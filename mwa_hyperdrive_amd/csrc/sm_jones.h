// 2x2 complex Jones helpers shared by the direct kernels.
#pragma once

#include "sm_internal.h"

namespace sm {

__device__ __forceinline__ void load_jones(const JonesF64 *p, double j[8]) {
    const double *q = reinterpret_cast<const double *>(p);
#pragma unroll
    for (int e = 0; e < 8; ++e) j[e] = q[e];
}

// (A S) B^H in the reference's order (src/model/cpu.rs:343; marlu Jones Mul and h()).
__device__ __forceinline__ void sandwich(const double a[8], const double s[8], const double b[8], double o[8]) {
    double t[8];
    // t = A S
    t[0] = (a[0] * s[0] - a[1] * s[1]) + (a[2] * s[4] - a[3] * s[5]);
    t[1] = (a[0] * s[1] + a[1] * s[0]) + (a[2] * s[5] + a[3] * s[4]);
    t[2] = (a[0] * s[2] - a[1] * s[3]) + (a[2] * s[6] - a[3] * s[7]);
    t[3] = (a[0] * s[3] + a[1] * s[2]) + (a[2] * s[7] + a[3] * s[6]);
    t[4] = (a[4] * s[0] - a[5] * s[1]) + (a[6] * s[4] - a[7] * s[5]);
    t[5] = (a[4] * s[1] + a[5] * s[0]) + (a[6] * s[5] + a[7] * s[4]);
    t[6] = (a[4] * s[2] - a[5] * s[3]) + (a[6] * s[6] - a[7] * s[7]);
    t[7] = (a[4] * s[3] + a[5] * s[2]) + (a[6] * s[7] + a[7] * s[6]);
    // B^H = [conj b00, conj b10; conj b01, conj b11]
    const double h00r = b[0], h00i = -b[1], h01r = b[4], h01i = -b[5];
    const double h10r = b[2], h10i = -b[3], h11r = b[6], h11i = -b[7];
    o[0] = (t[0] * h00r - t[1] * h00i) + (t[2] * h10r - t[3] * h10i);
    o[1] = (t[0] * h00i + t[1] * h00r) + (t[2] * h10i + t[3] * h10r);
    o[2] = (t[0] * h01r - t[1] * h01i) + (t[2] * h11r - t[3] * h11i);
    o[3] = (t[0] * h01i + t[1] * h01r) + (t[2] * h11i + t[3] * h11r);
    o[4] = (t[4] * h00r - t[5] * h00i) + (t[6] * h10r - t[7] * h10i);
    o[5] = (t[4] * h00i + t[5] * h00r) + (t[6] * h10i + t[7] * h10r);
    o[6] = (t[4] * h01r - t[5] * h01i) + (t[6] * h11r - t[7] * h11i);
    o[7] = (t[4] * h01i + t[5] * h01r) + (t[6] * h11i + t[7] * h11r);
}


// Aliases used by the tiled kernels.
__device__ __forceinline__ void load_jones_t(const JonesF64 *p, double j[8]) { load_jones(p, j); }
__device__ __forceinline__ void sandwich_t(const double a[8], const double s[8], const double b[8], double o[8]) {
    sandwich(a, s, b, o);
}

}  // namespace sm
