// 2x2 complex Jones helpers shared by the direct kernels.
#pragma once

#include "sm_internal.h"

namespace sm {

// One packed coefficient as a single 16-byte load (a scalar load when the address is
// wave-uniform).
// The same for a wave-uniform p, through the constant address space so that the compiler
// issues scalar loads (the coefficients are never written while a kernel runs).
__device__ __forceinline__ void load_packed_uniform(const PackedCoeff *p, double &value, int &n1, int &n2) {
    typedef int int4v __attribute__((ext_vector_type(4)));
    typedef __attribute__((address_space(4))) const int4v cint4v;
    const int4v w = *(const cint4v *)p;
    value = __hiloint2double(w.y, w.x);
    n1 = w.z;
    n2 = w.w;
}

__device__ __forceinline__ void load_packed(const PackedCoeff *p, double &value, int &n1, int &n2) {
    const int4 w = *reinterpret_cast<const int4 *>(p);
    value = __hiloint2double(w.y, w.x);
    n1 = w.z;
    n2 = w.w;
}

// One part (real or imaginary) of sum_k i^((n1 + n2) mod 4) c_k B_n1(x) B_n2(y), with B
// linearly interpolated at cell (ix, iy), fractions (wx, wy) (src/model/cpu.rs:633-695):
// the terms of that part in source order, their signs folded into the values. Table reads
// of four coefficients are issued before any of them is used.
__device__ __forceinline__ double shapelet_part(const PackedCoeff *__restrict__ cp, int count,
                                                const double *__restrict__ sbf, int ix, int iy, double wx,
                                                double wy) {
    double acc = 0.0;
    int j = 0;
    for (; j + 4 <= count; j += 4) {
        double val[4], xl[4], xh[4], yl[4], yh[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            int n1, n2;
            load_packed(cp + j + u, val[u], n1, n2);
            const int o1 = n1 * SM_SBF_L, o2 = n2 * SM_SBF_L;
            xl[u] = sbf[o1 + ix];
            xh[u] = sbf[o1 + ix + 1];
            yl[u] = sbf[o2 + iy];
            yh[u] = sbf[o2 + iy + 1];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double u_value = xl[u] + (xh[u] - xl[u]) * wx;
            const double v_value = yl[u] + (yh[u] - yl[u]) * wy;
            acc += (val[u] * u_value) * v_value;
        }
    }
    for (; j < count; ++j) {
        double val;
        int n1, n2;
        load_packed(cp + j, val, n1, n2);
        const int o1 = n1 * SM_SBF_L, o2 = n2 * SM_SBF_L;
        const double u_value = sbf[o1 + ix] + (sbf[o1 + ix + 1] - sbf[o1 + ix]) * wx;
        const double v_value = sbf[o2 + iy] + (sbf[o2 + iy + 1] - sbf[o2 + iy]) * wy;
        acc += (val * u_value) * v_value;
    }
    return acc;
}

// The envelope's real and imaginary parts for one component (range from launch_pack_coeffs).
__device__ __forceinline__ void shapelet_sum(const PackedCoeff *__restrict__ packed, int4 range,
                                             const double *__restrict__ sbf, int ix, int iy, double wx, double wy,
                                             double &er, double &ei) {
    er = shapelet_part(packed + range.x, range.y, sbf, ix, iy, wx, wy);
    ei = shapelet_part(packed + range.x + range.y, range.z, sbf, ix, iy, wx, wy);
}

__device__ __forceinline__ void load_jones(const JonesF64 *p, double j[8]) {
    const double *q = reinterpret_cast<const double *>(p);
#pragma unroll
    for (int e = 0; e < 8; ++e) j[e] = q[e];
}

// A class coherency J_r S J_0^H, or its conjugate transpose J_0 S J_r^H (S is Hermitian)
// when herm: entry (i, j) from (j, i), conjugated.
template <class T> __device__ __forceinline__ void class_coh(const T *p, bool herm, T o[8]) {
    T v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = p[e];
    o[0] = v[0];
    o[1] = herm ? -v[1] : v[1];
    o[2] = herm ? v[4] : v[2];
    o[3] = herm ? -v[5] : v[3];
    o[4] = herm ? v[2] : v[4];
    o[5] = herm ? -v[3] : v[5];
    o[6] = v[6];
    o[7] = herm ? -v[7] : v[7];
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
