// Epoch handling in front of the modelling path: GPS time -> UTC / UT1 / TT, the local mean
// sidereal time, and the rotation of the array to the J2000 frame. The reference does this on
// the host in SkyModeller*::get_lst_uvws_latitude (src/model/cpu.rs:731-769,
// src/model/gpu.rs:1357-1398) through marlu 0.17.1's precession::{get_lmst, precess_time} and
// PrecessionInfo::precess_xyz (a dependency outside the reference tree). This file restates the
// published methods those follow:
//   - GPS = TAI - 19 s, TT = TAI + 32.184 s, UTC from the IERS leap-second table, UT1 = UTC + DUT1;
//   - Earth rotation angle (IAU 2000) and GMST (IAU 2006, IERS Conventions 2010 eq. 5.32);
//   - IAU 2006 precession angles zeta_A, z_A, theta_A with the IERS 2010 frame bias and the
//     IAU 2006 mean obliquity;
//   - nutation from the 18 largest terms of the IAU 1980 series (the rest are below 5 mas each);
//   - annual aberration of the phase centre, first order in v/c, Earth velocity from the
//     eccentric orbit.
// Host code only: a few hundred flops per timestep.
#include <cmath>
#include <cstring>

#include "skymodel.h"
#include "sm_time.h"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTau = 2.0 * kPi;
constexpr double kAs2R = kPi / 180.0 / 3600.0;
constexpr double kGpsEpochJd = 2444244.5;  // 1980-01-06T00:00:00 UTC
constexpr double kJ2000Jd = 2451545.0;

struct Mat {
    double a[3][3];
};

Mat mul(const Mat &x, const Mat &y) {
    Mat r{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += x.a[i][k] * y.a[k][j];
            r.a[i][j] = s;
        }
    return r;
}

Mat transpose(const Mat &x) {
    Mat r{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.a[i][j] = x.a[j][i];
    return r;
}

void apply(const Mat &m, const double v[3], double out[3]) {
    for (int i = 0; i < 3; ++i) out[i] = m.a[i][0] * v[0] + m.a[i][1] * v[1] + m.a[i][2] * v[2];
}

// Frame rotations about x, y, z by angle t (a vector's coordinates in the rotated frame).
Mat rot1(double t) {
    const double c = std::cos(t), s = std::sin(t);
    return Mat{{{1, 0, 0}, {0, c, s}, {0, -s, c}}};
}
Mat rot2(double t) {
    const double c = std::cos(t), s = std::sin(t);
    return Mat{{{c, 0, -s}, {0, 1, 0}, {s, 0, c}}};
}
Mat rot3(double t) {
    const double c = std::cos(t), s = std::sin(t);
    return Mat{{{c, s, 0}, {-s, c, 0}, {0, 0, 1}}};
}

double wrap(double x) {
    x = std::fmod(x, kTau);
    return x < 0.0 ? x + kTau : x;
}

// JD at 0h of a Gregorian date.
double julian_day(int y, int m, int d) {
    const int a = (14 - m) / 12, yy = y + 4800 - a, mm = m + 12 * a - 3;
    return d + (153 * mm + 2) / 5 + 365 * yy + yy / 4 - yy / 100 + yy / 400 - 32045 - 0.5;
}

// TAI - UTC from the UTC date it takes effect (IERS Bulletin C), from the GPS epoch on.
struct Leap {
    int y, m, d, tai_utc;
};
constexpr Leap kLeaps[] = {{1981, 7, 1, 20}, {1982, 7, 1, 21}, {1983, 7, 1, 22}, {1985, 7, 1, 23}, {1988, 1, 1, 24},
                           {1990, 1, 1, 25}, {1991, 1, 1, 26}, {1992, 7, 1, 27}, {1993, 7, 1, 28}, {1994, 7, 1, 29},
                           {1996, 1, 1, 30}, {1997, 7, 1, 31}, {1999, 1, 1, 32}, {2006, 1, 1, 33}, {2009, 1, 1, 34},
                           {2012, 7, 1, 35}, {2015, 7, 1, 36}, {2017, 1, 1, 37}};

int tai_minus_utc(double gps) {
    int out = 19;
    for (const Leap &l : kLeaps)
        if (gps >= (julian_day(l.y, l.m, l.d) - kGpsEpochJd) * 86400.0 + (l.tai_utc - 19)) out = l.tai_utc;
    return out;
}

// Days since JD 2451545.0 on the UTC scale, and TT Julian centuries since J2000.0.
double utc_days(double gps) { return (gps - (tai_minus_utc(gps) - 19)) / 86400.0 - (kJ2000Jd - kGpsEpochJd); }
double tt_centuries(double gps) { return ((gps + 19.0 + 32.184) / 86400.0 - (kJ2000Jd - kGpsEpochJd)) / 36525.0; }

double gmst06(double tu, double t) {
    const double frac = tu - std::floor(tu);
    const double era = kTau * std::fmod(frac + 0.7790572732640 + 0.00273781191135448 * tu, 1.0);
    const double poly =
        (0.014506 + (4612.156534 + (1.3915817 + (-0.00000044 + (-0.000029956 + (-0.0000000368) * t) * t) * t) * t) * t) *
        kAs2R;
    return wrap(era + poly);
}

double lmst_at(double lon, double gps, double dut1) {
    return wrap(gmst06(utc_days(gps) + dut1 / 86400.0, tt_centuries(gps)) + lon);
}

double mean_obliquity(double t) {
    return (84381.406 + (-46.836769 + (-0.0001831 + (0.00200340 + (-0.000000576 - 0.0000000434 * t) * t) * t) * t) * t) *
           kAs2R;
}

// IAU 2006 precession, mean J2000 -> mean of date.
Mat precession(double t) {
    const double zeta =
        (2.650545 + (2306.083227 + (0.2988499 + (0.01801828 + (-0.000005971 - 0.0000003173 * t) * t) * t) * t) * t) * kAs2R;
    const double z =
        (-2.650545 + (2306.077181 + (1.0927348 + (0.01826837 + (-0.000028596 - 0.0000002904 * t) * t) * t) * t) * t) * kAs2R;
    const double theta = ((2004.191903 + (-0.4294934 + (-0.04182264 + (-0.000007089 - 0.0000001274 * t) * t) * t) * t) * t) *
                         kAs2R;
    return mul(rot3(-z), mul(rot2(theta), rot3(-zeta)));
}

// GCRS -> mean J2000 (xi0, eta0, d alpha0).
Mat frame_bias() { return mul(rot1(0.0068192 * kAs2R), mul(rot2(-0.0166170 * kAs2R), rot3(-0.01460 * kAs2R))); }

// Multipliers of (D, M, M', F, Omega); d psi = (a + b t) sin, d eps = (c + d t) cos, 0.0001".
struct NutTerm {
    signed char d, m, mp, f, om;
    double a, b, c, e;
};
constexpr NutTerm kNut[] = {
    {0, 0, 0, 0, 1, -171996.0, -174.2, 92025.0, 8.9}, {-2, 0, 0, 2, 2, -13187.0, -1.6, 5736.0, -3.1},
    {0, 0, 0, 2, 2, -2274.0, -0.2, 977.0, -0.5},      {0, 0, 0, 0, 2, 2062.0, 0.2, -895.0, 0.5},
    {0, 1, 0, 0, 0, 1426.0, -3.4, 54.0, -0.1},        {0, 0, 1, 0, 0, 712.0, 0.1, -7.0, 0.0},
    {-2, 1, 0, 2, 2, -517.0, 1.2, 224.0, -0.6},       {0, 0, 0, 2, 1, -386.0, -0.4, 200.0, 0.0},
    {0, 0, 1, 2, 2, -301.0, 0.0, 129.0, -0.1},        {-2, -1, 0, 2, 2, 217.0, -0.5, -95.0, 0.3},
    {-2, 0, 1, 0, 0, -158.0, 0.0, 0.0, 0.0},          {-2, 0, 0, 2, 1, 129.0, 0.1, -70.0, 0.0},
    {0, 0, -1, 2, 2, 123.0, 0.0, -53.0, 0.0},         {2, 0, 0, 0, 0, 63.0, 0.0, 0.0, 0.0},
    {0, 0, 1, 0, 1, 63.0, 0.1, -33.0, 0.0},           {2, 0, -1, 2, 2, -59.0, 0.0, 26.0, 0.0},
    {0, 0, -1, 0, 1, -58.0, -0.1, 32.0, 0.0},         {0, 0, 1, 2, 1, -51.0, 0.0, 27.0, 0.0},
};

Mat nutation(double t) {
    const double deg = kPi / 180.0;
    const double D = (297.85036 + 445267.111480 * t - 0.0019142 * t * t + t * t * t / 189474.0) * deg;
    const double M = (357.52772 + 35999.050340 * t - 0.0001603 * t * t - t * t * t / 300000.0) * deg;
    const double Mp = (134.96298 + 477198.867398 * t + 0.0086972 * t * t + t * t * t / 56250.0) * deg;
    const double F = (93.27191 + 483202.017538 * t - 0.0036825 * t * t + t * t * t / 327270.0) * deg;
    const double Om = (125.04452 - 1934.136261 * t + 0.0020708 * t * t + t * t * t / 450000.0) * deg;
    double dpsi = 0.0, deps = 0.0;
    for (const NutTerm &n : kNut) {
        const double arg = n.d * D + n.m * M + n.mp * Mp + n.f * F + n.om * Om;
        dpsi += (n.a + n.b * t) * std::sin(arg);
        deps += (n.c + n.e * t) * std::cos(arg);
    }
    dpsi *= 1e-4 * kAs2R;
    deps *= 1e-4 * kAs2R;
    const double eps = mean_obliquity(t);
    return mul(rot1(-(eps + deps)), mul(rot3(-dpsi), rot1(eps)));
}

// The Earth's orbital velocity over c in the J2000 frame.
void earth_velocity(double t, const Mat &pb, double v[3]) {
    const double deg = kPi / 180.0;
    const double L0 = 280.46646 + 36000.76983 * t + 0.0003032 * t * t;
    const double M = (357.52911 + 35999.05029 * t - 0.0001537 * t * t) * deg;
    const double C = (1.914602 - 0.004817 * t - 0.000014 * t * t) * std::sin(M) +
                     (0.019993 - 0.000101 * t) * std::sin(2 * M) + 0.000289 * std::sin(3 * M);
    const double L = (L0 + C) * deg;
    const double e = 0.016708634 - 0.000042037 * t - 0.0000001267 * t * t;
    const double pi = (102.93735 + 1.71946 * t + 0.00046 * t * t) * deg;
    const double kappa = 20.49552 * kAs2R;
    const double v_ecl[3] = {kappa * (std::sin(L) - e * std::sin(pi)), kappa * (-std::cos(L) + e * std::cos(pi)), 0.0};
    double v_date[3];
    apply(rot1(-mean_obliquity(t)), v_ecl, v_date);
    apply(transpose(pb), v_date, v);
}

void unit(double ra, double dec, double p[3]) {
    p[0] = std::cos(dec) * std::cos(ra);
    p[1] = std::cos(dec) * std::sin(ra);
    p[2] = std::sin(dec);
}

void to_radec(const double p[3], double &ra, double &dec) {
    const double n = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    ra = wrap(std::atan2(p[1], p[0]));
    double z = p[2] / n;
    dec = std::asin(z > 1.0 ? 1.0 : (z < -1.0 ? -1.0 : z));
}

}  // namespace

namespace sm {
int tai_minus_utc_at(double gps_seconds) { return tai_minus_utc(gps_seconds); }
double utc_jd_of_gps(double gps_seconds) { return utc_days(gps_seconds) + kJ2000Jd; }
}  // namespace sm

extern "C" double sm_get_lmst(double array_longitude_rad, double gps_seconds, double dut1_seconds) {
    return lmst_at(array_longitude_rad, gps_seconds, dut1_seconds);
}

extern "C" const char *sm_precess_time(double array_longitude_rad, double array_latitude_rad, double pc_ra,
                                       double pc_dec, double gps_seconds, double dut1_seconds,
                                       SmPrecessionInfo *out) {
    if (out == nullptr) return "null argument";
    if (!std::isfinite(gps_seconds) || !std::isfinite(dut1_seconds)) return "non-finite time";
    const double t = tt_centuries(gps_seconds);
    const double lmst = lmst_at(array_longitude_rad, gps_seconds, dut1_seconds);
    const Mat pb = mul(precession(t), frame_bias());
    const Mat rot = transpose(mul(nutation(t), pb));  // true of date -> J2000
    double z[3], zj[3];
    unit(lmst, array_latitude_rad, z);
    apply(rot, z, zj);
    double lmst_j2000, lat_j2000;
    to_radec(zj, lmst_j2000, lat_j2000);
    // Aberrated phase centre: p + v - (p . v) p.
    double p[3], v[3];
    unit(pc_ra, pc_dec, p);
    earth_velocity(t, pb, v);
    const double pv = p[0] * v[0] + p[1] * v[1] + p[2] * v[2];
    const double q[3] = {p[0] + v[0] - pv * p[0], p[1] + v[1] - pv * p[1], p[2] + v[2] - pv * p[2]};
    double ra_ab, dec_ab;
    to_radec(q, ra_ab, dec_ab);
    std::memcpy(out->rotation_matrix, rot.a, sizeof(rot.a));
    out->hadec_j2000 = HADec{wrap(lmst_j2000 - ra_ab), dec_ab};
    out->lmst = lmst;
    out->lmst_j2000 = lmst_j2000;
    out->array_latitude_j2000 = lat_j2000;
    return nullptr;
}

extern "C" const char *sm_precess_xyz(const SmPrecessionInfo *info, const XYZ *xyzs, XYZ *out, int32_t n) {
    if (n <= 0) return nullptr;
    if (info == nullptr || xyzs == nullptr || out == nullptr) return "null argument";
    const double c = std::cos(info->lmst), s = std::sin(info->lmst);
    const double c2 = std::cos(info->lmst_j2000), s2 = std::sin(info->lmst_j2000);
    const double(*r)[3] = info->rotation_matrix;
    for (int32_t i = 0; i < n; ++i) {
        const XYZ a = xyzs[i];
        // Local frame at lmst -> celestial frame of date -> J2000 -> local frame at lmst_j2000.
        const double cel[3] = {c * a.x - s * a.y, s * a.x + c * a.y, a.z};
        double j[3];
        for (int k = 0; k < 3; ++k) j[k] = r[k][0] * cel[0] + r[k][1] * cel[1] + r[k][2] * cel[2];
        out[i] = XYZ{c2 * j[0] + s2 * j[1], -s2 * j[0] + c2 * j[1], j[2]};
    }
    return nullptr;
}
