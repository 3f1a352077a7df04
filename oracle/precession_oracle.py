"""Test infrastructure: a CPU restatement of the epoch handling in front of the modelling path.
Only the tests (and the bench's CPU baseline) use it, as the checker.

The reference turns a timestamp into the modeller's (LST, tile positions, latitude) in
SkyModellerCpu/Gpu::get_lst_uvws_latitude (src/model/cpu.rs:731-769,
src/model/gpu.rs:1357-1398) through marlu 0.17.1's `precession::{get_lmst, precess_time}` and
`PrecessionInfo::precess_xyz`, with hifitime's GPS/UTC/TT scales. Neither crate is in the
reference tree, so this module restates the published methods they follow:

- time scales: GPS = TAI - 19 s, TT = TAI + 32.184 s, UTC from the IERS leap-second table
  (Bulletin C), UT1 = UTC + DUT1;
- Earth rotation angle (IAU 2000, Capitaine et al. 2000) and GMST (IAU 2006, IERS Conventions
  2010 eq. 5.32);
- precession: IAU 2006 (P03) angles zeta_A, z_A, theta_A, P = R3(-z_A) R2(theta_A) R3(-zeta_A),
  with the IERS 2010 frame bias (xi0, eta0, d alpha0) and the IAU 2006 mean obliquity;
- nutation: the 18 largest terms of the IAU 1980 series (amplitudes and Delaunay arguments as
  tabulated in Meeus, Astronomical Algorithms, ch. 22), N = R1(-(eps + d eps)) R3(-d psi)
  R1(eps). The terms left out are below 5 mas each;
- annual aberration of the phase centre to first order in v/c, with the Earth's velocity from
  the eccentric orbit (kappa = 20.49552", Meeus ch. 23 and ch. 25 for the Sun's longitude).

Pins (tests/test_oracle_precession.py):
- the reference's vis-simulate uvfits group parameters (src/cli/vis_simulate/tests.rs:235-316:
  first and last cross-correlation u, v, w of observation 1090008640), which the uvfits writer
  forms from these tile positions and `hadec_j2000`. Agreement is within 1e-7 of the UVW
  magnitude, the precision the file stores them at. Precession alone misses by 7.6e-5, and
  precession with nutation but no aberration misses by 4.9e-5;
- Meeus's worked examples 21.b (precession), 22.a (nutation) and 23.a (nutation and
  aberration in RA/Dec);
- invariants: orthonormal matrices, P = 1 at J2000.0, the sidereal rate, and the round trip
  through precess_xyz.

`lmst_j2000` and `array_latitude_j2000` (the J2000 coordinates of the zenith) are what the
model uses for az/el. The uvfits numbers do not constrain them: they cancel from the UVWs.
They come from the same matrix, which is pinned. Whether marlu also aberrates the zenith is
parity unpinned (this restatement does not).
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

AS2R = math.pi / 180.0 / 3600.0
TAU = 2.0 * math.pi
GPS_EPOCH_JD = 2444244.5          # 1980-01-06T00:00:00 UTC
J2000_JD = 2451545.0
DAYS_GPS_EPOCH_TO_J2000 = J2000_JD - GPS_EPOCH_JD

# TAI - UTC from the UTC date it takes effect (IERS Bulletin C), from the GPS epoch on.
LEAP_SECONDS = [
    ((1981, 7, 1), 20), ((1982, 7, 1), 21), ((1983, 7, 1), 22), ((1985, 7, 1), 23), ((1988, 1, 1), 24),
    ((1990, 1, 1), 25), ((1991, 1, 1), 26), ((1992, 7, 1), 27), ((1993, 7, 1), 28), ((1994, 7, 1), 29),
    ((1996, 1, 1), 30), ((1997, 7, 1), 31), ((1999, 1, 1), 32), ((2006, 1, 1), 33), ((2009, 1, 1), 34),
    ((2012, 7, 1), 35), ((2015, 7, 1), 36), ((2017, 1, 1), 37),
]


def julian_day(y: int, m: int, d: int) -> float:
    """JD at 0h of a Gregorian date (Fliegel and Van Flandern)."""
    a = (14 - m) // 12
    yy = y + 4800 - a
    mm = m + 12 * a - 3
    return d + (153 * mm + 2) // 5 + 365 * yy + yy // 4 - yy // 100 + yy // 400 - 32045 - 0.5


def tai_minus_utc(gps_seconds: float) -> int:
    """TAI - UTC at a GPS time: the step at each table date happens at that date's
    GPS time, (days since the GPS epoch) 86400 + (new TAI - UTC) - 19."""
    out = 19
    for (y, m, d), ls in LEAP_SECONDS:
        if gps_seconds >= (julian_day(y, m, d) - GPS_EPOCH_JD) * 86400.0 + (ls - 19):
            out = ls
    return out


def utc_days_since_j2000(gps_seconds: float) -> float:
    """UTC as days since JD 2451545.0 (UTC), for a GPS time."""
    return (gps_seconds - (tai_minus_utc(gps_seconds) - 19)) / 86400.0 - DAYS_GPS_EPOCH_TO_J2000


def tt_centuries(gps_seconds: float) -> float:
    """TT Julian centuries since J2000.0 (TT = GPS + 19 s + 32.184 s)."""
    return ((gps_seconds + 19.0 + 32.184) / 86400.0 - DAYS_GPS_EPOCH_TO_J2000) / 36525.0


def earth_rotation_angle(tu: float) -> float:
    """ERA = 2 pi (0.7790572732640 + 1.00273781191135448 Tu), Tu = UT1 days since J2000.0;
    the whole days of Tu are dropped first to keep the fraction's precision."""
    f = tu - math.floor(tu)
    return (TAU * ((f + 0.7790572732640 + 0.00273781191135448 * tu) % 1.0)) % TAU


def gmst06(tu: float, t_tt: float) -> float:
    """IAU 2006 Greenwich mean sidereal time from ERA (UT1) and TT centuries."""
    poly = (0.014506 + (4612.156534 + (1.3915817 + (-0.00000044 + (-0.000029956 + (-0.0000000368) * t_tt) * t_tt)
                                       * t_tt) * t_tt) * t_tt) * AS2R
    return (earth_rotation_angle(tu) + poly) % TAU


def get_lmst(array_longitude_rad: float, gps_seconds: float, dut1_seconds: float = 0.0) -> float:
    """marlu::precession::get_lmst: GMST at the epoch's UT1 plus the longitude, in [0, 2 pi)."""
    tu = utc_days_since_j2000(gps_seconds) + dut1_seconds / 86400.0
    return (gmst06(tu, tt_centuries(gps_seconds)) + array_longitude_rad) % TAU


def r1(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[1.0, 0.0, 0.0], [0.0, c, s], [0.0, -s, c]])


def r2(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, 0.0, -s], [0.0, 1.0, 0.0], [s, 0.0, c]])


def r3(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, s, 0.0], [-s, c, 0.0], [0.0, 0.0, 1.0]])


def precession_matrix(t: float) -> np.ndarray:
    """IAU 2006 precession, mean J2000 equator/equinox -> mean of date."""
    zeta = (2.650545 + (2306.083227 + (0.2988499 + (0.01801828 + (-0.000005971 - 0.0000003173 * t) * t) * t) * t)
            * t) * AS2R
    z = (-2.650545 + (2306.077181 + (1.0927348 + (0.01826837 + (-0.000028596 - 0.0000002904 * t) * t) * t) * t)
         * t) * AS2R
    theta = ((2004.191903 + (-0.4294934 + (-0.04182264 + (-0.000007089 - 0.0000001274 * t) * t) * t) * t) * t) * AS2R
    return r3(-z) @ r2(theta) @ r3(-zeta)


def frame_bias_matrix() -> np.ndarray:
    """GCRS -> mean J2000 frame bias (IERS Conventions 2010: xi0, eta0, d alpha0)."""
    xi0, eta0, da0 = -0.0166170 * AS2R, -0.0068192 * AS2R, -0.01460 * AS2R
    return r1(-eta0) @ r2(xi0) @ r3(da0)


def mean_obliquity(t: float) -> float:
    """IAU 2006 mean obliquity of the ecliptic."""
    return (84381.406 + (-46.836769 + (-0.0001831 + (0.00200340 + (-0.000000576 - 0.0000000434 * t) * t) * t) * t)
            * t) * AS2R


# The 18 largest IAU 1980 nutation terms: multipliers of (D, M, M', F, Omega), then
# d psi = (a + b t) sin(arg) and d eps = (c + d t) cos(arg), in units of 0.0001".
NUTATION_TERMS = [
    (0, 0, 0, 0, 1, -171996.0, -174.2, 92025.0, 8.9),
    (-2, 0, 0, 2, 2, -13187.0, -1.6, 5736.0, -3.1),
    (0, 0, 0, 2, 2, -2274.0, -0.2, 977.0, -0.5),
    (0, 0, 0, 0, 2, 2062.0, 0.2, -895.0, 0.5),
    (0, 1, 0, 0, 0, 1426.0, -3.4, 54.0, -0.1),
    (0, 0, 1, 0, 0, 712.0, 0.1, -7.0, 0.0),
    (-2, 1, 0, 2, 2, -517.0, 1.2, 224.0, -0.6),
    (0, 0, 0, 2, 1, -386.0, -0.4, 200.0, 0.0),
    (0, 0, 1, 2, 2, -301.0, 0.0, 129.0, -0.1),
    (-2, -1, 0, 2, 2, 217.0, -0.5, -95.0, 0.3),
    (-2, 0, 1, 0, 0, -158.0, 0.0, 0.0, 0.0),
    (-2, 0, 0, 2, 1, 129.0, 0.1, -70.0, 0.0),
    (0, 0, -1, 2, 2, 123.0, 0.0, -53.0, 0.0),
    (2, 0, 0, 0, 0, 63.0, 0.0, 0.0, 0.0),
    (0, 0, 1, 0, 1, 63.0, 0.1, -33.0, 0.0),
    (2, 0, -1, 2, 2, -59.0, 0.0, 26.0, 0.0),
    (0, 0, -1, 0, 1, -58.0, -0.1, 32.0, 0.0),
    (0, 0, 1, 2, 1, -51.0, 0.0, 27.0, 0.0),
]


def nutation(t: float, nterms: int = len(NUTATION_TERMS)):
    """(d psi, d eps) in radians at TT centuries t."""
    D = math.radians(297.85036 + 445267.111480 * t - 0.0019142 * t * t + t ** 3 / 189474.0)
    M = math.radians(357.52772 + 35999.050340 * t - 0.0001603 * t * t - t ** 3 / 300000.0)
    Mp = math.radians(134.96298 + 477198.867398 * t + 0.0086972 * t * t + t ** 3 / 56250.0)
    F = math.radians(93.27191 + 483202.017538 * t - 0.0036825 * t * t + t ** 3 / 327270.0)
    Om = math.radians(125.04452 - 1934.136261 * t + 0.0020708 * t * t + t ** 3 / 450000.0)
    dpsi = deps = 0.0
    for (a, b, c, d, e, ps, pst, es, est) in NUTATION_TERMS[:nterms]:
        arg = a * D + b * M + c * Mp + d * F + e * Om
        dpsi += (ps + pst * t) * math.sin(arg)
        deps += (es + est * t) * math.cos(arg)
    return dpsi * 1e-4 * AS2R, deps * 1e-4 * AS2R


def nutation_matrix(t: float) -> np.ndarray:
    dpsi, deps = nutation(t)
    eps = mean_obliquity(t)
    return r1(-(eps + deps)) @ r3(-dpsi) @ r1(eps)


def bpn_matrix(t: float) -> np.ndarray:
    """J2000 (GCRS) -> true equator and equinox of date: N P B."""
    return nutation_matrix(t) @ precession_matrix(t) @ frame_bias_matrix()


def earth_velocity_ecliptic(t: float) -> np.ndarray:
    """The Earth's orbital velocity over c from the eccentric orbit, in the ecliptic of date:
    kappa (sin L - e sin pi, -cos L + e cos pi, 0), with L the Sun's true geometric
    longitude and pi the longitude of perihelion."""
    L0 = 280.46646 + 36000.76983 * t + 0.0003032 * t * t
    M = math.radians(357.52911 + 35999.05029 * t - 0.0001537 * t * t)
    C = ((1.914602 - 0.004817 * t - 0.000014 * t * t) * math.sin(M) + (0.019993 - 0.000101 * t) * math.sin(2 * M)
         + 0.000289 * math.sin(3 * M))
    L = math.radians(L0 + C)
    e = 0.016708634 - 0.000042037 * t - 0.0000001267 * t * t
    pi_ = math.radians(102.93735 + 1.71946 * t + 0.00046 * t * t)
    kappa = 20.49552 * AS2R
    return kappa * np.array([math.sin(L) - e * math.sin(pi_), -math.cos(L) + e * math.cos(pi_), 0.0])


def earth_velocity_j2000(t: float) -> np.ndarray:
    """earth_velocity_ecliptic in the mean equator of date, then rotated to J2000."""
    v_date = r1(-mean_obliquity(t)) @ earth_velocity_ecliptic(t)
    return (precession_matrix(t) @ frame_bias_matrix()).T @ v_date


def unit(ra: float, dec: float) -> np.ndarray:
    return np.array([math.cos(dec) * math.cos(ra), math.cos(dec) * math.sin(ra), math.sin(dec)])


def to_radec(p: np.ndarray):
    p = p / np.linalg.norm(p)
    return math.atan2(p[1], p[0]) % TAU, math.asin(max(-1.0, min(1.0, p[2])))


def aberrate_vector(p: np.ndarray, v: np.ndarray) -> np.ndarray:
    """Apparent direction of unit vector p for an observer moving at v (over c), to first
    order: p + v - (p . v) p."""
    return p + v - (p @ v) * p


def aberrate(ra: float, dec: float, t: float):
    """Apparent J2000 direction under annual aberration."""
    return to_radec(aberrate_vector(unit(ra, dec), earth_velocity_j2000(t)))


@dataclass
class PrecessionInfo:
    """marlu::precession::PrecessionInfo: `rotation_matrix` takes true-of-date vectors to
    J2000; `hadec_j2000` is the aberrated phase centre's (HA, Dec) at lmst_j2000."""
    rotation_matrix: np.ndarray
    hadec_j2000: tuple
    lmst: float
    lmst_j2000: float
    array_latitude_j2000: float

    def precess_xyz(self, xyzs) -> np.ndarray:
        """Tile XYZs in the local frame at lmst -> the local frame at lmst_j2000 after the
        rotation to J2000: rotate by lmst into the celestial frame of date, apply the
        matrix, rotate back by lmst_j2000."""
        xyzs = np.asarray(xyzs, np.float64).reshape(-1, 3)
        c, s = math.cos(self.lmst), math.sin(self.lmst)
        cel = np.stack([c * xyzs[:, 0] - s * xyzs[:, 1], s * xyzs[:, 0] + c * xyzs[:, 1], xyzs[:, 2]], axis=1)
        j = cel @ self.rotation_matrix.T
        c2, s2 = math.cos(self.lmst_j2000), math.sin(self.lmst_j2000)
        return np.stack([c2 * j[:, 0] + s2 * j[:, 1], -s2 * j[:, 0] + c2 * j[:, 1], j[:, 2]], axis=1)


def precess_time(array_longitude_rad: float, array_latitude_rad: float, phase_centre, gps_seconds: float,
                 dut1_seconds: float = 0.0) -> PrecessionInfo:
    """marlu::precession::precess_time (called at src/model/cpu.rs:733-739)."""
    t = tt_centuries(gps_seconds)
    lmst = get_lmst(array_longitude_rad, gps_seconds, dut1_seconds)
    rot = bpn_matrix(t).T
    lmst_j2000, lat_j2000 = to_radec(rot @ unit(lmst, array_latitude_rad))
    ra_ab, dec_ab = aberrate(phase_centre[0], phase_centre[1], t)
    return PrecessionInfo(rot, ((lmst_j2000 - ra_ab) % TAU, dec_ab), lmst, lmst_j2000, lat_j2000)


def xyzs_to_cross_uvws(xyzs, ha: float, dec: float) -> np.ndarray:
    """Baseline UVWs (tile i minus tile j, upper triangle) at (HA, Dec)."""
    x = np.asarray(xyzs, np.float64).reshape(-1, 3)
    s_ha, c_ha, s_d, c_d = math.sin(ha), math.cos(ha), math.sin(dec), math.cos(dec)
    u = s_ha * x[:, 0] + c_ha * x[:, 1]
    v = -s_d * c_ha * x[:, 0] + s_d * s_ha * x[:, 1] + c_d * x[:, 2]
    w = c_d * c_ha * x[:, 0] - c_d * s_ha * x[:, 1] + s_d * x[:, 2]
    tile = np.stack([u, v, w], axis=1)
    i, j = np.triu_indices(len(x), 1)
    return tile[i] - tile[j]


def lst_uvws_latitude(xyzs, phase_centre, array_longitude_rad, array_latitude_rad, gps_seconds, dut1_seconds,
                      precess: bool):
    """get_lst_uvws_latitude (src/model/cpu.rs:731-769): (LST, cross UVWs, latitude, tile
    XYZs the UVWs came from). With precession the LST and latitude are the zenith's J2000
    coordinates and the UVWs come from the precessed XYZs at the J2000 phase centre."""
    if precess:
        info = precess_time(array_longitude_rad, array_latitude_rad, phase_centre, gps_seconds, dut1_seconds)
        lst, lat, xyz = info.lmst_j2000, info.array_latitude_j2000, info.precess_xyz(xyzs)
    else:
        lst = get_lmst(array_longitude_rad, gps_seconds, dut1_seconds)
        lat, xyz = array_latitude_rad, np.asarray(xyzs, np.float64).reshape(-1, 3)
    return lst, xyzs_to_cross_uvws(xyz, lst - phase_centre[0], phase_centre[1]), lat, xyz
