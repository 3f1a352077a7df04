"""Pins of the epoch and precession restatement (oracle/precession_oracle.py), the checker for
the modeller's timestamp entry points.

- The reference's vis-simulate uvfits group parameters (src/cli/vis_simulate/tests.rs:225-316,
  fixture made by tests/golden/extract_precession_goldens.py): the first and last
  cross-correlation UVWs, and the first row's Julian date. The uvfits writer forms the UVWs
  from the precessed tile XYZs and the aberrated J2000 phase centre. Dropping the nutation or
  the aberration misses them by more than 4e-5 of the UVW; the full form is within 1e-7, the
  precision the file stores them at.
- Meeus, Astronomical Algorithms (2nd ed.), worked examples 12.a (GMST), 21.b (precession),
  22.a (nutation) and 23.a (nutation and aberration in RA/Dec).
- Invariants: orthonormal rotations, no precession at J2000.0, the Earth's rotation rate, GPS
  and UTC at a leap second."""

import json
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import precession_oracle as P

GOLD = json.loads((Path(__file__).parent / "golden" / "precession_goldens.json").read_text())
C_LIGHT = 299792458.0


def writer_uvw(info, xyzs, i, j):
    """UVW of tiles (i, j) as the uvfits writer forms it: precessed XYZs at hadec_j2000."""
    px = info.precess_xyz(np.asarray(xyzs)[[i, j]])
    ha, dec = info.hadec_j2000
    return P.xyzs_to_cross_uvws(px, ha, dec)[0]


def uvfits_cases():
    xyzs = np.asarray(GOLD["tile_xyzs"])
    lon, lat = math.radians(GOLD["array_longitude_deg"]), math.radians(GOLD["array_latitude_deg"])
    pc = tuple(math.radians(x) for x in GOLD["phase_centre_deg"])
    for key in ("first_cross", "last_cross"):
        case = GOLD[key]
        gps = GOLD["timestamps_gps"][case["timestep"]]
        yield xyzs, lon, lat, pc, gps, case


@pytest.mark.parametrize("variant", ["full", "no nutation", "no aberration", "no precession"])
def test_uvfits_group_param_uvws(variant):
    """The full restatement meets the stored u, v, w to 1e-7 of the UVW magnitude; each
    dropped term misses by far more."""
    worst = 0.0
    for xyzs, lon, lat, pc, gps, case in uvfits_cases():
        expected = np.asarray(case["uvw_seconds"]) * C_LIGHT
        info = P.precess_time(lon, lat, pc, gps, GOLD["dut1_seconds"])
        if variant == "no nutation":
            info.rotation_matrix = (P.precession_matrix(P.tt_centuries(gps)) @ P.frame_bias_matrix()).T
        elif variant == "no aberration":
            info.hadec_j2000 = ((info.lmst_j2000 - pc[0]) % P.TAU, pc[1])
        elif variant == "no precession":
            info.rotation_matrix = np.eye(3)
        i, j = case["tiles"]
        uvw = writer_uvw(info, xyzs, i, j)
        worst = max(worst, np.abs(uvw - expected).max() / np.abs(expected).max())
    print(f"{variant}: worst relative UVW error {worst:.3e}")
    if variant == "full":
        assert worst <= 1e-7
    else:
        assert worst >= 4e-5


def test_uvfits_first_row_date():
    """GPS -> UTC with the 2014 leap-second count: the row's Julian date to the precision of
    the single-precision group parameter (1.5e-8 d)."""
    jd = P.utc_days_since_j2000(GOLD["timestamps_gps"][0]) + P.J2000_JD
    assert abs(jd - GOLD["first_row_jd"]) <= 2e-8


def test_uvfits_baseline_numbers():
    """The fixture's tile pairs are the rows' baselines: 256 (i + 1) + (j + 1)."""
    for key in ("first_cross", "last_cross"):
        i, j = GOLD[key]["tiles"]
        assert GOLD[key]["baseline"] == 256 * (i + 1) + (j + 1)


def test_meeus_12a_gmst():
    """1987 April 10, 0h UT: GMST 13h10m46.3668s (IAU 1982). The IAU 2006 expression differs by
    its 0.0145" offset and the -0.3"/cy precession-rate change, 0.0033 s here."""
    tu = P.julian_day(1987, 4, 10) - P.J2000_JD
    g = P.gmst06(tu, (tu + 55.0 / 86400.0) / 36525.0)
    seconds = math.degrees(g) / 15.0 * 3600.0
    assert abs(seconds - (13 * 3600 + 10 * 60 + 46.3668)) <= 0.005


def test_meeus_21b_precession():
    """theta Persei to 2028 Nov 13.19 TD: the rotation composition with the IAU 1976 angles
    gives Meeus's 41.547214 deg, 49.348483 deg; the IAU 2006 angles differ from it by the model
    change only (< 0.15")."""
    t = (2462088.69 - P.J2000_JD) / 36525.0
    p0 = P.unit(math.radians(41.054063), math.radians(49.227750))
    AS = P.AS2R
    zeta = (2306.2181 * t + 0.30188 * t * t + 0.017998 * t ** 3) * AS
    z = (2306.2181 * t + 1.09468 * t * t + 0.018203 * t ** 3) * AS
    theta = (2004.3109 * t - 0.42665 * t * t - 0.041833 * t ** 3) * AS
    ra, dec = P.to_radec(P.r3(-z) @ P.r2(theta) @ P.r3(-zeta) @ p0)
    assert abs(math.degrees(ra) - 41.547214) <= 1e-6
    assert abs(math.degrees(dec) - 49.348483) <= 1e-6
    ra6, dec6 = P.to_radec(P.precession_matrix(t) @ p0)
    sep = math.acos(min(1.0, P.unit(ra, dec) @ P.unit(ra6, dec6)))
    assert 0.0 < sep / AS <= 0.15


def test_meeus_22a_nutation():
    """1987 April 10, 0h TD: d psi = -3.788", d eps = +9.443" (the full 1980 series); the 18
    largest terms are within 0.005". The IAU 2006 mean obliquity is 0.042" below IAU 1980's
    23 deg 26' 27.407"."""
    t = (P.julian_day(1987, 4, 10) - P.J2000_JD) / 36525.0
    dpsi, deps = P.nutation(t)
    assert abs(dpsi / P.AS2R - (-3.788)) <= 0.005
    assert abs(deps / P.AS2R - 9.443) <= 0.005
    eps80 = (23 * 3600 + 26 * 60 + 27.407)
    assert abs(P.mean_obliquity(t) / P.AS2R - (eps80 - 0.042)) <= 0.005


def test_meeus_23a_nutation_and_aberration():
    """theta Persei, 2028 Nov 13.19 TD, mean place of date 41.5472 deg, 49.3485 deg: nutation
    moves it by +15.843" in RA and +6.217" in Dec (0.03" allowed for the truncated series),
    aberration by +30.045" and +6.697"."""
    t = (2462088.69 - P.J2000_JD) / 36525.0
    ra0, dec0 = math.radians(41.5472), math.radians(49.3485)
    p = P.unit(ra0, dec0)
    ran, decn = P.to_radec(P.nutation_matrix(t) @ p)
    assert abs(math.degrees(ran - ra0) * 3600 - 15.843) <= 0.03
    assert abs(math.degrees(decn - dec0) * 3600 - 6.217) <= 0.03
    v = P.r1(-P.mean_obliquity(t)) @ P.earth_velocity_ecliptic(t)
    raa, deca = P.to_radec(P.aberrate_vector(p, v))
    assert abs(math.degrees(raa - ra0) * 3600 - 30.045) <= 0.005
    assert abs(math.degrees(deca - dec0) * 3600 - 6.697) <= 0.005


def test_rotations_orthonormal_and_identity_at_j2000():
    for t in (-0.3, 0.0, 0.1446, 0.5):
        for m in (P.precession_matrix(t), P.nutation_matrix(t), P.bpn_matrix(t)):
            np.testing.assert_allclose(m @ m.T, np.eye(3), atol=1e-15)
            assert abs(np.linalg.det(m) - 1.0) <= 1e-15
    np.testing.assert_allclose(P.precession_matrix(0.0), np.eye(3), atol=1e-16)


def test_earth_rotation_rate_and_lmst_range():
    """LMST advances 1.00273781191135448 turns per UT1 day plus the GMST polynomial's
    4612.157" per Julian century, and stays in [0, 2 pi)."""
    lon = math.radians(116.67081523611111)
    gps = 1090008644.0
    a, b = P.get_lmst(lon, gps), P.get_lmst(lon, gps + 60.0)
    rate = ((b - a) % P.TAU) / 60.0
    expected = P.TAU * 1.00273781191135448 / 86400.0 + 4612.156534 * P.AS2R / (36525.0 * 86400.0)
    assert abs(rate - expected) <= 5e-13
    for g in np.linspace(1e9, 1.4e9, 17):
        assert 0.0 <= P.get_lmst(lon, g, 0.3) < P.TAU
    # DUT1 shifts UT1 only: 1 s of DUT1 is 1.0027 s of rotation.
    d = (P.get_lmst(lon, gps, 1.0) - P.get_lmst(lon, gps, 0.0)) % P.TAU
    assert abs(d - P.TAU * 1.00273781191135448 / 86400.0) <= 1e-12


def test_gps_utc_leap_seconds():
    """GPS 1167264018 is 2017-01-01T00:00:00 UTC (TAI - UTC 37 s); one GPS second before the
    leap second it is 23:59:59 on 2016-12-31 (36 s)."""
    assert P.utc_days_since_j2000(1167264018.0) == pytest.approx(6209.5, abs=1e-12)
    assert P.tai_minus_utc(1167264018.0) == 37
    assert P.tai_minus_utc(1167264016.0) == 36
    assert (P.utc_days_since_j2000(1167264016.0) - 6209.5) * 86400.0 == pytest.approx(-1.0, abs=1e-6)
    assert P.tai_minus_utc(0.0) == 19


def test_precess_xyz_keeps_lengths_and_zenith():
    """precess_xyz is a rotation (baseline lengths kept), and lmst_j2000 / latitude_j2000 are
    the J2000 coordinates of the zenith of date."""
    xyzs = np.asarray(GOLD["tile_xyzs"])
    lon, lat = math.radians(GOLD["array_longitude_deg"]), math.radians(GOLD["array_latitude_deg"])
    info = P.precess_time(lon, lat, (0.0, math.radians(-27.0)), GOLD["timestamps_gps"][0])
    px = info.precess_xyz(xyzs)
    np.testing.assert_allclose(np.linalg.norm(px, axis=1), np.linalg.norm(xyzs, axis=1), rtol=1e-14)
    zen = P.bpn_matrix(P.tt_centuries(GOLD["timestamps_gps"][0])) @ P.unit(info.lmst_j2000, info.array_latitude_j2000)
    np.testing.assert_allclose(zen, P.unit(info.lmst, lat), atol=1e-15)
    # Precession over 14.5 years moves the zenith by about 0.2 deg.
    assert 0.1 < math.degrees(abs(info.lmst_j2000 - info.lmst)) < 0.4
