"""The uvfits writer (sm_uvfits_*: host code) against what the reference's vis-simulate check
reads back from its file (src/cli/vis_simulate/tests.rs:103-316): PCOUNT 7, three floats per
polarisation, four polarisations, the channel count, PZERO5 = 2456860.5, PTYPE4 'BASELINE',
GCOUNT rows, the AN table's tile names (Tile011 ... Tile168) and un-precessed geodetic
positions (tile 0 = 456.2500494643639, -149.78500366210938, 68.04598669887378), and the first
and last cross rows' group parameters (u, v, w in seconds; baseline 258 and 32640; the date as
jd_zero-relative part and remainder; integration 8 s). The visibilities here are synthetic, so
the file's data must equal them in the file's polarisation order XX, YY, XY, YX."""

import json
import math
from pathlib import Path

import numpy as np
import pytest

from mwa_hyperdrive_amd.precession import Epoch
from mwa_hyperdrive_amd.uvfits import UvfitsWriter, cross_baseline_tiles
from tests.uvfits_reader import decode_baseline, read_uvfits

GOLD = json.loads((Path(__file__).parent / "golden" / "precession_goldens.json").read_text())
C_LIGHT = 299792458.0
MWA_HEIGHT = 377.827


def write_case(path, nchan=2, vis_seed=1, precess=True):
    xyz = np.asarray(GOLD["tile_xyzs"])
    names = GOLD["tile_names"]
    pairs = cross_baseline_tiles(len(xyz))
    freqs = [182.415e6 - 80e3, 182.415e6][:nchan] if nchan <= 2 else 182.335e6 + 80e3 * np.arange(nchan)
    pc = tuple(math.radians(x) for x in GOLD["phase_centre_deg"])
    arr = (math.radians(GOLD["array_longitude_deg"]), math.radians(GOLD["array_latitude_deg"]), MWA_HEIGHT)
    rng = np.random.default_rng(vis_seed)
    vis = [rng.normal(size=(len(freqs), len(pairs), 8)).astype(np.float32) for _ in GOLD["timestamps_gps"]]
    with UvfitsWriter(path, len(vis), pairs, freqs, pc, arr, names, xyz, 8.0, Epoch(GOLD["timestamps_gps"][0]),
                      dut1=0.0, apply_precession=precess, obs_name="1090008640") as w:
        for t, gps in enumerate(GOLD["timestamps_gps"]):
            w.write_timestep(Epoch(gps), vis[t], np.full((len(freqs), len(pairs)), 64.0, np.float32))
    return vis, pairs


def test_header_and_antenna_table(tmp_path):
    path = tmp_path / "model.uvfits"
    vis, pairs = write_case(path)
    h0, params, data, tables = read_uvfits(path)
    assert int(h0["PCOUNT"]) == 7
    assert int(h0["NAXIS2"]) == 3 and int(h0["NAXIS3"]) == 4 and int(h0["NAXIS4"]) == 2
    assert float(h0["PZERO5"]) == 2456860.5
    assert h0["PTYPE4"] == "BASELINE"
    assert int(h0["GCOUNT"]) == 2 * len(pairs)
    ah, an = tables["AIPS AN"]
    assert an["ANNAME"][0] == "Tile011" and an["ANNAME"][1] == "Tile012" and an["ANNAME"][127] == "Tile168"
    assert an["ANNAME"] == GOLD["tile_names"]
    np.testing.assert_array_equal(np.asarray(an["STABXYZ"]), np.asarray(GOLD["tile_xyzs"]))
    assert tuple(an["STABXYZ"][0]) == (456.2500494643639, -149.78500366210938, 68.04598669887378)


def test_group_parameters_of_first_and_last_rows(tmp_path):
    path = tmp_path / "model.uvfits"
    vis, pairs = write_case(path)
    h0, params, data, _ = read_uvfits(path)
    jd_zero = float(h0["PZERO5"])
    for key, row in (("first_cross", 0), ("last_cross", len(params) - 1)):
        case = GOLD[key]
        p = params[row].astype(np.float64)
        want = np.asarray(case["uvw_seconds"])
        assert np.abs(p[:3] - want).max() <= 1e-7 * np.abs(want).max()
        assert p[3] == case["baseline"]
        assert decode_baseline(p[3]) == (case["tiles"][0] + 1, case["tiles"][1] + 1)
        assert p[6] == 8.0
    # The first row's date: the f32 part plus PZERO5 gives the Julian date the test states
    # (itself the f32 rounding); the second DATE parameter restores the rest, so both parts give
    # 20:10:28.000 UTC to well under a millisecond. The test's own check of that second part is
    # f32 epsilon absolute (it states 5.276111e-9).
    assert float(params[0, 4]) + jd_zero == GOLD["first_row_jd"]
    jd = float(params[0, 4]) + float(params[0, 5]) + jd_zero
    exact = 2456860.5 - (3 * 3600 + 49 * 60 + 32) / 86400.0
    assert abs(jd - exact) * 86400.0 <= 1e-5
    assert abs(float(params[0, 5]) - 5.276111e-9) <= 1.1920929e-07
    # The last timestep is 8 s later.
    jd_last = float(params[-1, 4]) + float(params[-1, 5]) + jd_zero
    assert abs((jd_last - jd) * 86400.0 - 8.0) <= 1e-4


def test_visibilities_and_weights_in_file_order(tmp_path):
    path = tmp_path / "model.uvfits"
    vis, pairs = write_case(path)
    _, _, data, _ = read_uvfits(path)
    # data [group][dec][ra][if][chan][pol][3]
    d = data.reshape(data.shape[0], -1, 4, 3)
    order = [0, 3, 1, 2]  # XX, YY, XY, YX from Jones XX, XY, YX, YY
    for t in range(len(vis)):
        rows = d[t * len(pairs):(t + 1) * len(pairs)]  # [baseline][chan][pol][3]
        for pol, j in enumerate(order):
            np.testing.assert_array_equal(rows[:, :, pol, 0], vis[t][:, :, 2 * j].T)
            np.testing.assert_array_equal(rows[:, :, pol, 1], vis[t][:, :, 2 * j + 1].T)
        assert (rows[:, :, :, 2] == 64.0).all()


def test_without_precession_the_uvws_are_at_the_lmst(tmp_path):
    """apply_precession off: the rows' UVWs are the unprecessed XYZs at the phase centre's hour
    angle, which misses the stored (precessed) values by far more than their precision."""
    path = tmp_path / "model.uvfits"
    write_case(path, precess=False)
    _, params, _, _ = read_uvfits(path)
    want = np.asarray(GOLD["first_cross"]["uvw_seconds"])
    assert np.abs(params[0, :3].astype(np.float64) - want).max() > 1e-4 * np.abs(want).max()


def test_bad_arguments_are_errors(tmp_path):
    from mwa_hyperdrive_amd import _lib
    with pytest.raises(_lib.SkyModelError):
        UvfitsWriter(tmp_path / "x.uvfits", 1, [[1, 0]], [1e8, 1.1e8], (0.0, 0.0), (0.0, 0.0, 0.0), ["a", "b"],
                     np.zeros((2, 3)), 1.0, 1e9)
    w = UvfitsWriter(tmp_path / "y.uvfits", 1, [[0, 1]], [1e8], (0.0, 0.0), (0.0, 0.0, 0.0), ["a", "b"],
                     np.zeros((2, 3)), 1.0, 1e9)
    with pytest.raises(_lib.SkyModelError, match="fewer timesteps"):
        w.finish()
