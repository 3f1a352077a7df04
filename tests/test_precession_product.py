"""The library's timestamp functions (sm_get_lmst, sm_precess_time, sm_precess_xyz: host code,
no GPU) against the oracle restatement and the reference's vis-simulate uvfits numbers
(src/cli/vis_simulate/tests.rs:225-316)."""

import ctypes as C
import json
import math
from pathlib import Path

import numpy as np
import pytest

from mwa_hyperdrive_amd import _lib
from mwa_hyperdrive_amd.precession import Epoch, get_lmst, precess_time
from oracle import precession_oracle as P

GOLD = json.loads((Path(__file__).parent / "golden" / "precession_goldens.json").read_text())
LON = math.radians(GOLD["array_longitude_deg"])
LAT = math.radians(GOLD["array_latitude_deg"])
PC = tuple(math.radians(x) for x in GOLD["phase_centre_deg"])
# GPS times from 1981 to 2031, with a DUT1 each; the first two straddle the 2017 leap second.
EPOCHS = [1167264016.0, 1167264018.0, 1090008644.0, 1.0e9, 3.0e8, 1.4e9, 1.6e9]
DUT1S = [-0.4, 0.0, 0.25, 0.0, 0.7, -0.1, 0.3]


def test_struct_size():
    assert C.sizeof(_lib.SmPrecessionInfo) == 14 * 8


@pytest.mark.parametrize("gps,dut1", list(zip(EPOCHS, DUT1S)))
def test_lmst_and_precession_match_oracle(gps, dut1):
    assert abs(get_lmst(LON, Epoch(gps), dut1) - P.get_lmst(LON, gps, dut1)) <= 2e-12
    for pc in (PC, (1.2, 0.4), (5.9, -1.3)):
        got = precess_time(LON, LAT, pc, Epoch(gps), dut1)
        ref = P.precess_time(LON, LAT, pc, gps, dut1)
        np.testing.assert_allclose(got.rotation_matrix, ref.rotation_matrix, rtol=0, atol=1e-15)
        assert abs(got.lmst - ref.lmst) <= 2e-12
        for a, b in ((got.lmst_j2000, ref.lmst_j2000), (got.hadec_j2000[0], ref.hadec_j2000[0])):
            assert abs((a - b + math.pi) % (2 * math.pi) - math.pi) <= 2e-12
        assert abs(got.array_latitude_j2000 - ref.array_latitude_j2000) <= 1e-14
        assert abs(got.hadec_j2000[1] - ref.hadec_j2000[1]) <= 1e-14
        xyzs = np.asarray(GOLD["tile_xyzs"])
        np.testing.assert_allclose(got.precess_xyz(xyzs), ref.precess_xyz(xyzs), rtol=0, atol=1e-9)


def test_uvfits_group_param_uvws_through_the_library():
    """The library's precess_time and precess_xyz give the stored u, v, w of the first and last
    cross-correlation rows to 1e-7 of the UVW magnitude."""
    xyzs = np.asarray(GOLD["tile_xyzs"])
    for key in ("first_cross", "last_cross"):
        case = GOLD[key]
        info = precess_time(LON, LAT, PC, Epoch(GOLD["timestamps_gps"][case["timestep"]]), GOLD["dut1_seconds"])
        i, j = case["tiles"]
        px = info.precess_xyz(xyzs[[i, j]])
        uvw = P.xyzs_to_cross_uvws(px, *info.hadec_j2000)[0]
        expected = np.asarray(case["uvw_seconds"]) * 299792458.0
        assert np.abs(uvw - expected).max() <= 1e-7 * np.abs(expected).max()


def test_precess_time_rejects_bad_arguments():
    L = _lib.lib()
    info = _lib.SmPrecessionInfo()
    assert L.sm_precess_time(LON, LAT, 0.0, 0.0, float("nan"), 0.0, C.byref(info)) is not None
    assert L.sm_precess_time(LON, LAT, 0.0, 0.0, 1e9, 0.0, None) is not None
