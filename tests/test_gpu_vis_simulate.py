"""The product on the reference's vis-simulate check (src/cli/vis_simulate/tests.rs:103-316):
a modeller over the 128 tiles of observation 1090008640 with the 100-source list, no beam,
DUT1 0 and precession on, called with each timestamp. The first and last cross rows' XX of
the first channel must match the values the reference test states (to their 7 significant
digits; its own epsilon is 1e-3), and every visibility the oracle at the same timestamp within
the FP64 bar."""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from mwa_hyperdrive_amd.modeller import NoBeam, SkyModeller
from mwa_hyperdrive_amd.precession import Epoch
from oracle import oracle as orc
from oracle import precession_oracle as P
from tests.helpers import rel_err, vis_simulate_case

pytestmark = pytest.mark.gpu


def test_vis_simulate_rows_through_the_product():
    case = vis_simulate_case()
    m = SkyModeller(NoBeam(), case["srclist"], 0, case["xyzs"], case["freqs"], [], case["phase_centre"],
                    case["longitude"], case["latitude"], dut1=0.0, apply_precession=True)
    nbl = m.num_baselines
    for row, bl, key in ((0, 0, "first"), (1, nbl - 1, "last")):
        v32, uvws, v64 = m.model_timestep(Epoch(case["gps"][row]), want_f64=True)
        got = v32[0, bl, :2].astype(np.float64)
        print(f"{key} row: {got} against {case[key]}")
        assert np.abs(got - np.asarray(case[key])).max() <= 3e-5
        lst, ouvws, lat, uxyz = P.lst_uvws_latitude(case["xyzs"], case["phase_centre"], case["longitude"],
                                                    case["latitude"], case["gps"][row], 0.0, True)
        np.testing.assert_allclose(uvws, ouvws, rtol=0, atol=1e-9)
        o = orc.Obs(case["xyzs"], case["freqs"], case["phase_centre"], lst, lat, uvw_xyzs=uxyz)
        _, o64 = orc.model_timestep(o, case["srclist"])
        assert rel_err(v64, o64) <= 1e-12


def test_vis_simulate_uvfits_end_to_end(tmp_path):
    """Model every timestep through the product and write the uvfits file as vis-simulate does
    (weights 64 = 80 kHz / 10 kHz x 8 s); the first three floats of the first and last cross
    rows, and their group parameters, are what the reference test reads back."""
    import json
    import math
    from pathlib import Path

    from mwa_hyperdrive_amd.uvfits import UvfitsWriter, cross_baseline_tiles
    from tests.uvfits_reader import read_uvfits

    gold = json.loads((Path(__file__).parent / "golden" / "precession_goldens.json").read_text())
    case = vis_simulate_case()
    m = SkyModeller(NoBeam(), case["srclist"], 0, case["xyzs"], case["freqs"], [], case["phase_centre"],
                    case["longitude"], case["latitude"], dut1=0.0, apply_precession=True)
    pairs = cross_baseline_tiles(len(case["xyzs"]))
    path = tmp_path / "model.uvfits"
    weights = np.full((len(case["freqs"]), len(pairs)), 64.0, np.float32)
    with UvfitsWriter(path, len(case["gps"]), pairs, case["freqs"], case["phase_centre"],
                      (case["longitude"], case["latitude"], 377.827), gold["tile_names"], case["xyzs"], 8.0,
                      Epoch(case["gps"][0]), dut1=0.0, apply_precession=True, obs_name="1090008640") as w:
        for gps in case["gps"]:
            v32, _ = m.model_timestep(Epoch(gps))
            w.write_timestep(Epoch(gps), v32, weights)
    h0, params, data, _ = read_uvfits(path)
    flat = data.reshape(data.shape[0], -1)
    for row, key, gkey in ((0, "first", "first_cross"), (-1, "last", "last_cross")):
        got = flat[row, :3].astype(np.float64)
        want = np.array(case[key] + [64.0])
        print(f"{key} row in the file: {got} against {want}")
        assert np.abs(got - want).max() <= 3e-5
        uvw = np.asarray(gold[gkey]["uvw_seconds"])
        assert np.abs(params[row, :3] - uvw).max() <= 1e-7 * np.abs(uvw).max()
        assert params[row, 3] == gold[gkey]["baseline"]
    assert float(params[0, 4]) + float(h0["PZERO5"]) == gold["first_row_jd"]
