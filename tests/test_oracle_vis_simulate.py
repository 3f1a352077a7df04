"""End-to-end pin of the timestamp path on the reference's own visibilities: vis-simulate of
observation 1090008640 with the 100-source list, no beam, two 80 kHz channels about the
metafits centre frequency, two 8 s timesteps, precession on (src/cli/vis_simulate/tests.rs:
103-316). The oracle models at the J2000 LMST, latitude and precessed tile XYZs of each
timestamp (oracle/precession_oracle.py) and must give the first and last cross rows' XX of
the first channel to the 7 significant digits the test states them with (the test's own
epsilon is 1e-3). Without precession, or at the coarse-channel centre (182.400 MHz) instead of
FREQCENT, the same rows miss by more than 0.05."""

import numpy as np
import pytest

from oracle import oracle as orc
from oracle import precession_oracle as P
from tests.helpers import vis_simulate_case


def oracle_row_xx(case, row, precess, freqs=None):
    """XX of channel 0 of the first (row 0) or last (row 1) cross row, from the two tiles."""
    i, j = (0, 1) if row == 0 else (len(case["xyzs"]) - 2, len(case["xyzs"]) - 1)
    lst, _, lat, uxyz = P.lst_uvws_latitude(case["xyzs"], case["phase_centre"], case["longitude"], case["latitude"],
                                            case["gps"][row], 0.0, precess)
    o = orc.Obs(case["xyzs"][[i, j]], case["freqs"] if freqs is None else freqs, case["phase_centre"], lst, lat,
                uvw_xyzs=uxyz[[i, j]] if precess else None)
    v32, _ = orc.model_timestep(o, case["srclist"])
    return np.array([v32[0, 0, 0], v32[0, 0, 1]], np.float64)


@pytest.mark.parametrize("row", [0, 1], ids=["first", "last"])
def test_vis_simulate_rows(row):
    case = vis_simulate_case()
    want = np.asarray(case["first" if row == 0 else "last"])
    got = oracle_row_xx(case, row, True)
    print(f"row {row}: {got} against {want}")
    assert np.abs(got - want).max() <= 3e-5
    assert np.abs(oracle_row_xx(case, row, False) - want).max() > 0.05
    coarse_centre = np.array([182.4e6 - 80e3, 182.4e6])
    assert np.abs(oracle_row_xx(case, row, True, coarse_centre) - want).max() > 0.05
