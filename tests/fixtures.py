"""Shared fixture builders, restating the reference's test fixtures.

`ObsParams` follows src/model/tests/mod.rs:241-348 (3 tiles at (0,1,0), (0,-1,0),
(1,0,0); 150/175/200 MHz; phase centre (0, -27) deg; LST 0; MWA latitude). The
component builders follow mod.rs:125-239.
"""

import json
import math
from pathlib import Path

import numpy as np

from mwa_hyperdrive_amd.srclist import (
    CurvedPowerLaw,
    FluxDensity,
    FluxList,
    Gaussian,
    Point,
    PowerLaw,
    Shapelet,
    Source,
    SourceComponent,
    SourceList,
    radec_from_degrees,
)

GOLDEN = json.loads((Path(__file__).parent / "golden" / "model_goldens.json").read_text())
MWA_LAT_RAD = -0.4660608448386394
MWA_LONG_RAD = 2.0362898668561042
PHASE_CENTRE = radec_from_degrees(0.0, -27.0)
OFF_PHASE_CENTRE = radec_from_degrees(1.0, -27.0)
FREQS = [150e6, 175e6, 200e6]
XYZS = [[0.0, 1.0, 0.0], [0.0, -1.0, 0.0], [1.0, 0.0, 0.0]]


def get_list():
    return FluxList([FluxDensity(150e6, 1.0), FluxDensity(175e6, 3.0), FluxDensity(200e6, 2.0)])


def get_power_law():
    return PowerLaw(-0.8, FluxDensity(150e6, 1.0))


def get_curved_power_law():
    return CurvedPowerLaw(-0.8, FluxDensity(150e6, 1.0), 0.03)


FLUX = {"list": get_list, "power_law": get_power_law, "curved_power_law": get_curved_power_law}


def get_point(pos, flux):
    return SourceComponent(pos[0], pos[1], Point(), FLUX[flux]())


def get_gaussian(pos, flux):
    return SourceComponent(pos[0], pos[1], Gaussian(0.0, 0.0, 0.0), FLUX[flux]())


def get_gaussian2(pos, flux):
    return SourceComponent(pos[0], pos[1], Gaussian(1.0, 0.5, 0.25), FLUX[flux]())


def get_shapelet(pos, flux):
    return SourceComponent(pos[0], pos[1], Shapelet(0.0, 0.0, 0.0, [(0, 0, 1.0)]), FLUX[flux]())


BUILDERS = {"point": get_point, "gaussian": get_gaussian, "shapelet": get_shapelet, "gaussian2": get_gaussian2}
TYPE_MASK = {"point": 1, "gaussian": 2, "gaussian2": 2, "shapelet": 4}


def single(builder, pos, flux):
    return SourceList([("s", Source([BUILDERS[builder](pos, flux)]))])


def golden(name):
    return np.asarray(GOLDEN["goldens"][name], dtype=np.float32)


# (golden checker, component builder, position, flux type) for every no-beam golden run by
# the reference's CPU and GPU tests (src/model/tests/cpu.rs:57-112,221-246; gpu.rs).
NO_BEAM_CASES = []
for _builder in ("point", "gaussian", "shapelet"):
    for _flux in ("list", "power_law", "curved_power_law"):
        NO_BEAM_CASES.append((f"test_{_flux}_zenith_visibilities", _builder, PHASE_CENTRE, _flux))
        NO_BEAM_CASES.append((f"test_{_flux}_off_zenith_visibilities", _builder, OFF_PHASE_CENTRE, _flux))


def multiple_components(builder):
    """gaussian_multiple_components / shapelet_multiple_components (cpu.rs:306-330,400-438)."""
    return SourceList([
        ("multi", Source([
            BUILDERS[builder](radec_from_degrees(1.0, -27.0), "list"),
            BUILDERS[builder](radec_from_degrees(1.1, -27.0), "list"),
        ]))
    ])


def vis_max_rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = np.abs(b.reshape(-1, 2) @ np.array([1.0, 1j])).max()
    diff = np.abs((a - b).reshape(-1, 2) @ np.array([1.0, 1j])).max()
    return diff / scale if scale > 0 else diff
