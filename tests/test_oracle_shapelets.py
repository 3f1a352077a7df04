"""Pins the oracle's shapelet envelope (src/model/cpu.rs:633-695) where the reference has no
golden vector: nonzero axes and basis orders above 0.

* n1 = n2 = 0 with unit coefficient is the Gaussian envelope of the same (maj, min, pa):
  B_0(x) = exp(-x^2 / 2) at x = x_uv maj sqrt(pi^2 / (2 ln 2)) equals
  exp(-(pi / 2)^2 / ln 2 maj^2 x_uv^2), up to the table's linear interpolation (<= dx^2 / 8).
  A wrong position scale, rotation or axis assignment fails this.
* (n1, n2) = (1, 0) multiplies the n = 0 envelope by i sqrt(2) x (B_1 = sqrt(2) x B_0 and the
  factor i^(n1 + n2)). A wrong power of i or basis order fails this.
* Swapping n1 <-> n2 together with maj <-> min and pa -> pa + pi/2 maps (x, y) to (y, -x), so
  the envelope changes by (-1)^n2. Pairing n1 with the wrong axis, or dropping the minus sign
  on the minor axis, fails this.
The components sit at the phase centre, where the shapelet (u, v) equal the baseline (u, v).
"""

import math

import numpy as np

from mwa_hyperdrive_amd.srclist import Gaussian, PowerLaw, FluxDensity, Shapelet, Source, SourceComponent, SourceList
from oracle import oracle as orc
from tests.fixtures import MWA_LAT_RAD, PHASE_CENTRE

XYZS = [[0.0, 0.0, 0.0], [30.0, 40.0, 0.0], [-50.0, 20.0, 10.0], [10.0, -70.0, 5.0], [-25.0, -35.0, -3.0]]
FREQS = [150e6, 175e6, 200e6]
LST = 0.3
MAJ, MIN, PA = math.radians(2.0), math.radians(0.8), 0.7


def _obs():
    return orc.Obs(XYZS, FREQS, PHASE_CENTRE, LST, MWA_LAT_RAD)


def _sky(shape):
    c = SourceComponent(PHASE_CENTRE[0], PHASE_CENTRE[1], shape, PowerLaw(-0.8, FluxDensity(150e6, 2.0)))
    return SourceList([("s", Source([c]))])


def _vis(shape, mask):
    return orc.model_timestep(_obs(), _sky(shape), mask=mask)[1]  # f64 [freq][baseline][8]


def _xy_uv():
    """Rotated baseline (x, y) in wavelengths, [freq][baseline]."""
    uvw = np.asarray(orc.cross_uvws(XYZS, LST - PHASE_CENTRE[0], PHASE_CENTRE[1])).reshape(-1, 3)
    f = np.asarray(FREQS)[:, None] / 299792458.0
    u, v = uvw[None, :, 0] * f, uvw[None, :, 1] * f
    return u * math.sin(PA) + v * math.cos(PA), u * math.cos(PA) - v * math.sin(PA)


def _cplx(v, pol=0):
    return v[..., 2 * pol] + 1j * v[..., 2 * pol + 1]


def test_n0_shapelet_is_the_gaussian_envelope():
    g = _vis(Gaussian(MAJ, MIN, PA), 2)
    s = _vis(Shapelet(MAJ, MIN, PA, [(0, 0, 1.0)]), 4)
    scale = np.abs(_cplx(g)).max()
    assert np.abs(_cplx(s) - _cplx(g)).max() <= 2e-5 * scale
    # The envelope is far from 1 on the longer baselines, so this is not a trivial match.
    x, _ = _xy_uv()
    assert (np.abs(x) * MAJ * math.sqrt(math.pi ** 2 / (2 * math.log(2)))).max() > 1.0


def test_first_order_basis_and_power_of_i():
    v0 = _cplx(_vis(Shapelet(MAJ, MIN, PA, [(0, 0, 1.0)]), 4))
    v1 = _cplx(_vis(Shapelet(MAJ, MIN, PA, [(1, 0, 1.0)]), 4))
    x, _ = _xy_uv()
    xt = x * MAJ * math.sqrt(math.pi ** 2 / (2 * math.log(2)))
    want = 1j * math.sqrt(2.0) * xt * v0
    assert np.abs(v1 - want).max() <= 1e-4 * np.abs(v0).max()


def test_order_swap_symmetry():
    for n1, n2 in [(3, 1), (2, 5), (4, 0)]:
        a = _vis(Shapelet(MAJ, MIN, PA, [(n1, n2, 0.7)]), 4)
        b = _vis(Shapelet(MIN, MAJ, PA + math.pi / 2, [(n2, n1, 0.7)]), 4)
        scale = np.abs(a).max()
        assert scale > 0
        np.testing.assert_allclose(b, (-1) ** n2 * a, rtol=0, atol=1e-9 * scale)
