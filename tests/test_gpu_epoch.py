"""The modeller's timestamp entry points against the oracle. The modeller is made with a DUT1
and with or without precession (new_sky_modeller's dut1 / apply_precession,
src/model/mod.rs:215-227). The entry points are sm_lst_uvws_latitude,
sm_model_timestep_epoch_with(_device) and sm_model_timestep_autos_epoch_with.

The oracle models at the (LST, latitude) and with the cross-UVW tile XYZs that
get_lst_uvws_latitude gives (src/model/cpu.rs:731-769, restated in
oracle/precession_oracle.py). With precession, the shapelet UVs keep the unprecessed XYZs
(src/model/cpu.rs:782-785). Sky: points, Gaussians and shapelets, with per-tile beams
(dead dipoles). Bar: the FP64 bar, 1e-12 of max |V|."""

import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from mwa_hyperdrive_amd import synthetic
from mwa_hyperdrive_amd.modeller import AnalyticBeam, SkyModeller
from mwa_hyperdrive_amd.precession import Epoch
from oracle import oracle as orc
from oracle import precession_oracle as P
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

GPS = 1090008644.0
LON = math.radians(116.67081523611111)
DUT1 = 0.3


def _setup():
    base = synthetic.shrink(synthetic.config(5, scale=0.004), tiles_n=40, chans=7, timesteps=1)
    return base


def _modeller(base, precess):
    return SkyModeller(AnalyticBeam(base.delays, base.gains), base.srclist, 0, base.xyzs, base.freqs, [],
                       base.phase_centre, LON, base.latitude, dut1=DUT1, apply_precession=precess)


@pytest.mark.parametrize("precess", [False, True], ids=["no-precession", "precession"])
def test_epoch_entry_points_vs_oracle(precess):
    base = _setup()
    m = _modeller(base, precess)
    info = m.info()
    assert info["num_points"] > 0 and info["num_gaussians"] > 0 and info["num_shapelets"] > 0
    assert info["num_unique_tiles"] > 1
    lst, uvws, lat, uvw_xyz = P.lst_uvws_latitude(base.xyzs, base.phase_centre, LON, base.latitude, GPS, DUT1,
                                                  precess)
    g_lst, g_uvws, g_lat = m.lst_uvws_latitude(Epoch(GPS))
    assert abs(g_lst - lst) <= 2e-12 and abs(g_lat - lat) <= 1e-14
    np.testing.assert_allclose(g_uvws, uvws, rtol=0, atol=1e-9)
    o = orc.Obs(base.xyzs, base.freqs, base.phase_centre, lst, lat, beam="analytic", delays=base.delays,
                gains=base.gains, uvw_xyzs=uvw_xyz if precess else None)
    o32, o64 = orc.model_timestep(o, base.srclist)
    v32, v_uvws, v64 = m.model_timestep(Epoch(GPS), want_f64=True)
    err = rel_err(v64, o64)
    print(f"precess={precess}: LST {lst:.9f} rad, latitude {lat:.9f} rad, rel err {err:.3e}")
    assert err <= 1e-12
    np.testing.assert_allclose(v_uvws, uvws, rtol=0, atol=1e-9)
    # The device form runs the same kernels on the same inputs. Split component loops add their
    # FP64 partial sums atomically, in an order that varies from call to call, so the two
    # agree to rounding rather than bit for bit.
    shape = (len(base.freqs), base.num_baselines, 8)
    d32 = torch.zeros(shape, dtype=torch.float32, device="cuda")
    d64 = torch.zeros(shape, dtype=torch.float64, device="cuda")
    m.model_timestep_with_device(Epoch(GPS), d32, d64)
    torch.cuda.synchronize()
    assert rel_err(d64.cpu().numpy(), v64) <= 1e-14
    assert rel_err(d32.cpu().numpy(), v32) <= 1e-6
    # Autos at the same LST and latitude.
    a32 = m.model_timestep_autos_with(Epoch(GPS))
    r32, r64 = orc.model_autos(o, base.srclist)
    assert rel_err(a32, r64) <= 1e-6


def test_precession_changes_the_model():
    """The two settings model different skies: 14.5 years of precession move the J2000 zenith
    by about 0.19 deg, far above the FP64 bar."""
    base = _setup()
    v_no = _modeller(base, False).model_timestep(Epoch(GPS), want_f64=True)[2]
    v_pr = _modeller(base, True).model_timestep(Epoch(GPS), want_f64=True)[2]
    assert rel_err(v_pr, v_no) > 1e-3
