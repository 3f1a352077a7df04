"""Edge cases of the SkyModeller handle against the oracle:
- flagged tiles (src/model/mod.rs:215-227: unflagged XYZs, the flagged set, beams indexed by
  the total array);
- unevenly spaced channels and sky sizes that are not a multiple of the 16-component step
  through the split-precision path;
- a modeller re-used across skies of different sizes (update_with_a_source)."""

import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from mwa_hyperdrive_amd import _lib, synthetic
from mwa_hyperdrive_amd.modeller import AnalyticBeam, NoBeam, SkyModeller
from oracle import oracle as orc
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


def _flagged_setup(n_total=20, flagged=(2, 7, 13), comps=60):
    base = synthetic.shrink(synthetic.config(2, scale=comps / 8000.0), tiles_n=n_total, chans=11, timesteps=1)
    keep = [t for t in range(n_total) if t not in flagged]
    delays = np.zeros((n_total, 16), np.uint32)
    gains = np.ones((n_total, 32))
    delays[4, 3] = 32       # an unflagged tile with a dead dipole
    delays[7, 9] = 32       # a flagged one: must not matter
    gains[11, 20] = 0.0     # another unflagged tile, dead Y dipole
    return base, keep, list(flagged), delays, gains


def _gpu(base, keep, flagged, delays, gains, factored=None):
    m = SkyModeller(AnalyticBeam(delays, gains), base.srclist, 0, base.xyzs[keep], base.freqs, flagged,
                    base.phase_centre, 0.0, base.latitude)
    nbl = len(keep) * (len(keep) - 1) // 2
    shape = (len(base.freqs), nbl, 8)
    if factored is None:
        _, _, v64 = m.model_timestep(float(base.lsts[0]), base.latitude, want_f64=True)
        return v64, m
    d32 = torch.zeros(shape, dtype=torch.float32, device="cuda")
    d64 = torch.zeros(shape, dtype=torch.float64, device="cuda")
    _lib.call("sm_model_timestep_factored", m._h, float(base.lsts[0]), base.latitude, factored,
              C.c_void_p(d32.data_ptr()), C.c_void_p(d64.data_ptr()))
    torch.cuda.synchronize()
    return d64.cpu().numpy(), m


def _oracle(base, keep, flagged, delays, gains):
    o = orc.Obs(base.xyzs[keep], base.freqs, base.phase_centre, float(base.lsts[0]), base.latitude,
                flagged_tiles=flagged, beam="analytic", delays=delays, gains=gains)
    return orc.model_timestep(o, base.srclist)[1]


def test_flagged_tiles_direct_and_factored():
    base, keep, flagged, delays, gains = _flagged_setup()
    o64 = _oracle(base, keep, flagged, delays, gains)
    v64, m = _gpu(base, keep, flagged, delays, gains)
    err = rel_err(v64, o64)
    print(f"flagged tiles, FP64 direct: rel err {err:.3e}, unique tiles {m.info()['num_unique_tiles']}")
    assert err <= 1e-12
    assert m.info()["num_unique_tiles"] == 3
    f64, _ = _gpu(base, keep, flagged, delays, gains, factored=0)
    err = rel_err(f64, o64)
    print(f"flagged tiles, split precision fp16: rel err {err:.3e}")
    assert err <= 1e-6


@pytest.mark.parametrize("n_comps", [1, 15, 17, 33])
def test_factored_uneven_channels_and_ragged_component_counts(n_comps):
    obs = synthetic.shrink(synthetic.config(3, scale=n_comps / 20000), tiles_n=30, chans=1, timesteps=1)
    # Channels with uneven gaps (and a ragged last 8-channel group): the phasor kernel takes
    # one FP64 sincos per channel here.
    obs.freqs = np.array([167.00e6, 167.04e6, 167.12e6, 167.13e6, 168.00e6, 168.04e6, 168.08e6, 170.5e6,
                          170.54e6, 171.0e6, 172.2e6])
    m = SkyModeller(NoBeam(), obs.srclist, 0, obs.xyzs, obs.freqs, [], obs.phase_centre, 0.0, obs.latitude)
    assert m.info()["num_points"] == obs.num_components
    o = orc.Obs(obs.xyzs, obs.freqs, obs.phase_centre, float(obs.lsts[0]), obs.latitude)
    o64 = orc.model_timestep(o, obs.srclist)[1]
    shape = (len(obs.freqs), obs.num_baselines, 8)
    d32 = torch.zeros(shape, dtype=torch.float32, device="cuda")
    d64 = torch.zeros(shape, dtype=torch.float64, device="cuda")
    _lib.call("sm_model_timestep_factored", m._h, float(obs.lsts[0]), obs.latitude, 0,
              C.c_void_p(d32.data_ptr()), C.c_void_p(d64.data_ptr()))
    torch.cuda.synchronize()
    err = rel_err(d64.cpu().numpy(), o64)
    print(f"{obs.num_components} components, uneven channels: rel err {err:.3e}")
    assert err <= 1e-6


def test_factored_reused_across_sky_sizes():
    """A big sky, then one small source through update_with_a_source, then the big sky's
    first source again: every call must match a fresh oracle run (no stale buffers)."""
    obs = synthetic.shrink(synthetic.config(3, scale=700 / 20000), tiles_n=40, chans=9, timesteps=1)
    m = SkyModeller(NoBeam(), obs.srclist, 0, obs.xyzs, obs.freqs, [], obs.phase_centre, 0.0, obs.latitude)
    shape = (len(obs.freqs), obs.num_baselines, 8)
    names = list(obs.srclist.keys())
    from mwa_hyperdrive_amd.srclist import SourceList
    for step, name in enumerate([None, names[3], names[0]]):
        if name is None:
            sky = obs.srclist
        else:
            sky = SourceList([(name, obs.srclist[name])])
            m.update_with_a_source(obs.srclist[name], obs.phase_centre)
        o = orc.Obs(obs.xyzs, obs.freqs, obs.phase_centre, float(obs.lsts[0]), obs.latitude)
        o64 = orc.model_timestep(o, sky)[1]
        d32 = torch.zeros(shape, dtype=torch.float32, device="cuda")
        d64 = torch.zeros(shape, dtype=torch.float64, device="cuda")
        _lib.call("sm_model_timestep_factored", m._h, float(obs.lsts[0]), obs.latitude, 0,
                  C.c_void_p(d32.data_ptr()), C.c_void_p(d64.data_ptr()))
        torch.cuda.synchronize()
        err = rel_err(d64.cpu().numpy(), o64)
        print(f"call {step}: {m.info()['num_points']} points, rel err {err:.3e}")
        assert err <= 1e-6


def test_per_tile_beam_classes_fp64_and_split_vs_oracle():
    """Per-tile beams with several beam rows, as in the combined 256-antenna case: 40 tiles,
    7 with a dead dipole, two pairs of them sharing the same dead dipole (one beam row each),
    all three component types. The FP64 path splits the baselines by beam-row class: pairs of
    two common-row tiles on the tile-pair kernel, pairs of a class-r tile and a common-row
    tile on the row-by-column kernel with J_r S J_0^H per class (both orientations), the
    remaining pairs with per-visibility J_p S J_q^H, and the shapelets through the per-class
    coherency table. With the class tables disabled (sm_set_class_table_limit(0)) the class-r
    pairs take the shared-coherency baseline kernel per segment and the shapelets form
    J_p S J_q^H per visibility; that path runs here too. Every baseline must meet the FP64 bar;
    the split-precision path the fp16 bar."""
    base = synthetic.shrink(synthetic.config(2, scale=120 / 8000.0), tiles_n=40, chans=11, timesteps=1)
    n = 40
    delays = np.zeros((n, 16), np.uint32)
    gains = np.ones((n, 32))
    for t, d in ((2, 5), (9, 5), (17, 11), (30, 11), (35, 0)):
        delays[t, d] = 32
    gains[12, 3] = 0.0
    gains[21, 40 % 32] = 0.0
    m = SkyModeller(AnalyticBeam(delays, gains), base.srclist, 0, base.xyzs, base.freqs, [], base.phase_centre, 0.0,
                    base.latitude)
    assert m.info()["num_unique_tiles"] == 6
    o = orc.Obs(base.xyzs, base.freqs, base.phase_centre, float(base.lsts[0]), base.latitude, beam="analytic",
                delays=delays, gains=gains)
    o64 = orc.model_timestep(o, base.srclist)[1]
    _, _, v64 = m.model_timestep(float(base.lsts[0]), base.latitude, want_f64=True)
    err = rel_err(v64, o64)
    print(f"per-tile classes, FP64 direct: rel err {err:.3e}")
    assert err <= 1e-12
    try:
        _lib.call("sm_set_class_table_limit", 0)
        _, _, s64 = m.model_timestep(float(base.lsts[0]), base.latitude, want_f64=True)
    finally:
        _lib.call("sm_set_class_table_limit", -1)
    err = rel_err(s64, o64)
    print(f"per-tile classes, FP64 per-segment path: rel err {err:.3e}")
    assert err <= 1e-12
    shape = (len(base.freqs), n * (n - 1) // 2, 8)
    d32 = torch.zeros(shape, dtype=torch.float32, device="cuda")
    d64 = torch.zeros(shape, dtype=torch.float64, device="cuda")
    _lib.call("sm_model_timestep_factored", m._h, float(base.lsts[0]), base.latitude, 0,
              C.c_void_p(d32.data_ptr()), C.c_void_p(d64.data_ptr()))
    torch.cuda.synchronize()
    err = rel_err(d64.cpu().numpy(), o64)
    print(f"per-tile classes, split precision fp16: rel err {err:.3e}")
    assert err <= 1e-6


def flagged_channel_freqs(n_coarse=3, fine=32, width=40e3, start=167e6):
    """Unflagged fine channels of an MWA-like band: per 32-channel coarse channel the two edge
    channels and the centre channel are flagged, as hyperdrive's default flags do."""
    keep = [c * fine + k for c in range(n_coarse) for k in range(fine) if k not in (0, fine - 1, fine // 2)]
    return start + width * np.asarray(keep, np.float64)


@pytest.mark.parametrize("beam", ["shared", "per-tile"])
def test_flagged_channels_vs_oracle(beam):
    """Flagged fine channels leave gaps in the channel list: the direct kernels run on runs of
    evenly spaced channels, so every channel keeps the evenly spaced kernels. FP64 path at the
    FP64 bar, the precision path at 1e-6."""
    base = synthetic.shrink(synthetic.config(5, scale=0.004), tiles_n=40, chans=1, timesteps=1)
    base.freqs = flagged_channel_freqs()
    if beam == "shared":
        base.delays = np.zeros((40, 16), np.uint32)
        base.gains = np.ones((40, 32))
    o = orc.Obs(base.xyzs, base.freqs, base.phase_centre, float(base.lsts[0]), base.latitude, beam="analytic",
                delays=base.delays, gains=base.gains)
    o64 = orc.model_timestep(o, base.srclist)[1]
    m = SkyModeller(AnalyticBeam(base.delays, base.gains), base.srclist, 0, base.xyzs, base.freqs, [],
                    base.phase_centre, 0.0, base.latitude)
    _, _, v64 = m.model_timestep(float(base.lsts[0]), base.latitude, want_f64=True)
    err = rel_err(v64, o64)
    print(f"{beam} beam, {len(base.freqs)} unflagged channels: FP64 rel err {err:.3e}")
    assert err <= 1e-12
    shape = (len(base.freqs), base.num_baselines, 8)
    d32 = torch.zeros(shape, dtype=torch.float32, device="cuda")
    d64 = torch.zeros(shape, dtype=torch.float64, device="cuda")
    _lib.call("sm_model_timestep_factored", m._h, float(base.lsts[0]), base.latitude, 0,
              C.c_void_p(d32.data_ptr()), C.c_void_p(d64.data_ptr()))
    torch.cuda.synchronize()
    err = rel_err(d64.cpu().numpy(), o64)
    print(f"{beam} beam: precision path rel err {err:.3e}")
    assert err <= 1e-6
