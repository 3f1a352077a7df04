"""Parity of the HIP product path with the reference's golden vectors and the CPU oracle.

Bars: the f32 drop-in output equals the reference goldens bit for bit (epsilon 0, as
src/model/tests/gpu.rs:43-46 demands of hyperdrive's FP64 GPU path); the f64 output is
within 1e-12 of max|V| of the oracle's f64 accumulators. Every test runs the product
through its C ABI (libskymodel_hip.so) on cuda:0.
"""

import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from mwa_hyperdrive_amd import synthetic
from mwa_hyperdrive_amd.modeller import AnalyticBeam, NoBeam, SkyModeller
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
from tests.fixtures import (
    FREQS,
    MWA_LAT_RAD,
    NO_BEAM_CASES,
    PHASE_CENTRE,
    XYZS,
    golden,
    multiple_components,
    single,
)
from tests.helpers import f32_ulps, gpu_model, oracle_model, rel_err

pytestmark = pytest.mark.gpu

TOL_F64 = 1e-12  # relative to max|V|


@pytest.fixture(params=[0, 1], ids=["fast-kernels", "baseline-per-lane"])
def variant(request):
    """Runs a test with the fastest kernels (tile-pair phasors, channel-per-lane shapelets)
    and again with the baseline-per-lane kernels."""
    from mwa_hyperdrive_amd import _lib
    _lib.call("sm_set_direct_variant", request.param)
    yield request.param
    _lib.call("sm_set_direct_variant", 0)


def fixture_modeller(srclist, pols=0):
    return SkyModeller(NoBeam(), srclist, pols, XYZS, FREQS, [], PHASE_CENTRE, 0.0, MWA_LAT_RAD)


@pytest.mark.parametrize("checker,builder,pos,flux", NO_BEAM_CASES)
def test_goldens_bit_exact(checker, builder, pos, flux, variant):
    m = fixture_modeller(single(builder, pos, flux))
    vis, _ = m.model_timestep(0.0, MWA_LAT_RAD)
    np.testing.assert_array_equal(vis, golden(checker))


@pytest.mark.parametrize("builder,checker", [("gaussian", "test_multiple_gaussian_components"),
                                             ("shapelet", "test_multiple_shapelet_components")])
def test_multiple_components_goldens(builder, checker):
    m = fixture_modeller(multiple_components(builder))
    vis, _ = m.model_timestep(0.0, MWA_LAT_RAD)
    np.testing.assert_array_equal(vis, golden(checker))


def test_autos_golden():
    for builder in ("point", "gaussian", "shapelet"):
        m = fixture_modeller(single(builder, radec_from_degrees(1.0, -27.0), "list"))
        autos = m.model_timestep_autos_with(0.0, MWA_LAT_RAD)
        np.testing.assert_array_equal(autos, golden("test_model_timestep_autos_with_point"))


def test_update_with_a_source_matches_off_zenith_golden():
    m = fixture_modeller(single("point", PHASE_CENTRE, "list"))
    src = Source([SourceComponent(*radec_from_degrees(1.0, -27.0), Point(),
                                  FluxList([FluxDensity(150e6, 1.0), FluxDensity(175e6, 3.0),
                                            FluxDensity(200e6, 2.0)]))])
    m.update_with_a_source(src, PHASE_CENTRE)
    vis, _ = m.model_timestep(0.0, MWA_LAT_RAD)
    np.testing.assert_array_equal(vis, golden("test_list_off_zenith_visibilities"))


def _mixed_obs(beam, n_tiles=24, chans=None, dead=False, seed=7, comps=60):
    obs = synthetic.shrink(synthetic.config(2, scale=comps / 8000.0), tiles_n=n_tiles, chans=37, timesteps=2)
    if chans is not None:
        obs.freqs = np.asarray(chans, np.float64)
    obs.beam = beam
    if beam == "analytic":
        obs.delays = np.zeros((n_tiles, 16), np.uint32)
        obs.gains = np.ones((n_tiles, 32))
        if dead:
            rng = np.random.default_rng(seed)
            for t in rng.choice(n_tiles, size=5, replace=False):
                obs.delays[t, rng.integers(0, 16)] = 32
    return obs


@pytest.mark.parametrize("beam,dead", [("none", False), ("analytic", False), ("analytic", True)])
def test_mixed_sky_vs_oracle(beam, dead, variant):
    """Points, Gaussians and shapelets with power-law, curved and list fluxes; identity,
    shared analytic and per-tile (dead-dipole) analytic beams; 37 channels so the last chunk
    is ragged."""
    obs = _mixed_obs(beam, dead=dead)
    for t in range(2):
        v32, v64, _, m = gpu_model(obs, t)
        o32, o64 = oracle_model(obs, t)
        err = rel_err(v64, o64)
        print(f"variant={variant} beam={beam} dead={dead} t={t} rel err {err:.3e} unique tiles {m.info()['num_unique_tiles']}")
        assert err <= TOL_F64
        assert f32_ulps(v32, o32) <= 2.0 or rel_err(v32, o32) < 1e-7
    if dead:
        assert m.info()["num_unique_tiles"] > 1


@pytest.mark.parametrize("case", ["high_order", "wide_window", "far_channels"])
def test_shapelet_staging_exits_vs_oracle(case, variant):
    """Shapelets that leave the shared-beam shapelet kernel's staged table window, mixed with
    shapelets that stay in it: basis orders above the 16 staged rows (high_order), windows
    wider than the 24 staged cells with adjacent channels under half a cell apart
    (wide_window: 140 channels of 40 kHz, majors of 5-20 arcmin), and adjacent channels
    many cells apart with positions clamped at the table's edge (far_channels: 1.28 MHz
    channels, 30 arcmin majors)."""
    obs = _mixed_obs("analytic", comps=60)
    rng = np.random.default_rng(11)
    arcmin = math.radians(1.0 / 60.0)
    if case == "high_order":
        maj = rng.uniform(0.5, 5.0, 8) * arcmin
        orders = [(18, 3), (2, 17), (20, 20), (0, 16)]
    elif case == "wide_window":
        obs.freqs = 167e6 + 40e3 * np.arange(140, dtype=np.float64)
        maj = np.linspace(5.0, 20.0, 8) * arcmin
        orders = [(0, 0), (3, 1), (6, 4), (10, 10)]
    else:
        obs.freqs = 150e6 + 1.28e6 * np.arange(70, dtype=np.float64)
        maj = np.full(8, 30.0) * arcmin
        orders = [(0, 0), (1, 2), (5, 5), (8, 0)]
    names = list(obs.srclist.keys())
    for i, name in enumerate(names[:8]):
        comp = obs.srclist[name].components[0]
        coeffs = [(n1, n2, float(rng.normal())) for n1, n2 in orders] + [(1, 1, 0.5)]
        comp.comp_type = Shapelet(float(maj[i]), float(maj[i] * 0.6), float(rng.uniform(0, np.pi)), coeffs)
    v32, v64, _, _ = gpu_model(obs, 0)
    o32, o64 = oracle_model(obs, 0)
    err = rel_err(v64, o64)
    print(f"case={case} variant={variant} rel err {err:.3e}")
    assert err <= TOL_F64


def test_uneven_channels_vs_oracle(variant):
    """Channels that are not evenly spaced take the per-channel sincos/exp kernel."""
    freqs = 167e6 + 40e3 * np.array([0, 1, 2, 4, 5, 9, 10, 11, 12, 13, 20, 21], dtype=np.float64)
    obs = _mixed_obs("analytic", chans=freqs)
    v32, v64, _, _ = gpu_model(obs, 0)
    o32, o64 = oracle_model(obs, 0)
    assert rel_err(v64, o64) <= TOL_F64


def test_c1_full_size_vs_oracle(variant):
    """Configuration C1 in full: 128 tiles, 32 channels, 100 power-law points, no beam."""
    obs = synthetic.config(1)
    v32, v64, uvws, _ = gpu_model(obs, 0)
    o32, o64 = oracle_model(obs, 0)
    err = rel_err(v64, o64)
    print(f"variant={variant} C1 rel err {err:.3e}")
    assert err <= TOL_F64
    assert uvws.shape == (8128, 3)


def test_c2_full_size_channel_slice_vs_oracle(variant):
    """One C2 timestep at full size (8000 components, 8128 baselines, 768 channels) on the
    GPU; the oracle checks the first 8 channels, and linearity (model(A) + model(B) =
    model(A u B)) checks every channel."""
    from oracle import oracle as orc
    from tests.helpers import oracle_obs
    obs = synthetic.config(2)
    v32, v64, _, m = gpu_model(obs, 3)
    o32, o64 = orc.model_timestep_slice(oracle_obs(obs, 3), obs.srclist, 0, 8)
    err = rel_err(v64[:8], o64)
    print(f"variant={variant} C2 slice rel err {err:.3e}")
    assert err <= TOL_F64
    items = list(obs.srclist.items())
    a = synthetic.shrink(obs)
    a.srclist = SourceList(items[: len(items) // 2])
    b = synthetic.shrink(obs)
    b.srclist = SourceList(items[len(items) // 2:])
    _, a64, _, _ = gpu_model(a, 3)
    _, b64, _, _ = gpu_model(b, 3)
    assert rel_err(a64 + b64, v64) <= 1e-13


def test_pols_masking():
    """mask_pols (src/model/mod.rs:264-285) after modelling."""
    obs = _mixed_obs("analytic", n_tiles=8, comps=10)
    full, _, _, _ = gpu_model(obs, 0, pols=0)
    for pols, keep in [(1, [0]), (2, [3]), (3, [0, 3]), (4, [0, 1, 3])]:
        v, _, _, _ = gpu_model(obs, 0, pols=pols)
        for k in range(4):
            sl = v[..., 2 * k:2 * k + 2]
            if k in keep:
                np.testing.assert_array_equal(sl, full[..., 2 * k:2 * k + 2])
            else:
                assert np.all(sl == 0)


def test_accumulates_into_existing_visibilities():
    """model_timestep_with adds to what the buffer holds (src/model/mod.rs:154-159)."""
    obs = _mixed_obs("none", n_tiles=6, comps=8)
    beam = NoBeam()
    m = SkyModeller(beam, obs.srclist, 0, obs.xyzs, obs.freqs, [], obs.phase_centre, 0.0, obs.latitude)
    base = np.full((len(obs.freqs), obs.num_baselines, 8), 0.5, np.float32)
    once = np.zeros_like(base)
    m.model_timestep_with(0.0, obs.latitude, once)
    m.model_timestep_with(0.0, obs.latitude, base)
    np.testing.assert_allclose(base, once + np.float32(0.5), rtol=0, atol=1e-6)


def test_analytic_beam_vs_oracle():
    import ctypes as C
    from oracle import oracle as orc
    from mwa_hyperdrive_amd import _lib
    rng = np.random.default_rng(3)
    n_dirs = 300
    az = rng.uniform(0, 2 * np.pi, n_dirs)
    za = rng.uniform(0, np.pi / 2, n_dirs)
    freqs = np.array([150e6, 167.04e6, 197e6])
    delays = rng.integers(0, 32, size=(2, 16)).astype(np.uint32)
    delays[1, 3] = 32
    gains = rng.uniform(0.5, 1.0, size=(2, 32))
    d_az = torch.tensor(az, dtype=torch.float64, device="cuda")
    d_za = torch.tensor(za, dtype=torch.float64, device="cuda")
    out = torch.zeros((2, 3, n_dirs, 8), dtype=torch.float64, device="cuda")
    _lib.call("sm_analytic_beam", 2, delays.ctypes.data, gains.ctypes.data, 3, freqs.ctypes.data, n_dirs,
              C.c_void_p(d_az.data_ptr()), C.c_void_p(d_za.data_ptr()), C.c_void_p(out.data_ptr()))
    got = out.cpu().numpy()
    for t in range(2):
        for fi, f in enumerate(freqs):
            for d in range(0, n_dirs, 7):
                want = orc.analytic_beam(az[d], za[d], f, delays[t], gains[t])
                np.testing.assert_allclose(got[t, fi, d], want, rtol=0, atol=1e-13)


def test_analytic_beam_vs_oracle_frequency_blocks():
    """More frequencies than one beam block holds (8), in a count that leaves a partial block,
    unevenly spaced, with a dead dipole: every (frequency, direction) against the oracle."""
    import ctypes as C
    from oracle import oracle as orc
    from mwa_hyperdrive_amd import _lib
    rng = np.random.default_rng(4)
    n_dirs = 70
    az = rng.uniform(0, 2 * np.pi, n_dirs)
    za = rng.uniform(0, np.pi / 2, n_dirs)
    freqs = np.sort(rng.uniform(70e6, 300e6, 19))
    delays = rng.integers(0, 32, size=(1, 16)).astype(np.uint32)
    delays[0, 9] = 32
    gains = rng.uniform(0.5, 1.0, size=(1, 32))
    d_az = torch.tensor(az, dtype=torch.float64, device="cuda")
    d_za = torch.tensor(za, dtype=torch.float64, device="cuda")
    out = torch.zeros((1, freqs.size, n_dirs, 8), dtype=torch.float64, device="cuda")
    _lib.call("sm_analytic_beam", 1, delays.ctypes.data, gains.ctypes.data, freqs.size, freqs.ctypes.data, n_dirs,
              C.c_void_p(d_az.data_ptr()), C.c_void_p(d_za.data_ptr()), C.c_void_p(out.data_ptr()))
    got = out.cpu().numpy()
    for fi, f in enumerate(freqs):
        for d in range(n_dirs):
            want = orc.analytic_beam(az[d], za[d], f, delays[0], gains[0])
            np.testing.assert_allclose(got[0, fi, d], want, rtol=0, atol=1e-13)


def test_device_buffer_entry_matches_host_entry():
    obs = _mixed_obs("analytic", n_tiles=16, comps=30)
    v32, v64, _, m = gpu_model(obs, 1)
    d32 = torch.zeros(v32.shape, dtype=torch.float32, device="cuda")
    d64 = torch.zeros(v64.shape, dtype=torch.float64, device="cuda")
    m.model_timestep_with_device(float(obs.lsts[1]), obs.latitude, d32, d64)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d32.cpu().numpy(), v32)
    np.testing.assert_array_equal(d64.cpu().numpy(), v64)
