"""Reference behaviours of the product path checked against the reference's own numbers and
against the oracle, on the GPU:

- flux estimation as the product does it (the handle's host estimation and the device's
  power-law / curved-power-law extrapolation), driven through visibilities of a point at the
  phase centre with no beam, where V = S exactly: the reference's KATs for negative, single-
  entry and sign-changing lists (src/srclist/types/flux_density/tests.rs:31-217,
  src/srclist/types/components/tests.rs:84-104);
- the curved-power-law re-reference to 150 MHz (src/model/gpu.rs:664-680), read back from the
  device buffers as the reference's regression test does (src/model/tests/gpu.rs:558-601);
- a mixed sky whose fluxes take every estimation branch, against the oracle at 1e-12;
- the full-size 256-tile case (60 500 components, 16 beam rows) and the full-size
  20 000-point case, every component on the GPU, checked against the oracle on a sample of
  baselines that covers every beam-row category and every 64-tile column group.
"""

import ctypes as C
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from mwa_hyperdrive_amd import _lib, synthetic
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
from oracle import oracle as orc
from tests.fixtures import MWA_LAT_RAD, PHASE_CENTRE, XYZS
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

EPS = np.finfo(float).eps
SM_POINT, SM_LIST, SM_POWER_LAW, SM_CURVED_POWER_LAW = 0, 0, 1, 2


@pytest.fixture(params=[0, 1], ids=["fast-kernels", "baseline-per-lane"])
def variant(request):
    _lib.call("sm_set_direct_variant", request.param)
    yield request.param
    _lib.call("sm_set_direct_variant", 0)


def zenith_flux(flux, freqs):
    """The product's f64 visibility XX (real part) of one point at the phase centre, no beam:
    the instrumental XX = I - Q of the flux density the product estimates at each frequency."""
    pt = SourceComponent(*PHASE_CENTRE, Point(), flux)
    m = SkyModeller(NoBeam(), SourceList([("z", Source([pt]))]), 0, XYZS, freqs, [], PHASE_CENTRE, 0.0, MWA_LAT_RAD)
    _, _, v64 = m.model_timestep(0.0, MWA_LAT_RAD, want_f64=True)
    # Every baseline sees the same value: the phase is zero at the phase centre.
    assert np.all(v64[:, :, 0] == v64[:, :1, 0])
    return v64[:, 0, 0]


@pytest.mark.parametrize("fds,cases", [
    # estimate_with_negative_fds (flux_density/tests.rs:31-85)
    ([(150e6, -1.0), (175e6, -3.0), (200e6, -2.0)],
     [(140e6, -0.611583722518741), (150e6, -1.0), (160e6, -1.584007188344133), (190e6, -2.3370702845538984),
      (210e6, -1.72460308893938)]),
    # estimate_with_negative_to_positive_fds (flux_density/tests.rs:87-131): a sign change is
    # a linear fit between the bracketing entries, extrapolated linearly outside
    ([(100e6, -1.0), (200e6, 1.0)], [(90e6, -1.0 - 0.2), (150e6, 0.0), (210e6, 1.0 + 0.2)]),
    # estimate_with_negative_to_positive_fds_2 (flux_density/tests.rs:133-217)
    ([(100e6, -1.0), (150e6, -0.5), (200e6, 1.0)],
     [(90e6, -1.1973550404744007), (145e6, -0.5298337000434852), (155e6, -0.5 + 1.5 / 50.0 * 5.0),
      (210e6, 1.0 + 1.5 / 50.0 * 10.0)]),
])
def test_list_flux_kats_through_product(fds, cases):
    flux = FluxList([FluxDensity(f, i) for f, i in fds])
    got = zenith_flux(flux, [f for f, _ in cases])
    for (f, want), g in zip(cases, got):
        assert abs(g - want) <= 4 * EPS * max(1.0, abs(want)), (f, g, want)


def test_two_entry_list_kat_through_product():
    """components/tests.rs:84-104: a two-entry same-sign list estimated at 180 MHz with the
    spectral index fitted to both entries."""
    flux = FluxList([FluxDensity(80e6, 2.13017), FluxDensity(240e6, 0.33037)])
    si = math.log(0.33037 / 2.13017) / math.log(240e6 / 80e6)
    want = 2.13017 * (180e6 / 80e6) ** si
    (got,) = zenith_flux(flux, [180e6])
    assert abs(got - want) <= 4 * EPS * want


def test_single_entry_list_takes_the_default_spectral_index():
    """A one-entry list extrapolates with DEFAULT_SPEC_INDEX = -0.8 at every frequency, its own
    included (src/constants.rs:8, flux_density/mod.rs:176-185: no exact-match test on this
    branch)."""
    flux = FluxList([FluxDensity(150e6, 2.5, 0.25)])
    freqs = [120e6, 150e6, 150e6 + 5e-4, 180e6, 230e6]
    got = zenith_flux(flux, freqs)
    for f, g in zip(freqs, got):
        want = (2.5 - 0.25) * (f / 150e6) ** -0.8
        assert abs(g - want) <= 4 * EPS * want, (f, g, want)
    assert got[1] == 2.5 - 0.25


def test_curved_power_law_re_referenced_through_product():
    """test_curved_power_law_changing_ref_freq (src/model/tests/gpu.rs:558-601): a curved power
    law given at 200 MHz is stored at 150 MHz with a re-derived spectral index."""
    comp = SourceComponent(*PHASE_CENTRE, Point(), CurvedPowerLaw(-0.8, FluxDensity(200e6, 1.0), 0.03))
    m = SkyModeller(NoBeam(), SourceList([("zenith", Source([comp]))]), 0, XYZS, [150e6], [], PHASE_CENTRE, 0.0,
                    MWA_LAT_RAD)
    fd = (C.c_double * 8)()
    si, q = C.c_double(), C.c_double()
    _lib.call("sm_get_flux_params", m._h, SM_POINT, SM_CURVED_POWER_LAW, 0, 0, fd, C.byref(si), C.byref(q))
    assert abs(fd[0] - 1.261912575563708) <= EPS
    assert abs(si.value - -0.8172609243471072) <= EPS
    assert q.value == 0.03
    # The visibilities over a band, with a beam, from the re-referenced parameters on the
    # device equal the oracle's direct estimation from the 200 MHz reference.
    obs = synthetic.shrink(synthetic.config(2, scale=30 / 8000.0), tiles_n=12, chans=21, timesteps=1)
    sky = SourceList()
    rng = np.random.default_rng(4)
    for k, (ra, dec) in enumerate(zip(*synthetic.cap_positions(rng, 12, 10.0))):
        ref = [200e6, 120e6, 150e6, 171.3e6][k % 4]
        fdk = FluxDensity(ref, float(rng.lognormal()), 0.1, -0.05, 0.02)
        sky[f"c{k}"] = Source([SourceComponent(float(ra), float(dec), Point(),
                                               CurvedPowerLaw(float(rng.normal(-0.8, 0.2)), fdk,
                                                              float(rng.normal(0, 0.3))))])
    mm = SkyModeller(AnalyticBeam(obs.delays, obs.gains), sky, 0, obs.xyzs, obs.freqs, [], obs.phase_centre, 0.0,
                     obs.latitude)
    _, _, v64 = mm.model_timestep(float(obs.lsts[0]), obs.latitude, want_f64=True)
    o = orc.Obs(obs.xyzs, obs.freqs, obs.phase_centre, float(obs.lsts[0]), obs.latitude, beam="analytic",
                delays=obs.delays, gains=obs.gains)
    err = rel_err(v64, orc.model_timestep(o, sky)[1])
    print(f"curved power laws with 120/150/171.3/200 MHz references: rel err {err:.3e}")
    assert err <= 1e-12


def branch_sky(rng, n_tiles_extent=10.0):
    """Every flux-estimation branch on every component type: power laws and curved power laws
    referenced away from 150 MHz, full-Stokes; lists with one entry, with negative entries,
    with sign changes inside the band, with an entry on a channel and with entries only on one
    side of the band."""
    lists = [
        [(150e6, 1.7, 0.1, 0.05, -0.02)],
        [(168e6, -1.0, 0, 0, 0), (168.5e6, 2.0, 0.1, 0, 0), (170e6, 1.5, 0, 0.2, 0)],
        [(150e6, -1.0, 0, 0, 0), (175e6, -3.0, 0, 0, 0), (200e6, -2.0, 0, 0, 0)],
        [(167.04e6, 3.0, 0.3, -0.1, 0.05), (167.6e6, 2.5, 0.2, 0, 0)],
        [(120e6, 4.0, 0, 0, 0), (140e6, 3.0, 0, 0, 0)],
        [(190e6, 0.5, 0, 0, 0), (200e6, -0.25, 0, 0, 0)],
        [(100e6, -1.0, 0, 0, 0), (150e6, -0.5, 0, 0, 0), (200e6, 1.0, 0, 0, 0)],
    ]
    fluxes = []
    for ent in lists:
        fluxes.append(FluxList([FluxDensity(*e) for e in ent]))
    for ref in (120e6, 200e6, 167.5e6):
        fluxes.append(PowerLaw(float(rng.normal(-0.8, 0.3)), FluxDensity(ref, float(rng.lognormal()), 0.1, 0.05, -0.1)))
        fluxes.append(CurvedPowerLaw(float(rng.normal(-0.8, 0.3)), FluxDensity(ref, float(rng.lognormal()), -0.1, 0.0,
                                                                               0.05), float(rng.normal(0, 0.3))))
    sky = SourceList()
    ra, dec = synthetic.cap_positions(rng, 3 * len(fluxes), n_tiles_extent)
    arcmin = math.radians(1 / 60)
    k = 0
    for ty in range(3):
        for flux in fluxes:
            if ty == 0:
                ct = Point()
            elif ty == 1:
                ct = Gaussian(3 * arcmin, 1.5 * arcmin, float(rng.uniform(0, math.pi)))
            else:
                ct = Shapelet(3 * arcmin, 2 * arcmin, float(rng.uniform(0, math.pi)),
                              [(int(rng.integers(0, 6)), int(rng.integers(0, 6)), float(rng.normal())) for _ in range(6)])
            sky[f"b{k}"] = Source([SourceComponent(float(ra[k]), float(dec[k]), ct, flux)])
            k += 1
    return sky


@pytest.mark.parametrize("beam", ["none", "per-tile"])
def test_flux_branches_mixed_sky_vs_oracle(beam, variant):
    rng = np.random.default_rng(21)
    nt = 20
    xyzs = synthetic.tiles(rng, nt, 1200.0)
    freqs = 167e6 + 40e3 * np.arange(40)  # 167.00 .. 168.56 MHz
    sky = branch_sky(rng)
    delays = gains = None
    if beam == "per-tile":
        delays = np.zeros((nt, 16), np.uint32)
        gains = np.ones((nt, 32))
        delays[3, 7] = delays[11, 2] = delays[15, 2] = 32
    m = SkyModeller(AnalyticBeam(delays, gains) if beam != "none" else NoBeam(), sky, 0, xyzs, freqs, [],
                    synthetic.PHASE_CENTRE, 0.0, MWA_LAT_RAD)
    _, _, v64 = m.model_timestep(0.2, MWA_LAT_RAD, want_f64=True)
    o = orc.Obs(xyzs, freqs, synthetic.PHASE_CENTRE, 0.2, MWA_LAT_RAD, beam="analytic" if beam != "none" else "none",
                delays=delays, gains=gains)
    err = rel_err(v64, orc.model_timestep(o, sky)[1])
    print(f"variant={variant} beam={beam}: every flux branch, rel err {err:.3e}")
    assert err <= 1e-12


# --------------------------------------------------------------------------- full size
def baseline_sample(nt, extra_tiles, n_random, seed):
    """Every baseline of the given tiles plus n_random others, ascending."""
    rng = np.random.default_rng(seed)
    nbl = nt * (nt - 1) // 2
    pick = set(int(b) for b in rng.choice(nbl, size=n_random, replace=False))
    for t in extra_tiles:
        for u in range(nt):
            if u == t:
                continue
            i, j = min(t, u), max(t, u)
            pick.add(i * nt - i * (i + 1) // 2 + (j - i - 1))
    return np.array(sorted(pick), np.int32)


def c5_sample(obs):
    """Baselines of the 256-tile case that reach every per-tile-beam path: all baselines of two
    tiles with a dead dipole (beam row r with each class-0 tile, in both baseline orientations
    and in all four 64-tile column groups, and with the other rows) and of one class-0 tile,
    plus random baselines."""
    dead = [t for t in range(obs.num_tiles) if np.any(obs.delays[t] == 32)]
    live = [t for t in range(obs.num_tiles) if t not in dead]
    return baseline_sample(obs.num_tiles, [dead[0], dead[len(dead) // 2], live[-1]], 300, 55)


def device_timestep(m, obs, t, factored=None):
    shape = (len(obs.freqs), obs.num_baselines, 8)
    d32 = torch.zeros(shape, dtype=torch.float32, device="cuda")
    d64 = torch.zeros(shape, dtype=torch.float64, device="cuda")
    if factored is None:
        m.model_timestep_with_device(float(obs.lsts[t]), obs.latitude, d32, d64)
    else:
        _lib.call("sm_model_timestep_factored", m._h, float(obs.lsts[t]), obs.latitude, factored,
                  C.c_void_p(d32.data_ptr()), C.c_void_p(d64.data_ptr()))
    torch.cuda.synchronize()
    return d64


@pytest.fixture(scope="module")
def c5_reference():
    obs = synthetic.config(5)
    t = 5
    bls = c5_sample(obs)
    o = orc.Obs(obs.xyzs, obs.freqs, obs.phase_centre, float(obs.lsts[t]), obs.latitude, beam="analytic",
                delays=obs.delays, gains=obs.gains)
    chans = [0, 1, 383]
    ref = np.concatenate([orc.model_timestep_baselines(o, obs.srclist, c, 1, bls)[1] for c in chans])
    m = SkyModeller(AnalyticBeam(obs.delays, obs.gains), obs.srclist, 0, obs.xyzs, obs.freqs, [], obs.phase_centre,
                    0.0, obs.latitude)
    info = m.info()
    assert info["num_unique_tiles"] >= 16
    assert (info["num_points"], info["num_gaussians"], info["num_shapelets"]) == (50000, 10000, 500)
    return obs, t, bls, chans, ref, m


@pytest.mark.parametrize("path", ["fp64", "fp64-segments", "split-fp16"])
def test_c5_full_size_vs_oracle_sample(c5_reference, path):
    """The 256-tile, 384-channel case with all 60 500 components and per-tile beams, modelled
    whole on the GPU: FP64 through the row-by-column and per-class paths, FP64 with the class
    tables disabled (the per-segment fallback), and points on the split-precision MFMA path."""
    obs, t, bls, chans, ref, m = c5_reference
    try:
        if path == "fp64-segments":
            _lib.call("sm_set_class_table_limit", 0)
        d64 = device_timestep(m, obs, t, factored=0 if path == "split-fp16" else None)
    finally:
        _lib.call("sm_set_class_table_limit", -1)
    got = d64[chans][:, torch.as_tensor(bls, device="cuda").long()].cpu().numpy()
    err = rel_err(got, ref)
    print(f"C5 {path}: {len(bls)} baselines x {len(chans)} channels, rel err {err:.3e}")
    assert err <= (1e-6 if path == "split-fp16" else 1e-12)


def test_c3_full_size_vs_oracle_sample():
    """The 20 000-point, 768-channel, 128-tile case, whole on the GPU through the FP64 direct
    kernel and the fp16 and bf16 split-precision MFMA paths, against the oracle on a baseline
    sample at the first, a middle and the last channel."""
    obs = synthetic.config(3)
    t = 9
    bls = baseline_sample(obs.num_tiles, [0, 77, 127], 400, 33)
    o = orc.Obs(obs.xyzs, obs.freqs, obs.phase_centre, float(obs.lsts[t]), obs.latitude)
    chans = [0, 400, 767]
    ref = np.concatenate([orc.model_timestep_baselines(o, obs.srclist, c, 1, bls)[1] for c in chans])
    m = SkyModeller(NoBeam(), obs.srclist, 0, obs.xyzs, obs.freqs, [], obs.phase_centre, 0.0, obs.latitude)
    idx = torch.as_tensor(bls, device="cuda").long()
    for name, fac, tol in (("fp64", None, 1e-12), ("fp16", 0, 1e-6), ("bf16", 1, 5e-6)):
        got = device_timestep(m, obs, t, fac)[chans][:, idx].cpu().numpy()
        err = rel_err(got, ref)
        print(f"C3 {name}: {len(bls)} baselines x {len(chans)} channels, rel err {err:.3e}")
        assert err <= tol
