"""Split-precision MFMA point path (sm_model_timestep_factored).

Bar: relative visibility error <= 1e-6 of max|V| against the FP64 oracle for fp16 hi+lo
operands. bf16 hi+lo keeps 16 significand bits and does not meet that bar (measured 3.4e-6 on
a C3 timestep): test_bf16_meets_the_1e6_bar states the bar for bf16 as an expected failure,
and the other tests hold bf16 to its measured bound (BF16_MEASURED, 5e-6) so that a
regression still shows. With per-tile beams, baselines outside the most common beam row take
the FP64 direct kernel and must then meet the FP64 bar."""

import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from mwa_hyperdrive_amd import _lib, synthetic
from mwa_hyperdrive_amd.modeller import AnalyticBeam, NoBeam, SkyModeller
from tests.helpers import gpu_model, oracle_model, rel_err

pytestmark = pytest.mark.gpu

FP16_TOL = 1e-6
BF16_MEASURED = 5e-6  # bf16's measured error bound, not the bar


def test_mfma_layout_selftest():
    rng = np.random.default_rng(0)
    a = rng.integers(-8, 8, size=(32, 16)).astype(np.float16)
    b = rng.integers(-8, 8, size=(16, 32)).astype(np.float16)
    da = torch.tensor(a, device="cuda")
    db = torch.tensor(b, device="cuda")
    dd = torch.zeros((32, 32), dtype=torch.float32, device="cuda")
    _lib.call("sm_selftest_mfma_layout", C.c_void_p(da.data_ptr()), C.c_void_p(db.data_ptr()),
              C.c_void_p(dd.data_ptr()))
    np.testing.assert_array_equal(dd.cpu().numpy(), a.astype(np.float32) @ b.astype(np.float32))


def factored(obs, t, dtype):
    beam = AnalyticBeam(obs.delays, obs.gains) if obs.beam == "analytic" else NoBeam()
    m = SkyModeller(beam, obs.srclist, 0, obs.xyzs, obs.freqs, [], obs.phase_centre, 0.0, obs.latitude)
    shape = (len(obs.freqs), obs.num_baselines, 8)
    d32 = torch.zeros(shape, dtype=torch.float32, device="cuda")
    d64 = torch.zeros(shape, dtype=torch.float64, device="cuda")
    _lib.call("sm_model_timestep_factored", m._h, float(obs.lsts[t]), obs.latitude, dtype,
              C.c_void_p(d32.data_ptr()), C.c_void_p(d64.data_ptr()))
    torch.cuda.synchronize()
    return d32.cpu().numpy(), d64.cpu().numpy(), m


@pytest.mark.parametrize("beam", ["none", "analytic"])
@pytest.mark.parametrize("n_tiles", [24, 70, 128])
def test_factored_vs_oracle(beam, n_tiles):
    """C3-like full-Stokes points; tile counts below, between and at 64-tile blocks."""
    obs = synthetic.shrink(synthetic.config(3, scale=300 / 20000), tiles_n=n_tiles, chans=12, timesteps=2)
    obs.beam = beam
    if beam == "analytic":
        obs.delays = np.zeros((n_tiles, 16), np.uint32)
        obs.gains = np.ones((n_tiles, 32))
    o32, o64 = oracle_model(obs, 1)
    for dtype, tol, name in [(0, FP16_TOL, "fp16"), (1, BF16_MEASURED, "bf16")]:
        v32, v64, m = factored(obs, 1, dtype)
        err = rel_err(v64, o64)
        print(f"{name} beam={beam} tiles={n_tiles}: rel err {err:.3e}")
        assert err <= tol
        assert m.info()["num_points"] == obs.num_components


@pytest.mark.xfail(reason="bf16 hi + lo carries 16 significand bits: about 3e-6, above the 1e-6 bar", strict=False)
def test_bf16_meets_the_1e6_bar():
    obs = synthetic.shrink(synthetic.config(3, scale=2000 / 20000), tiles_n=128, chans=16, timesteps=1)
    o32, o64 = oracle_model(obs, 0)
    _, v64, _ = factored(obs, 0, 1)
    err = rel_err(v64, o64)
    print(f"bf16 hi+lo, 2000 points, 128 tiles: rel err {err:.3e} against the 1e-6 bar")
    assert err <= FP16_TOL


@pytest.mark.parametrize("n_tiles,dead", [(20, [4]), (40, [3, 11, 25, 38]), (70, [0, 69]), (90, [1, 40, 64, 70, 89])])
def test_factored_per_tile_beams(n_tiles, dead):
    """Per-tile beams (tiles with one dead dipole each, as in the combined 256-antenna case).
    Every pair with a tile of the most common beam row takes the MFMA path: pairs inside that
    row with coherency J0 S J0^H, pairs with one other-row tile r with J_r S J0^H (or its
    conjugate transpose when the other-row tile comes second). Pairs of two other-row tiles
    take the own-beam pair kernel in FP32 with FP64 phases; they must meet the path's bar on
    their own, and differ from FP64 (it is FP32 that ran). 90 tiles put other-row tiles on
    both sides of the class-0 tiles and across a 64-tile block boundary."""
    obs = synthetic.shrink(synthetic.config(3, scale=100 / 20000), tiles_n=n_tiles, chans=9, timesteps=1)
    obs.beam = "analytic"
    obs.delays = np.zeros((n_tiles, 16), np.uint32)
    obs.gains = np.ones((n_tiles, 32))
    for k, t in enumerate(dead):
        obs.delays[t, (3 * k + 7) % 16] = 32
    o32, o64 = oracle_model(obs, 0)
    v32, v64, _ = factored(obs, 0, 0)
    err = rel_err(v64, o64)
    both = np.zeros(obs.num_baselines, bool)
    one = np.zeros(obs.num_baselines, bool)
    b = 0
    for i in range(n_tiles):
        for j in range(i + 1, n_tiles):
            both[b] = i in dead and j in dead
            one[b] = (i in dead) != (j in dead)
            b += 1
    scale = np.sqrt((o64.reshape(-1, 2) ** 2).sum(1)).max()

    def part(mask):
        if not mask.any():
            return 0.0
        d = v64[:, mask] - o64[:, mask]
        return np.sqrt((d.reshape(-1, 2) ** 2).sum(1)).max() / scale

    err_one, err_both = part(one), part(both)
    print(f"per-tile beams, {n_tiles} tiles, dead {dead}: rel err {err:.3e} "
          f"(one other-row tile {err_one:.3e}, two other-row tiles {err_both:.3e})")
    assert err <= FP16_TOL
    assert 0.0 < err_one <= FP16_TOL
    if both.any():
        assert 1e-12 < err_both <= FP16_TOL


def test_factored_c3_full_timestep_vs_fp64_direct():
    """One C3 timestep at full size (20 000 full-Stokes points, 8128 baselines, 768 channels):
    split-precision MFMA against the FP64 direct kernel (itself within 1e-12 of the oracle)."""
    obs = synthetic.config(3)
    m = SkyModeller(NoBeam(), obs.srclist, 0, obs.xyzs, obs.freqs, [], obs.phase_centre, 0.0, obs.latitude)
    shape = (len(obs.freqs), obs.num_baselines, 8)
    ref = torch.zeros(shape, dtype=torch.float64, device="cuda")
    d32 = torch.zeros(shape, dtype=torch.float32, device="cuda")
    m.model_timestep_with_device(float(obs.lsts[5]), obs.latitude, d32, ref)
    for dtype, tol, name in [(0, FP16_TOL, "fp16"), (1, BF16_MEASURED, "bf16")]:
        d64 = torch.zeros(shape, dtype=torch.float64, device="cuda")
        _lib.call("sm_model_timestep_factored", m._h, float(obs.lsts[5]), obs.latitude, dtype,
                  C.c_void_p(d32.data_ptr()), C.c_void_p(d64.data_ptr()))
        torch.cuda.synchronize()
        diff = (d64 - ref).reshape(-1, 2)
        scale = ref.reshape(-1, 2).norm(dim=1).max().item()
        err = diff.norm(dim=1).max().item() / scale
        print(f"C3 full timestep {name}: rel err vs FP64 direct {err:.3e}")
        assert err <= tol


@pytest.mark.parametrize("n_comps", [10, 23, 40])
def test_factored_kernel_organisations_agree(n_comps):
    """The three work-group organisations of the split-precision kernel (every wave generates
    and multiplies; 8 producer waves feed 4 MFMA waves; 4 producer waves feed 4 MFMA waves that
    read the K operand from pre-split planes) compute the same operands and products in the
    same order: FP64 outputs equal bit for bit, for 1, 2 and 3 steps of 16 components,
    diagonal and off-diagonal tile blocks (70 tiles), and FP32 spans of 1, 3 and 64 steps."""
    obs = synthetic.shrink(synthetic.config(3, scale=n_comps / 20000), tiles_n=70, chans=5, timesteps=1)
    try:
        for dtype in (0, 1):
            for seg in (1, 3, 64):
                _lib.call("sm_set_factored_segment", seg)
                outs = []
                for kernel in (0, 1, 2):
                    _lib.call("sm_set_factored_kernel", kernel)
                    outs.append(factored(obs, 0, dtype)[1])
                assert np.array_equal(outs[0], outs[1]), (dtype, seg)
                assert np.array_equal(outs[0], outs[2]), (dtype, seg, "planes")
                assert np.isfinite(outs[0]).all()
    finally:
        _lib.call("sm_set_factored_segment", 64)
        _lib.call("sm_set_factored_kernel", 1)


@pytest.mark.parametrize("beam", ["shared", "per-tile"])
def test_fp32_extended_components_vs_oracle(beam):
    """Gaussians and shapelets on the precision path run in FP32 with FP64 phases (tile-pair,
    row-by-column and channel-per-lane kernels). Bar: 1e-6 of max|V| against the FP64 oracle;
    the result must differ from the FP64 path's (it is the FP32 arithmetic that ran) by no
    more than that."""
    from mwa_hyperdrive_amd.srclist import Gaussian, Shapelet, SourceList
    base = synthetic.shrink(synthetic.config(5, scale=0.004), tiles_n=80, chans=37, timesteps=1)
    sky = SourceList([(k, s) for k, s in base.srclist.items()
                      if isinstance(s.components[0].comp_type, (Gaussian, Shapelet))])
    assert len(sky) > 20
    base.srclist = sky
    if beam == "shared":
        base.delays = np.zeros((80, 16), np.uint32)
        base.gains = np.ones((80, 32))
    o32, o64 = oracle_model(base, 0)
    v32, v64, m = factored(base, 0, 0)
    info = m.info()
    assert info["num_points"] == 0 and info["num_gaussians"] > 0 and info["num_shapelets"] > 0
    if beam == "per-tile":
        assert info["num_unique_tiles"] > 2
    f64 = gpu_model(base, 0)[1]
    err = rel_err(v64, o64)
    diff = rel_err(v64, f64)
    print(f"{beam} beam: FP32 Gaussians + shapelets rel err vs oracle {err:.3e}, vs FP64 path {diff:.3e}")
    assert err <= 1e-6
    assert diff > 1e-12
