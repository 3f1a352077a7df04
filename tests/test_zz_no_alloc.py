"""The allocation-free entry point sm_model_points_noalloc against sm_model_points_f64 (the
entry it is a variant of): the same inputs give the same visibilities bit for bit, on two
calls in a row (the second smaller, so the shared scratch array holds the first call's
contents), and a call needing more work memory than the fixed array fails with an error."""

import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from mwa_hyperdrive_amd import _lib, synthetic
from tests.test_gpu_abi import Buckets, addresses, dt, make_sky, p

pytestmark = pytest.mark.gpu


def _case(seed, nt, nchan, n_each, per_tile):
    from oracle import oracle as orc
    rng = np.random.default_rng(seed)
    xyzs = synthetic.tiles(rng, nt, 1000.0)
    freqs = 167e6 + 40e3 * np.arange(nchan)
    pc = synthetic.PHASE_CENTRE
    lst = 0.2
    comps = make_sky(rng, n_each, freqs, "point")
    b = Buckets(comps, freqs, pc, xyzs, lst)
    nbl = nt * (nt - 1) // 2
    ndir = len(b.order)
    if per_tile:
        delays = np.zeros((nt, 16), np.uint32)
        delays[1, 3] = 32
        delays[nt - 2, 9] = 32
        gains = np.ones((nt, 32))
        n_u, tile_map = orc.unique_tiles(nt, [], delays, gains)
        az_za = [orc.azel(lst - c.ra, c.dec, synthetic.MWA_LAT_RAD) for c in b.order]
        uniq = [list(tile_map).index(u) for u in range(n_u)]
        jones = np.array([[[orc.analytic_beam(az, np.pi / 2 - el, f, delays[t], gains[t]) for az, el in az_za]
                           for f in freqs] for t in uniq])
        freq_map = np.arange(nchan, dtype=np.int32)
        n_ufreq = nchan
    else:
        jones = np.zeros((1, 1, ndir, 8))
        jones[..., 0] = jones[..., 6] = 1.0
        tile_map = np.zeros(nt, np.int32)
        freq_map = np.zeros(nchan, np.int32)
        n_ufreq = 1
    a, keep = addresses(freqs, nbl, n_ufreq, tile_map, freq_map, np.arange(nt))
    n = b.n
    s = _lib.Points(n["pl"], b.g("pl_lmn"), b.g("pl_fds"), b.g("pl_sis"), n["cpl"], b.g("cpl_lmn"), b.g("cpl_fds"),
                    b.g("cpl_sis"), b.g("cpl_qs"), n["list"], b.g("list_lmn"), b.g("list_fds"))
    d_uvw = dt(orc.cross_uvws(xyzs, lst - pc[0], pc[1]))
    return dict(s=s, a=a, keep=keep, b=b, d_uvw=d_uvw, d_beam=dt(jones), shape=(nchan, nbl, 8))


def _run(fn, case):
    v32 = torch.zeros(case["shape"], dtype=torch.float32, device="cuda")
    v64 = torch.zeros(case["shape"], dtype=torch.float64, device="cuda")
    _lib.call(fn, C.byref(case["s"]), C.byref(case["a"]), p(case["d_uvw"]), p(case["d_beam"]), p(v32), p(v64))
    torch.cuda.synchronize()
    return v32.cpu().numpy(), v64.cpu().numpy()


@pytest.mark.parametrize("per_tile", [False, True], ids=["shared-beam", "per-tile-beams"])
def test_noalloc_matches_allocating_entry(per_tile):
    big = _case(3, 48, 40, 30, per_tile)
    small = _case(4, 20, 13, 5, per_tile)
    for case in (big, small):
        want32, want64 = _run("sm_model_points_f64", case)
        got32, got64 = _run("sm_model_points_noalloc", case)
        np.testing.assert_array_equal(got64, want64)
        np.testing.assert_array_equal(got32, want32)
        assert np.abs(want64).max() > 0


def test_noalloc_rejects_calls_beyond_the_scratch_array():
    case = _case(5, 16, 4000, 120, False)  # 360 components x 4008 padded channels x 64 B = 92 MB
    v32 = torch.zeros(case["shape"], dtype=torch.float32, device="cuda")
    with pytest.raises(_lib.SkyModelError, match="more work memory"):
        _lib.call("sm_model_points_noalloc", C.byref(case["s"]), C.byref(case["a"]), p(case["d_uvw"]),
                  p(case["d_beam"]), p(v32), None)
    torch.cuda.synchronize()
    assert float(v32.abs().max()) == 0.0
