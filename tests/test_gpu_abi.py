"""The kernel-level drop-ins (model_points / model_gaussians / model_shapelets / model_autos)
called exactly as hyperdrive's SkyModellerGpu calls them (src/model/gpu.rs:913-945,
1000-1150, 1172-1351): reference struct layouts, device pointers, beam Jones
[unique tile][unique freq][direction], shapelet UVs [baseline][component]."""

import ctypes as C
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from mwa_hyperdrive_amd import _lib
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
)
from mwa_hyperdrive_amd import synthetic
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

dev = "cuda"


@pytest.fixture(params=[0, 1], ids=["fast-kernels", "baseline-per-lane"])
def variant(request):
    _lib.call("sm_set_direct_variant", request.param)
    yield request.param
    _lib.call("sm_set_direct_variant", 0)


def dt(a, dtype=torch.float64):
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device=dev)


def p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def make_sky(rng, n_each, freqs, kind):
    """Components whose power laws are referenced at 150 MHz and whose lists hold an entry at
    every channel, so the struct inputs (FD at 150 MHz, list FDs) are exact."""
    comps = []
    ra, dec = synthetic.cap_positions(rng, 3 * n_each, 8.0)
    for k in range(3 * n_each):
        i150 = float(rng.lognormal())
        si = float(rng.normal(-0.8, 0.2))
        q = 0.05 * i150 * rng.normal(size=3)
        fd = FluxDensity(150e6, i150, *q.tolist())
        if k % 3 == 0:
            flux = PowerLaw(si, fd)
        elif k % 3 == 1:
            flux = CurvedPowerLaw(si, fd, float(rng.normal(0, 0.2)))
        else:
            flux = FluxList([FluxDensity(f, i150 * (f / 150e6) ** si) for f in freqs])
        if kind == "point":
            ct = Point()
        elif kind == "gaussian":
            ct = Gaussian(float(rng.uniform(1e-4, 1e-3)), float(rng.uniform(1e-4, 5e-4)), float(rng.uniform(0, 3)))
        else:
            ct = Shapelet(float(rng.uniform(1e-4, 1e-3)), float(rng.uniform(1e-4, 5e-4)), float(rng.uniform(0, 3)),
                          [(int(rng.integers(0, 6)), int(rng.integers(0, 6)), float(rng.normal())) for _ in range(5)])
        comps.append(SourceComponent(float(ra[k]), float(dec[k]), ct, flux))
    return comps


def inst(fd):
    return [fd.i - fd.q, 0.0, fd.u, -fd.v, fd.u, fd.v, fd.i + fd.q, 0.0]


def lmn_rime(ra, dec, pc):
    d = ra - pc[0]
    l = math.cos(dec) * math.sin(d)
    m = math.sin(dec) * math.cos(pc[1]) - math.cos(dec) * math.sin(pc[1]) * math.cos(d)
    n = math.sin(dec) * math.sin(pc[1]) + math.cos(dec) * math.cos(pc[1]) * math.cos(d)
    return [2 * math.pi * l, 2 * math.pi * m, 2 * math.pi * (n - 1.0)]


class Buckets:
    """The nine-bucket split of SkyModellerGpu::new for one component type."""

    def __init__(self, comps, freqs, pc, xyzs, lst):
        from oracle import oracle as orc
        self.keep = []
        self.groups = {}
        for name, cls in (("pl", PowerLaw), ("cpl", CurvedPowerLaw), ("list", FluxList)):
            g = [c for c in comps if type(c.flux_type) is cls]
            self.groups[name] = g
        self.order = self.groups["pl"] + self.groups["cpl"] + self.groups["list"]
        self.n = {k: len(v) for k, v in self.groups.items()}
        self.t = {}
        for name, g in self.groups.items():
            if not g:
                continue
            self.t[name + "_lmn"] = dt([lmn_rime(c.ra, c.dec, pc) for c in g])
            if name != "list":
                self.t[name + "_fds"] = dt([inst(c.flux_type.fd) for c in g])
                self.t[name + "_sis"] = dt([c.flux_type.si for c in g])
            if name == "cpl":
                self.t["cpl_qs"] = dt([c.flux_type.q for c in g])
            if name == "list":
                self.t["list_fds"] = dt([[inst(c.flux_type.fds[f]) for c in g] for f in range(len(freqs))])
            if not isinstance(g[0].comp_type, Point):
                self.t[name + "_gps"] = dt([[c.comp_type.maj, c.comp_type.min, c.comp_type.pa] for c in g])
            if isinstance(g[0].comp_type, Shapelet):
                uvs = np.stack([orc.cross_uvws(xyzs, lst - c.ra, c.dec)[:, :2] for c in g], axis=1)
                self.t[name + "_uvs"] = dt(uvs)
                co = np.zeros(sum(len(c.comp_type.coeffs) for c in g), dtype=[("v", "f8"), ("n1", "u1"),
                                                                             ("n2", "u1"), ("pad", "u1", 6)])
                k = 0
                for c in g:
                    for n1, n2, v in c.comp_type.coeffs:
                        co[k] = (v, n1, n2, np.zeros(6, np.uint8))
                        k += 1
                self.t[name + "_coeffs"] = torch.tensor(co.view(np.uint8), device=dev)
                self.t[name + "_ncoeffs"] = dt([len(c.comp_type.coeffs) for c in g], torch.int32)

    def g(self, key):
        return p(self.t.get(key))


def addresses(freqs, nbl, n_ufreq, tile_map, freq_map, unflagged, sbf=None):
    a = _lib.Addresses()
    a.num_freqs = len(freqs)
    a.num_vis = len(freqs) * nbl
    a.num_baselines = nbl
    keep = [dt(freqs), dt(tile_map, torch.int32), dt(freq_map, torch.int32), dt(unflagged, torch.int32)]
    a.d_freqs, a.d_tile_map, a.d_freq_map = keep[0].data_ptr(), keep[1].data_ptr(), keep[2].data_ptr()
    a.d_tile_index_to_unflagged_tile_index_map = keep[3].data_ptr()
    a.num_unique_beam_freqs = n_ufreq
    if sbf is not None:
        keep.append(sbf)
        a.d_shapelet_basis_values = sbf.data_ptr()
    return a, keep


@pytest.mark.parametrize("kind", ["point", "gaussian", "shapelet"])
@pytest.mark.parametrize("beam", ["identity", "per_tile"])
def test_dropin_vs_oracle(kind, beam, variant):
    _dropin_case(kind, beam, variant, 167e6 + 40e3 * np.arange(19))


@pytest.mark.parametrize("kind", ["point", "gaussian", "shapelet"])
def test_dropin_gapped_channels_vs_oracle(kind, variant):
    """Two runs of 8 evenly spaced channels with a gap of two channels between them, then 3
    more: every fixed chunk of 8 (and 4) is evenly spaced, but a chunk of 6 channels straddles
    the gap, so the drop-ins must give every channel its own phase."""
    freqs = 167e6 + 40e3 * np.concatenate([np.arange(8), 10 + np.arange(8), 20 + np.arange(3)])
    _dropin_case(kind, "identity", variant, freqs)


def _dropin_case(kind, beam, variant, freqs):
    from oracle import oracle as orc
    rng = np.random.default_rng(11)
    nt = 12
    xyzs = synthetic.tiles(rng, nt, 800.0)
    pc = synthetic.PHASE_CENTRE
    lst = 0.3
    comps = make_sky(rng, 4, freqs, kind)
    b = Buckets(comps, freqs, pc, xyzs, lst)
    nbl = nt * (nt - 1) // 2
    ndir = len(b.order)
    delays = np.zeros((nt, 16), np.uint32)
    if beam == "per_tile":
        delays[3, 5] = 32
        delays[7, 0] = 32
    gains = np.ones((nt, 32))
    n_u, tile_map = orc.unique_tiles(nt, [], delays, gains)
    if beam == "identity":
        jones = np.zeros((1, 1, ndir, 8))
        jones[..., 0] = jones[..., 6] = 1.0
        tile_map = np.zeros(nt, np.int32)
        freq_map = np.zeros(len(freqs), np.int32)
        n_ufreq = 1
    else:
        az_za = [orc.azel(lst - c.ra, c.dec, synthetic.MWA_LAT_RAD) for c in b.order]
        uniq = [list(tile_map).index(u) for u in range(n_u)]
        jones = np.array([[[orc.analytic_beam(az, math.pi / 2 - el, f, delays[t], gains[t]) for az, el in az_za]
                           for f in freqs] for t in uniq])
        freq_map = np.arange(len(freqs), dtype=np.int32)
        n_ufreq = len(freqs)
    d_beam = dt(jones)
    uvws = orc.cross_uvws(xyzs, lst - pc[0], pc[1])
    d_uvw = dt(uvws)
    sbf = dt(orc.shapelet_basis_table()) if kind == "shapelet" else None
    a, keep = addresses(freqs, nbl, n_ufreq, tile_map, freq_map, np.arange(nt), sbf)
    vis = torch.zeros((len(freqs), nbl, 8), dtype=torch.float32, device=dev)
    vis64 = torch.zeros((len(freqs), nbl, 8), dtype=torch.float64, device=dev)
    n = b.n
    if kind == "point":
        s = _lib.Points(n["pl"], b.g("pl_lmn"), b.g("pl_fds"), b.g("pl_sis"), n["cpl"], b.g("cpl_lmn"),
                        b.g("cpl_fds"), b.g("cpl_sis"), b.g("cpl_qs"), n["list"], b.g("list_lmn"), b.g("list_fds"))
        fn = "sm_model_points_f64"
    elif kind == "gaussian":
        s = _lib.Gaussians(n["pl"], b.g("pl_lmn"), b.g("pl_fds"), b.g("pl_sis"), b.g("pl_gps"), n["cpl"],
                           b.g("cpl_lmn"), b.g("cpl_fds"), b.g("cpl_sis"), b.g("cpl_qs"), b.g("cpl_gps"), n["list"],
                           b.g("list_lmn"), b.g("list_fds"), b.g("list_gps"))
        fn = "sm_model_gaussians_f64"
    else:
        s = _lib.Shapelets(n["pl"], b.g("pl_lmn"), b.g("pl_fds"), b.g("pl_sis"), b.g("pl_gps"), b.g("pl_uvs"),
                           b.g("pl_coeffs"), b.g("pl_ncoeffs"), n["cpl"], b.g("cpl_lmn"), b.g("cpl_fds"),
                           b.g("cpl_sis"), b.g("cpl_qs"), b.g("cpl_gps"), b.g("cpl_uvs"), b.g("cpl_coeffs"),
                           b.g("cpl_ncoeffs"), n["list"], b.g("list_lmn"), b.g("list_fds"), b.g("list_gps"),
                           b.g("list_uvs"), b.g("list_coeffs"), b.g("list_ncoeffs"))
        fn = "sm_model_shapelets_f64"
    _lib.call(fn, C.byref(s), C.byref(a), p(d_uvw), p(d_beam), p(vis), p(vis64))
    # The f32-only drop-in accumulates the same values into a second buffer.
    vis2 = torch.zeros_like(vis)
    _lib.call({"point": "model_points", "gaussian": "model_gaussians", "shapelet": "model_shapelets"}[kind],
              C.byref(s), C.byref(a), p(d_uvw), p(d_beam), p(vis2))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(vis2.cpu().numpy(), vis.cpu().numpy())
    obs = orc.Obs(xyzs, freqs, pc, lst, synthetic.MWA_LAT_RAD, beam="none" if beam == "identity" else "analytic",
                  delays=delays if beam != "identity" else None, gains=gains if beam != "identity" else None)
    sl = SourceList([("s", Source(b.order))])
    # The oracle reverses the source order; a single source keeps the bucket order.
    o32, o64 = orc.model_timestep(obs, sl, mask={"point": 1, "gaussian": 2, "shapelet": 4}[kind])
    err = rel_err(vis64.cpu().numpy(), o64)
    print(f"variant={variant} {kind} {beam}: rel err {err:.3e}")
    assert err <= 1e-12


def test_dropin_autos_vs_oracle():
    from oracle import oracle as orc
    rng = np.random.default_rng(5)
    nt = 9
    xyzs = synthetic.tiles(rng, nt, 500.0)
    freqs = 170e6 + 40e3 * np.arange(5)
    pc = synthetic.PHASE_CENTRE
    comps = make_sky(rng, 2, freqs, "point")
    b = Buckets(comps, freqs, pc, xyzs, 0.0)
    ndir = len(b.order)
    jones = np.zeros((1, 1, ndir, 8))
    jones[..., 0] = jones[..., 6] = 1.0
    d_beam = dt(jones)
    n = b.n
    pts = _lib.Points(n["pl"], b.g("pl_lmn"), b.g("pl_fds"), b.g("pl_sis"), n["cpl"], b.g("cpl_lmn"), b.g("cpl_fds"),
                      b.g("cpl_sis"), b.g("cpl_qs"), n["list"], b.g("list_lmn"), b.g("list_fds"))
    gau = _lib.Gaussians(0, *([None] * 4), 0, *([None] * 5), 0, *([None] * 3))
    sha = _lib.Shapelets(0, *([None] * 7), 0, *([None] * 8), 0, *([None] * 6))
    d_freqs = dt(freqs)
    tm = dt(np.zeros(nt), torch.int32)
    fm = dt(np.zeros(len(freqs)), torch.int32)
    um = dt(np.arange(nt), torch.int32)
    vis = torch.zeros((len(freqs), nt, 8), dtype=torch.float32, device=dev)
    _lib.call("model_autos", nt, len(freqs), p(d_freqs), C.byref(pts), C.byref(gau), C.byref(sha), p(d_beam), p(tm),
              p(fm), 1, p(um), p(vis))
    torch.cuda.synchronize()
    obs = orc.Obs(xyzs, freqs, pc, 0.0, synthetic.MWA_LAT_RAD)
    o32, o64 = orc.model_autos(obs, SourceList([("s", Source(b.order))]))
    assert rel_err(vis.cpu().numpy(), o64) < 1e-7


def test_dropins_from_two_threads():
    """The kernel-level drop-ins keep their work buffers per calling thread: two threads that
    model different skies into their own buffers at the same time get the results each gets
    alone, bit for bit (the reference's SkyModellerGpu owns its buffers per modeller,
    src/model/gpu.rs:150-519)."""
    import threading

    rng = np.random.default_rng(19)
    nt = 24
    xyzs = synthetic.tiles(rng, nt, 900.0)
    freqs = 167e6 + 40e3 * np.arange(33)
    pc = synthetic.PHASE_CENTRE
    lst = 0.1
    nbl = nt * (nt - 1) // 2
    jobs = []
    for seed in (1, 2):
        comps = make_sky(np.random.default_rng(100 + seed), 20 * seed, freqs, "point")
        b = Buckets(comps, freqs, pc, xyzs, lst)
        n = b.n
        s = _lib.Points(n["pl"], b.g("pl_lmn"), b.g("pl_fds"), b.g("pl_sis"), n["cpl"], b.g("cpl_lmn"),
                        b.g("cpl_fds"), b.g("cpl_sis"), b.g("cpl_qs"), n["list"], b.g("list_lmn"), b.g("list_fds"))
        jones = np.zeros((1, 1, len(b.order), 8))
        jones[..., 0] = jones[..., 6] = 1.0
        from oracle import oracle as orc
        d_uvw = dt(orc.cross_uvws(xyzs, lst - pc[0], pc[1]))
        a, keep = addresses(freqs, nbl, 1, np.zeros(nt), np.zeros(len(freqs)), np.arange(nt))
        jobs.append(dict(b=b, s=s, a=a, keep=keep, d_beam=dt(jones), d_uvw=d_uvw))

    def run(job, out):
        out64 = torch.zeros((len(freqs), nbl, 8), dtype=torch.float64, device=dev)
        out32 = torch.zeros((len(freqs), nbl, 8), dtype=torch.float32, device=dev)
        for _ in range(3):
            _lib.call("sm_model_points_f64", C.byref(job["s"]), C.byref(job["a"]), p(job["d_uvw"]), p(job["d_beam"]),
                      p(out32), p(out64))
        out.append((out32, out64))

    alone = []
    for job in jobs:
        run(job, alone)
    torch.cuda.synchronize()
    together = [[], []]
    threads = [threading.Thread(target=run, args=(job, res)) for job, res in zip(jobs, together)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    torch.cuda.synchronize()
    for (a32, a64), res in zip(alone, together):
        (t32, t64), = res
        np.testing.assert_array_equal(t64.cpu().numpy(), a64.cpu().numpy())
        np.testing.assert_array_equal(t32.cpu().numpy(), a32.cpu().numpy())
