"""Benchmark: sky-model visibility prediction on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
                    [--no-split-precision] [--no-surrogate]
    python bench.py --surrogate-sweep   # held-out accuracy against network size and epochs

Headline workload (configuration C2): 128 tiles (8128 cross baselines), 768 channels x
40 kHz from 167 MHz, 14 timesteps of 8 s, 8000 components (5600 points, 2000 Gaussians,
400 shapelets) with power-law / curved / list fluxes, analytic beam with delays 0.
One step = the whole job: every timestep modelled through the SkyModeller device entry
(beam, coherency table, FP64 direct kernels), inputs resident in HBM, visibilities
accumulated on the device. `value` is that job's comp.vis/s.

Also reported, outside the timed steps:
  split_precision  C3 (20 000 full-Stokes points, 128 tiles, 768 ch, 14 timesteps): the FP64
                   direct kernel against the fp16 / bf16 hi+lo MFMA path, all 14 timesteps
                   each, with the error of one timestep against the FP64 path;
  surrogate        the MLP beam: a short training run, held-out Jones error, and FP16
                   inference against the exact beam on the C2 beam workload
                   (1 tile x 768 ch x 8000 directions).
Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Gcomp·vis/s, FP64 direct vs split-precision MFMA at ≤1e-6 rel err vs FP64"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
# Algorithmic FP64 work of the point tile-pair kernel per comp.vis: the 4 Jones entries'
# coherency x phasor accumulations, which with one beam on both sides take the Hermitian form
# (XX and YY real, YX = conj(XY): Re XX, Re XY, Im XY and Re YY times the complex phasor, 8
# FMAs = 16), the phasor's advance across the chunk's channels (6), the tile-phasor product at
# the chunk's first channel shared by its 4 channels (1.5), and the tile phasors themselves,
# shared by 256 baselines (under 0.5). The reference kernel's per comp.vis J S J^H (112 FLOP)
# and per-baseline sincos are not part of this algorithm: J S J^H is formed once per
# (component, channel) and the phase factorises over antennas.
FLOP_PER_COMPVIS_POINT = 24.0
# The same terms in the full complex form (4 complex multiply-adds = 32 instead of 16), as the
# shapelet kernel computes them.
FLOP_PER_COMPVIS_POINT_FULL = 40.0


def cpu_model_name() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def native_oracle():
    """The oracle built for this host's CPU (-march=native, as the CPU baseline is specified)
    into a private temporary folder; the tests use the portable build in oracle/."""
    import subprocess
    import tempfile

    from oracle import oracle as orc

    out = Path(tempfile.mkdtemp(prefix="skyoracle_native_")) / "libskyoracle.so"
    src = ROOT / "oracle" / "sky_oracle.c"
    cmd = ["gcc", "-O3", "-march=native", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-shared",
           "-o", str(out), str(src), "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        print(f"[bench] native oracle build failed, using the portable build:\n{r.stderr}", file=sys.stderr)
        return orc, "portable (-march=x86-64-v3)"
    orc._LIB_PATH = out
    orc._lib = None
    orc.lib()
    return orc, "-O3 -march=native -ffp-contract=off"


def cpu_baseline(obs, chans: int, threads: int) -> dict:
    """The oracle (the C restatement of SkyModellerCpu, same loop nest, f64) on the host's
    cores: C1 in full, and a bounded slice of C2 (the first `chans` channels of one timestep,
    all components), which is the reported value."""
    from mwa_hyperdrive_amd import synthetic

    orc, flags = native_oracle()
    c1 = synthetic.config(1)
    o1 = orc.Obs(c1.xyzs, c1.freqs, c1.phase_centre, float(c1.lsts[0]), c1.latitude)
    t0 = time.perf_counter()
    orc.model_timestep(o1, c1.srclist, num_threads=threads)
    dt1 = time.perf_counter() - t0
    work1 = c1.num_components * c1.num_baselines * len(c1.freqs)
    o = orc.Obs(obs.xyzs, obs.freqs, obs.phase_centre, float(obs.lsts[0]), obs.latitude, beam=obs.beam,
                delays=obs.delays, gains=obs.gains)
    t0 = time.perf_counter()
    orc.model_timestep_slice(o, obs.srclist, 0, chans, threads)
    dt = time.perf_counter() - t0
    work = obs.num_components * obs.num_baselines * chans
    print(f"[bench] CPU baseline: C1 in full {dt1:.2f} s, C2 slice {dt:.1f} s", file=sys.stderr, flush=True)
    return {
        "value": work / dt / 1e9,
        "unit": "Gcomp·vis/s",
        "cores": threads,
        "kind": "port",
        "sample": f"C2 timestep 0, channels 0-{chans - 1}, all {obs.num_components} components "
                  f"({work:.3e} comp.vis, {dt:.1f} s)",
        "c1_full": {"comp_vis": work1, "seconds": dt1, "gcompvis_per_s": work1 / dt1 / 1e9},
        "host": {"cpu_model": cpu_model_name(), "nproc": os.cpu_count(),
                 "affinity_cpus": len(os.sched_getaffinity(0)), "threads_used": threads, "oracle_flags": flags},
    }


def host_entry_c2(obs, timesteps: int = 2) -> dict:
    """C2 timesteps through the host-buffer entry (sm_model_timestep_with): the visibilities
    go host -> device and back (200 MB each way per timestep in f32), as the reference's
    SkyModellerGpu::model_timestep_with does (src/model/gpu.rs:1476-1496)."""
    import numpy as np

    from mwa_hyperdrive_amd.modeller import AnalyticBeam, SkyModeller

    m = SkyModeller(AnalyticBeam(obs.delays, obs.gains), obs.srclist, 0, obs.xyzs, obs.freqs, [],
                    obs.phase_centre, 0.0, obs.latitude)
    vis = np.zeros((len(obs.freqs), obs.num_baselines, 8), np.float32)
    m.model_timestep_with(float(obs.lsts[0]), obs.latitude, vis)  # warm-up (allocations)
    t0 = time.perf_counter()
    for t in range(timesteps):
        m.model_timestep_with(float(obs.lsts[t]), obs.latitude, vis)
    dt = (time.perf_counter() - t0) / timesteps
    work = obs.num_components * obs.num_baselines * len(obs.freqs)
    del m
    return {"ms_per_timestep": dt * 1e3, "gcompvis_per_s": work / dt / 1e9,
            "vis_bytes_each_way": vis.nbytes, "timesteps": timesteps}


FLOP_PER_COMPVIS_GAUSSIAN = FLOP_PER_COMPVIS_POINT + 21.0  # (kx, ky) 6, quadratic form 5, x K 1, exp 1, scale 8
KERNEL_NAMES = {
    "points": "k_direct_tiled<point> (FP64 tile-pair phasor kernel)",
    "gaussians": "k_direct_tiled<gaussian> (FP64 tile-pair phasor kernel)",
    "shapelets": "k_shapelet_chan (FP64 shapelet kernel, two channels per lane)",
}


def shapelet_flop(obs) -> float:
    """Executed FP64 work per shapelet comp.vis: the point terms, 34 for the rotation,
    positions and complex envelope, and 9 per coefficient (two table interpolations, the
    product and the complex accumulation), at the sky's mean coefficient count."""
    from mwa_hyperdrive_amd.srclist import Shapelet

    counts = [len(c.comp_type.coeffs) for src in obs.srclist.values() for c in src.components
              if isinstance(c.comp_type, Shapelet)]
    mean = sum(counts) / len(counts) if counts else 0.0
    return FLOP_PER_COMPVIS_POINT_FULL + 34.0 + 9.0 * mean


def split_precision() -> dict:
    """C3 points: FP64 direct vs the fp16 / bf16 hi+lo MFMA path over all 14 timesteps, each
    timed with HIP events around its kernels; errors of timestep 5 against the FP64 path."""
    import ctypes as C

    import torch

    from mwa_hyperdrive_amd import _lib, synthetic
    from mwa_hyperdrive_amd.modeller import NoBeam, SkyModeller

    obs = synthetic.config(3)
    m = SkyModeller(NoBeam(), obs.srclist, 0, obs.xyzs, obs.freqs, [], obs.phase_centre, 0.0, obs.latitude)
    shape = (len(obs.freqs), obs.num_baselines, 8)
    v32 = torch.zeros(shape, dtype=torch.float32, device="cuda")
    work = obs.num_components * obs.num_baselines * len(obs.freqs) * len(obs.lsts)
    res = {"workload": f"C3: {obs.num_components} full-Stokes points, {obs.num_baselines} baselines, "
                       f"{len(obs.freqs)} channels, {len(obs.lsts)} timesteps, no beam",
           "comp_vis": work}

    def run(kind, lst, d64=None):
        if kind == "fp64":
            m.model_timestep_with_device(lst, obs.latitude, v32, d64)
        else:
            _lib.call("sm_model_timestep_factored", m._h, lst, obs.latitude, kind, C.c_void_p(v32.data_ptr()),
                      C.c_void_p(d64.data_ptr()) if d64 is not None else None)

    ref = torch.zeros(shape, dtype=torch.float64, device="cuda")
    run("fp64", float(obs.lsts[5]), ref)
    torch.cuda.synchronize()
    scale = ref.reshape(-1, 2).norm(dim=1).max().item()
    for kind, name in (("fp64", "fp64_direct"), (0, "fp16_hi_lo"), (1, "bf16_hi_lo")):
        run(kind, float(obs.lsts[0]))  # warm-up (allocations)
        torch.cuda.synchronize()
        m.set_timing(True)
        t0 = time.perf_counter()
        for lst in obs.lsts:
            run(kind, float(lst))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        info = m.info()
        m.set_timing(False)
        kern_ms = info["kernel_ms_points"] if kind == "fp64" else info["kernel_ms_factored"]
        entry = {"kernel_ms": kern_ms, "wall_ms": wall * 1e3, "gcompvis_per_s_kernel": work / (kern_ms * 1e-3) / 1e9,
                 "gcompvis_per_s_wall": work / wall / 1e9}
        if kind != "fp64":
            d64 = torch.zeros(shape, dtype=torch.float64, device="cuda")
            run(kind, float(obs.lsts[5]), d64)
            torch.cuda.synchronize()
            entry["rel_err_vs_fp64"] = (d64 - ref).reshape(-1, 2).norm(dim=1).max().item() / scale
            del d64
        res[name] = entry
        print(f"[bench] split precision {name}: {entry['gcompvis_per_s_kernel']:.1f} Gcomp.vis/s (kernels)",
              file=sys.stderr, flush=True)
    # The same fp16 path with the K operand read from pre-split planes (factored kernel 2).
    try:
        _lib.call("sm_set_factored_kernel", 2)
        run(0, float(obs.lsts[0]))
        torch.cuda.synchronize()
        m.set_timing(True)
        t0 = time.perf_counter()
        for lst in obs.lsts:
            run(0, float(lst))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        info = m.info()
        m.set_timing(False)
        kern_ms = info["kernel_ms_factored"]
        d64 = torch.zeros(shape, dtype=torch.float64, device="cuda")
        run(0, float(obs.lsts[5]), d64)
        torch.cuda.synchronize()
        res["fp16_hi_lo_planes"] = {
            "kernel_ms": kern_ms, "wall_ms": wall * 1e3, "gcompvis_per_s_kernel": work / (kern_ms * 1e-3) / 1e9,
            "gcompvis_per_s_wall": work / wall / 1e9,
            "rel_err_vs_fp64": (d64 - ref).reshape(-1, 2).norm(dim=1).max().item() / scale}
        del d64
        print(f"[bench] split precision fp16 (K planes): {res['fp16_hi_lo_planes']['gcompvis_per_s_kernel']:.1f} "
              "Gcomp.vis/s (kernels)", file=sys.stderr, flush=True)
    finally:
        _lib.call("sm_set_factored_kernel", 1)
    base = res["fp64_direct"]["gcompvis_per_s_kernel"]
    for name in ("fp16_hi_lo", "bf16_hi_lo", "fp16_hi_lo_planes"):
        res[name]["speedup_vs_fp64_direct"] = res[name]["gcompvis_per_s_kernel"] / base
    # MFMA rate of the fp16 path, executed dense work: per channel and component, each 64 x 64
    # tile block computes 4 Jones entries x 64 x 128 outputs = 32 output tiles of 32 x 32
    # (diagonal blocks skip 8 of them), with 2 real K per component, 2 FLOP per MAC and 3
    # hi/lo products (hi*hi, hi*lo, lo*hi).
    g = (len(obs.xyzs) + 63) // 64
    tiles = g * 24 + (g * (g - 1) // 2) * 32
    flop = tiles * 32 * 32 * 2 * 2 * 3 * obs.num_components * len(obs.freqs) * len(obs.lsts)
    res["fp16_hi_lo"]["mfma_tflops_executed"] = flop / (res["fp16_hi_lo"]["kernel_ms"] * 1e-3) / 1e12
    res["fp16_hi_lo"]["mfma_peak_tflops"] = 2500.0
    del m, v32, ref
    torch.cuda.empty_cache()
    return res


def combined_c5(with_surrogate: bool = True) -> dict:
    """C5 (256 tiles, per-tile beams with dead dipoles, 50 000 points + 10 000 Gaussians + 500
    shapelets, 384 channels, 28 timesteps): every timestep through the FP64 direct path and
    through the precision path (fp16 points on MFMA by beam-row classes, Gaussians and shapelets
    in FP32 with FP64 phases), after one warm-up call each; the error of the latter against the
    former at one timestep."""
    import ctypes as C

    import torch

    from mwa_hyperdrive_amd import _lib, synthetic
    from mwa_hyperdrive_amd.modeller import AnalyticBeam, SkyModeller

    obs = synthetic.config(5)
    m = SkyModeller(AnalyticBeam(obs.delays, obs.gains), obs.srclist, 0, obs.xyzs, obs.freqs, [],
                    obs.phase_centre, 0.0, obs.latitude)
    shape = (len(obs.freqs), obs.num_baselines, 8)
    v32 = torch.zeros(shape, dtype=torch.float32, device="cuda")
    ref = torch.zeros(shape, dtype=torch.float64, device="cuda")
    out = torch.zeros(shape, dtype=torch.float64, device="cuda")
    lst = float(obs.lsts[3])
    nts = len(obs.lsts)
    work = obs.num_components * obs.num_baselines * len(obs.freqs)

    def fp64(d64, t_lst=None):
        m.model_timestep_with_device(lst if t_lst is None else t_lst, obs.latitude, v32, d64)

    def mixed(d64, t_lst=None):
        _lib.call("sm_model_timestep_factored", m._h, lst if t_lst is None else t_lst, obs.latitude, 0,
                  C.c_void_p(v32.data_ptr()), C.c_void_p(d64.data_ptr()) if d64 is not None else None)

    res = {"workload": f"C5: {obs.num_baselines} baselines, {len(obs.freqs)} channels, "
                       f"{obs.num_components} components, per-tile beams ({m.info()['num_unique_tiles']} beam rows), "
                       f"{nts} timesteps", "comp_vis": work * nts}
    for name, fn, buf in (("fp64_direct", fp64, ref), ("fp16_points_mfma", mixed, out)):
        fn(buf)  # warm-up; this call's output is the error reference
        torch.cuda.synchronize()
        m.set_timing(True)
        t0 = time.perf_counter()
        for t in range(nts):
            fn(None, float(obs.lsts[t]))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        info = m.info()
        m.set_timing(False)
        res[name] = {"wall_ms": wall * 1e3, "ms_per_timestep": wall * 1e3 / nts,
                     "gcompvis_per_s_wall": work * nts / wall / 1e9,
                     "kernel_ms_per_timestep": {k: info[f"kernel_ms_{k}"] / nts
                                                for k in ("points", "factored", "gaussians", "shapelets")}}
        print(f"[bench] C5 {name}: {wall * 1e3 / nts:.1f} ms per timestep", file=sys.stderr, flush=True)
    scale = ref.reshape(-1, 2).norm(dim=1).max().item()
    res["fp16_points_mfma"]["rel_err_vs_fp64"] = (out - ref).reshape(-1, 2).norm(dim=1).max().item() / scale
    res["speedup_end_to_end"] = res["fp64_direct"]["wall_ms"] / res["fp16_points_mfma"]["wall_ms"]
    if not with_surrogate:
        del m, v32, ref, out
        torch.cuda.empty_cache()
        return res
    # The same precision path with the beam from the MLP surrogate and the exact beam as per-tile
    # fallback above a 1e-3 Jones-error threshold (a short training run first).
    from mwa_hyperdrive_amd import surrogate as sg

    s = sg.BeamSurrogate(seed=1)
    tr = s.train(n_samples=1 << 21, epochs=4, batch=4096, lr=2e-3)
    m.set_beam_surrogate(s, 1e-3)
    out.zero_()
    mixed(out)
    torch.cuda.synchronize()
    m.set_beam_surrogate(s, 1e-3)
    m.set_timing(True)
    t0 = time.perf_counter()
    for t in range(nts):
        mixed(None, float(obs.lsts[t]))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    m.set_timing(False)
    fallbacks = m.set_beam_surrogate(None, 0.0)
    res["fp16_points_mfma_surrogate_beam"] = {
        "wall_ms": wall * 1e3, "ms_per_timestep": wall * 1e3 / nts, "gcompvis_per_s_wall": work * nts / wall / 1e9,
        "threshold": 1e-3, "tile_fallbacks": fallbacks,
        "tile_beams_evaluated": 3 * nts * m.info()["num_unique_tiles"], "train_seconds": tr["seconds"],
        "rel_err_vs_fp64": (out - ref).reshape(-1, 2).norm(dim=1).max().item() / scale}
    print(f"[bench] C5 surrogate beam: {wall * 1e3 / nts:.1f} ms per timestep, {fallbacks} tile fallbacks",
          file=sys.stderr, flush=True)
    del m, v32, ref, out, s
    torch.cuda.empty_cache()
    return res


def surrogate_bench(obs) -> dict:
    """Short training of the MLP beam, its held-out error, and FP16 inference against the exact
    beam for the C2 beam workload (delays 0, 768 channels, 8000 directions)."""
    import numpy as np
    import torch

    from mwa_hyperdrive_amd import surrogate as sg
    from mwa_hyperdrive_amd.modeller import analytic_beam_device

    s = sg.BeamSurrogate(seed=1)
    tr = s.train(n_samples=1 << 21, epochs=12, batch=4096, lr=2e-3)
    xh, yh = sg.samples(99, 0, 100000)
    d = s.forward(xh) - yh
    rms, mx = d.pow(2).mean().sqrt().item(), d.abs().max().item()
    rng = np.random.default_rng(5)
    n_dirs = 8000
    az = torch.tensor(rng.uniform(0, 2 * np.pi, n_dirs), device="cuda")
    za = torch.tensor(np.arccos(rng.uniform(0.3, 1.0, n_dirs)), device="cuda")
    delays = np.zeros((1, 16), np.uint32)
    gains = np.ones((1, 32))
    freqs = np.asarray(obs.freqs, np.float64)
    evals = n_dirs * freqs.size
    exact = analytic_beam_device(delays, gains, freqs, az, za)
    j, _ = s.beam(delays, gains, freqs, az, za, threshold=1e30)
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        analytic_beam_device(delays, gains, freqs, az, za)
    torch.cuda.synchronize()
    t_exact = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        s.beam(delays, gains, freqs, az, za, threshold=1e30)
    torch.cuda.synchronize()
    t_sur = (time.perf_counter() - t0) / reps
    err_c2 = (j - exact).abs().max().item()
    res = {
        "network": "53 -> 4 x 256 SiLU -> 8, FP32 training / FP16 MFMA inference",
        "train": {"samples": 1 << 21, "epochs": 12, "batch": 4096, "delays": "pointing (mode 1)", "steps": tr["steps"],
                  "seconds": tr["seconds"]},
        "heldout_jones_err": {"rms": rms, "max": mx, "samples": 100000},
        "c2_beam": {"evaluations": evals, "max_jones_err": err_c2,
                    "exact_ms": t_exact * 1e3, "surrogate_ms": t_sur * 1e3,
                    "exact_evals_per_s": evals / t_exact, "surrogate_evals_per_s": evals / t_sur,
                    "surrogate_speedup": t_exact / t_sur},
        "note": "exact = this build's analytic beam (about 6 FP64 sincos per evaluation); the FEE beam the "
                "surrogate targets is unavailable offline",
    }
    print(f"[bench] surrogate: held-out rms {rms:.2e}, exact {t_exact * 1e3:.2f} ms vs surrogate "
          f"{t_sur * 1e3:.2f} ms for {evals} evaluations", file=sys.stderr, flush=True)
    return res


def load_traffic(workload: str):
    p = ROOT / "profiles" / "traffic.json"
    if not p.exists():
        return None
    try:
        return json.loads(p.read_text()).get(workload)
    except (OSError, ValueError):
        return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-chans", type=int, default=16)
    ap.add_argument("--timesteps", type=int, default=0, help="limit timesteps per step (profiling runs only)")
    ap.add_argument("--no-split-precision", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--probe", action="store_true",
                    help="counter runs only: one plain 1 GiB device copy (the calibration kernel), one C2 "
                         "timestep (FP64) and one C3 timestep (fp16 split precision), no JSON")
    ap.add_argument("--surrogate-sweep", action="store_true",
                    help="held-out Jones error of the beam surrogate against width, depth and epochs; "
                         "one JSON line, no benchmark")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} (WORLD_SIZE {world}) is not supported. "
              "The predictor shards by frequency with no collective; see the design notes.", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from mwa_hyperdrive_amd import synthetic
    from mwa_hyperdrive_amd.modeller import AnalyticBeam, SkyModeller

    torch.cuda.set_device(0)
    if args.surrogate_sweep:
        from mwa_hyperdrive_amd import surrogate as sg

        runs = ([(256, 4, e) for e in (4, 16, 64)] + [(w, 4, 16) for w in (64, 128, 512)]
                + [(256, d, 16) for d in (2, 6)] + [(1024, 4, 8)])
        rows = sg.accuracy_sweep(runs, log=lambda t: print(f"[sweep] {t}", file=sys.stderr, flush=True))
        print(json.dumps({"surrogate_sweep": rows, "samples": 1 << 21, "batch": 4096,
                          "delays": "pointing (mode 1)", "heldout": 100000}), flush=True)
        return 0
    if args.probe:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        obs = synthetic.config(2)
        m = SkyModeller(AnalyticBeam(obs.delays, obs.gains), obs.srclist, 0, obs.xyzs, obs.freqs, [],
                        obs.phase_centre, 0.0, obs.latitude)
        vis = torch.zeros((len(obs.freqs), obs.num_baselines, 8), dtype=torch.float32, device="cuda")
        m.model_timestep_with_device(float(obs.lsts[0]), obs.latitude, vis)
        torch.cuda.synchronize()
        del m, vis
        # One C3 timestep through the fp16 split-precision MFMA path.
        import ctypes as C

        from mwa_hyperdrive_amd import _lib
        from mwa_hyperdrive_amd.modeller import NoBeam
        c3 = synthetic.config(3)
        m3 = SkyModeller(NoBeam(), c3.srclist, 0, c3.xyzs, c3.freqs, [], c3.phase_centre, 0.0, c3.latitude)
        v3 = torch.zeros((len(c3.freqs), c3.num_baselines, 8), dtype=torch.float32, device="cuda")
        _lib.call("sm_model_timestep_factored", m3._h, float(c3.lsts[0]), c3.latitude, 0, C.c_void_p(v3.data_ptr()),
                  None)
        torch.cuda.synchronize()
        print("[bench] probe done", file=sys.stderr, flush=True)
        return 0
    obs = synthetic.config(2)
    nts = args.timesteps or len(obs.lsts)
    lsts = [float(x) for x in obs.lsts[:nts]]
    m = SkyModeller(AnalyticBeam(obs.delays, obs.gains), obs.srclist, 0, obs.xyzs, obs.freqs, [],
                    obs.phase_centre, 0.0, obs.latitude)
    info = m.info()
    vis = torch.zeros((len(obs.freqs), obs.num_baselines, 8), dtype=torch.float32, device="cuda")

    def step():
        for lst in lsts:
            m.model_timestep_with_device(lst, obs.latitude, vis)

    print(f"[bench] C2: {obs.num_components} comps ({info['num_points']} points, {info['num_gaussians']} "
          f"Gaussians, {info['num_shapelets']} shapelets), {obs.num_baselines} baselines, {len(obs.freqs)} "
          f"channels, {nts} timesteps", file=sys.stderr, flush=True)
    for w in range(args.warmup):
        step()
        torch.cuda.synchronize()
        print(f"[bench] warmup {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    m.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step()
        torch.cuda.synchronize()
        print(f"[bench] step {k + 1}/{args.steps} done at {time.perf_counter() - t0:.2f} s", file=sys.stderr,
              flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    info = m.info()
    m.set_timing(False)
    work = obs.num_components * obs.num_baselines * len(obs.freqs) * nts
    value = work * args.steps / elapsed / 1e9
    launches_per_type = args.steps * nts
    nbf = obs.num_baselines * len(obs.freqs)
    per_kernel = {}
    for ty, key, flop in (("points", "num_points", FLOP_PER_COMPVIS_POINT),
                          ("gaussians", "num_gaussians", FLOP_PER_COMPVIS_GAUSSIAN),
                          ("shapelets", "num_shapelets", shapelet_flop(obs))):
        ms = info[f"kernel_ms_{ty}"] / launches_per_type
        fl = flop * info[key] * nbf
        per_kernel[ty] = {"avg_launch_ms": ms, "flop_per_comp_vis": flop,
                          "achieved_tflops": fl / (ms * 1e-3) / 1e12 if ms > 0 else None}
    dom = max(per_kernel, key=lambda k: per_kernel[k]["avg_launch_ms"])
    achieved = per_kernel[dom]["achieved_tflops"]
    traffic = load_traffic(f"C2_{dom}")
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "Gcomp·vis/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "C2: FP64 direct SkyModeller drop-in, 128 tiles, 768 ch x 40 kHz from 167 MHz, "
                        f"{nts} timesteps, 8000 components (70% points / 25% Gaussians / 5% shapelets), "
                        "analytic beam, delays 0",
            "comp_vis_per_step": work,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "fp64",
            "kernel": KERNEL_NAMES[dom],
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": traffic,
            "avg_launch_ms": per_kernel[dom]["avg_launch_ms"],
            "flop_per_comp_vis": per_kernel[dom]["flop_per_comp_vis"],
        },
        "per_kernel": per_kernel,
        "kernels_ms_per_step": {
            "points": info["kernel_ms_points"] / args.steps,
            "gaussians": info["kernel_ms_gaussians"] / args.steps,
            "shapelets": info["kernel_ms_shapelets"] / args.steps,
        },
    }
    if traffic is not None:
        out["roofline"]["hbm_gbs"] = traffic / (per_kernel[dom]["avg_launch_ms"] * 1e-3) / 1e9
        out["roofline"]["hbm_frac"] = out["roofline"]["hbm_gbs"] / PEAK_HBM_GBS
    del m, vis
    torch.cuda.empty_cache()
    out["host_entry_c2"] = host_entry_c2(obs)
    print(f"[bench] host entry: {out['host_entry_c2']['ms_per_timestep']:.1f} ms per C2 timestep",
          file=sys.stderr, flush=True)
    if not args.no_split_precision:
        out["split_precision"] = split_precision()
        out["combined_c5"] = combined_c5(with_surrogate=not args.no_surrogate)
    if not args.no_surrogate:
        out["surrogate"] = surrogate_bench(obs)
    if not args.no_cpu_baseline:
        # Every CPU this process may run on, up to the 16 a one-GPU box allots per GPU.
        threads = min(16, len(os.sched_getaffinity(0)))
        print(f"[bench] CPU baseline: oracle on {threads} threads ({cpu_model_name()}, nproc {os.cpu_count()}), "
              f"{args.cpu_chans} channels", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(obs, args.cpu_chans, threads)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
