"""Benchmark: sky-model visibility prediction on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Workload (configuration C2): 128 tiles (8128 cross baselines), 768 channels x 40 kHz
from 167 MHz, 14 timesteps of 8 s, 8000 components (5600 points, 2000 Gaussians,
400 shapelets) with power-law / curved / list fluxes, analytic beam with delays 0.
One step = the whole job: every timestep modelled through the SkyModeller device entry
(beam, coherency table, FP64 direct kernels), inputs resident in HBM, visibilities
accumulated on the device. Prints one JSON line.
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
# Algorithmic FP64 work of the point direct kernel per comp.vis (sincos counted as one op
# each): phase u L + v M + w N (5) + sin, cos (2) + coherency x phasor (24) + accumulate (8).
# The J S J^H product (112 FLOP) is computed once per (component, channel) by the
# coherency kernel, not per baseline.
FLOP_PER_COMPVIS_POINT = 39.0


def cpu_model_name() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(obs, chans: int, threads: int) -> dict:
    from oracle import oracle as orc

    o = orc.Obs(obs.xyzs, obs.freqs, obs.phase_centre, float(obs.lsts[0]), obs.latitude, beam=obs.beam,
                delays=obs.delays, gains=obs.gains)
    t0 = time.perf_counter()
    orc.model_timestep_slice(o, obs.srclist, 0, chans, threads)
    dt = time.perf_counter() - t0
    work = obs.num_components * obs.num_baselines * chans
    return {
        "value": work / dt / 1e9,
        "unit": "Gcomp·vis/s",
        "cores": threads,
        "kind": "port",
        "sample": f"C2 timestep 0, channels 0-{chans - 1}, all {obs.num_components} components "
                  f"({work:.3e} comp.vis, {dt:.1f} s) on {cpu_model_name()}",
    }


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
    ms_points = info["kernel_ms_points"] / launches_per_type
    flop_launch = FLOP_PER_COMPVIS_POINT * info["num_points"] * obs.num_baselines * len(obs.freqs)
    achieved = flop_launch / (ms_points * 1e-3) / 1e12
    traffic = load_traffic("C2_points")
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
            "kernel": "k_direct<point> (FP64 direct point kernel)",
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": traffic,
            "avg_launch_ms": ms_points,
            "flop_per_comp_vis": FLOP_PER_COMPVIS_POINT,
        },
        "kernels_ms_per_step": {
            "points": info["kernel_ms_points"] / args.steps,
            "gaussians": info["kernel_ms_gaussians"] / args.steps,
            "shapelets": info["kernel_ms_shapelets"] / args.steps,
        },
    }
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        print(f"[bench] CPU baseline: oracle on {threads} threads, {args.cpu_chans} channels", file=sys.stderr,
              flush=True)
        out["cpu_baseline"] = cpu_baseline(obs, args.cpu_chans, threads)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
