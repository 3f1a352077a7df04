"""Builds the HIP library in-tree with hipcc for gfx950.

    python -m mwa_hyperdrive_amd.build

Produces mwa_hyperdrive_amd/libskymodel_hip.so, the C-ABI library declared in
include/skymodel.h (plus the factored and surrogate entry points).
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libskymodel_hip.so"
SOURCES = ["direct.hip", "direct_tiled.hip", "abi.hip", "modeller.hip", "factored.hip", "surrogate.hip", "peel.hip",
           "precession.cpp", "uvfits.cpp"]
ARCH = os.environ.get("SM_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> Path:
    srcs = [str(CSRC / s) for s in SOURCES if (CSRC / s).exists()]
    # Objects go to a private temporary folder, so concurrent builds never share files
    # and nothing but the library lands in the source tree.
    with tempfile.TemporaryDirectory(prefix="skymodel_build_") as tmpdir:
        return _build(srcs, Path(tmpdir), verbose)


def _build(srcs, tmpdir: Path, verbose: bool) -> Path:
    objs = []
    jobs = []
    for s in srcs:
        obj = str(tmpdir / Path(s).with_suffix(".o").name)
        cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-c",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-Wno-unused-command-line-argument", f"-I{ROOT / 'include'}", "-o", obj, s]
        jobs.append((cmd, obj))
    procs = [(subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, cwd=tmpdir), obj, cmd) for cmd, obj in jobs]
    for p, obj, cmd in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            raise RuntimeError(f"hipcc failed for {cmd[-1]}:\n{text}")
        if verbose and text.strip():
            print(text)
        objs.append(obj)
    tmp = str(tmpdir / LIB.name)
    link = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(link, capture_output=True, text=True, cwd=tmpdir)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    # Copy next to the target, then rename: the library is replaced in one step.
    staged = str(LIB) + f".{os.getpid()}.tmp"
    shutil.copyfile(tmp, staged)
    os.replace(staged, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
