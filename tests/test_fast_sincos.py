"""CPU check of the kernels' FP64 sin/cos (mwa_hyperdrive_amd/csrc/sm_math.h).

The header is compiled for the host with g++ and compared with glibc's sin/cos (which
Rust's f64::sin_cos also calls) over the phase range the kernels see, and at the
quadrant boundaries where a wrong reduction or sign would show."""

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
SRC = r"""
#include "sm_math.h"
extern "C" void run(int n, const double *x, double *s, double *c) {
    for (int i = 0; i < n; ++i) sm::fast_sincos(x[i], s[i], c[i]);
}
static const double kTab[128] = {SM_SINCOS32_TABLE};
extern "C" void run_tab(int n, const double *x, double *s, double *c) {
    for (int i = 0; i < n; ++i) sm::table_sincos(x[i], kTab, s[i], c[i]);
}
"""


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("sincos")
    (d / "t.cpp").write_text(SRC)
    so = d / "t.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", f"-I{ROOT / 'mwa_hyperdrive_amd' / 'csrc'}",
                    "-o", str(so), str(d / "t.cpp")], check=True)
    L = C.CDLL(str(so))
    L.run.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    L.run_tab.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    return L


@pytest.fixture(scope="module", params=["fast_sincos", "table_sincos"])
def fn(request, lib):
    """Each test runs on both forms: fast_sincos (the tile kernels' phasors) and table_sincos
    (the shapelet kernel's per-channel phases)."""
    return lib.run if request.param == "fast_sincos" else lib.run_tab


def run(f, x):
    x = np.ascontiguousarray(x, np.float64)
    s = np.empty_like(x)
    c = np.empty_like(x)
    f(len(x), x.ctypes.data, s.ctypes.data, c.ctypes.data)
    return s, c


def test_matches_libm_over_phase_range(fn):
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-2e4, 2e4, 200000), rng.uniform(-4, 4, 50000),
                        rng.uniform(-1e6, 1e6, 20000)])
    s, c = run(fn, x)
    assert np.abs(s - np.sin(x)).max() < 4e-16
    assert np.abs(c - np.cos(x)).max() < 4e-16


def test_quadrant_boundaries_and_exact_points(fn):
    k = np.arange(-4000, 4000, dtype=np.float64)
    for off in (0.0, 1e-9, -1e-9, np.pi / 4, -np.pi / 4):
        x = k * (np.pi / 2) + off
        s, c = run(fn, x)
        assert np.abs(s - np.sin(x)).max() < 4e-16
        assert np.abs(c - np.cos(x)).max() < 4e-16
    s, c = run(fn, np.array([0.0, -0.0]))
    assert s[0] == 0.0 and c[0] == 1.0 and c[1] == 1.0
    # Pythagorean identity far out.
    x = np.linspace(-3e4, 3e4, 10001)
    s, c = run(fn, x)
    assert np.abs(s * s + c * c - 1.0).max() < 5e-16
