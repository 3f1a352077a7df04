"""CPU-only checks of the C-ABI library: it loads without a GPU, exports every function
declared in include/*.h, and its struct layouts match the reference's FFI sizes."""

import ctypes as C
import re
from pathlib import Path

import pytest

from mwa_hyperdrive_amd import _lib

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = []
    for h in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, flags=re.M):
            name = m.group(1)
            if name not in ("sizeof",):
                names.append(name)
    return sorted(set(names))


def test_library_loads_and_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_functions()
    assert "model_points" in names and "sm_new" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_struct_sizes_match_reference_ffi():
    # src/gpu/types.h with FLOAT = double: ShapeletCoeff 16 B, Jones 64 B, LmnRime 24 B.
    from mwa_hyperdrive_amd.srclist import ShapeletCoeff, SmComponent
    assert C.sizeof(ShapeletCoeff) == 16
    assert C.sizeof(_lib.LmnRime) == 24
    assert C.sizeof(_lib.UVW) == 24
    assert C.sizeof(_lib.GaussianParams) == 24
    assert C.sizeof(_lib.Addresses) == 64
    assert C.sizeof(_lib.Points) == 96
    assert C.sizeof(SmComponent) == 88


def test_no_cpu_fallback_when_library_missing(tmp_path, monkeypatch):
    monkeypatch.setattr(_lib, "LIB_PATH", tmp_path / "absent.so")
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(_lib.SkyModelError):
        _lib.lib()
