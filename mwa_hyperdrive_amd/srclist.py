"""Sky-model source-list types and their packing for the C ABI.

Mirrors hyperdrive's `SourceList` / `Source` / `SourceComponent` / `FluxDensityType`
(src/srclist/types/source_list/mod.rs:26-172, source/mod.rs:17-31,
components/mod.rs:21-128, flux_density/mod.rs:20-145). Source-list parsing stays
outside this package; these classes are what a parser would produce.

`pack_source_list` lays the list out as the flat `SmComponent` array declared in
include/skymodel.h (components in source-list order plus per-source offsets); the
library itself reverses the source order, as `SkyModellerCpu::new` /
`SkyModellerGpu::new` do (src/model/cpu.rs:86-93, src/model/gpu.rs:230-241).
"""

from __future__ import annotations

import ctypes as C
import math
from collections import OrderedDict
from dataclasses import dataclass, field
from typing import List, Sequence, Tuple, Union

import numpy as np

POINT, GAUSSIAN, SHAPELET = 0, 1, 2
LIST, POWER_LAW, CURVED_POWER_LAW = 0, 1, 2


@dataclass
class FluxDensity:
    """Stokes flux densities at one frequency (flux_density/mod.rs:20-41)."""

    freq: float
    i: float
    q: float = 0.0
    u: float = 0.0
    v: float = 0.0


@dataclass
class PowerLaw:
    si: float
    fd: FluxDensity


@dataclass
class CurvedPowerLaw:
    si: float
    fd: FluxDensity
    q: float


@dataclass
class FluxList:
    fds: List[FluxDensity]


FluxDensityType = Union[PowerLaw, CurvedPowerLaw, FluxList]


@dataclass
class Point:
    pass


@dataclass
class Gaussian:
    maj: float  # radians
    min: float
    pa: float


@dataclass
class Shapelet:
    maj: float
    min: float
    pa: float
    coeffs: List[Tuple[int, int, float]]  # (n1, n2, value)


ComponentType = Union[Point, Gaussian, Shapelet]


@dataclass
class SourceComponent:
    ra: float  # radians
    dec: float
    comp_type: ComponentType
    flux_type: FluxDensityType


@dataclass
class Source:
    components: List[SourceComponent] = field(default_factory=list)


class SourceList(OrderedDict):
    """Ordered mapping of source name -> Source (an IndexMap in the reference)."""


def radec_from_degrees(ra_deg: float, dec_deg: float) -> Tuple[float, float]:
    """RADec::from_degrees: degrees * (PI / 180) as Rust's f64::to_radians computes it."""
    k = math.pi / 180.0
    return ra_deg * k, dec_deg * k


# ----------------------------------------------------------------------------- C layout
class SmFluxDensity(C.Structure):
    _fields_ = [("freq", C.c_double), ("i", C.c_double), ("q", C.c_double), ("u", C.c_double), ("v", C.c_double)]


class ShapeletCoeff(C.Structure):
    """Same layout as hyperdrive's FFI ShapeletCoeff (src/gpu/types.h:105-109)."""

    _fields_ = [("value", C.c_double), ("n1", C.c_uint8), ("n2", C.c_uint8)]


class SmComponent(C.Structure):
    _fields_ = [
        ("ra", C.c_double),
        ("dec", C.c_double),
        ("comp_type", C.c_int32),
        ("flux_type", C.c_int32),
        ("maj", C.c_double),
        ("min", C.c_double),
        ("pa", C.c_double),
        ("si", C.c_double),
        ("q", C.c_double),
        ("fds", C.POINTER(SmFluxDensity)),
        ("num_fds", C.c_int32),
        ("num_coeffs", C.c_int32),
        ("coeffs", C.POINTER(ShapeletCoeff)),
    ]


assert C.sizeof(SmComponent) == 88
assert C.sizeof(ShapeletCoeff) == 16


class PackedSourceList:
    """Flat C arrays for a source list; keeps every buffer alive while referenced."""

    def __init__(self, comps, offsets, keep):
        self.comps = comps
        self.offsets = offsets
        self._keep = keep

    @property
    def num_sources(self) -> int:
        return len(self.offsets) - 1

    @property
    def num_components(self) -> int:
        return int(self.offsets[-1])


def iter_components(srclist: Union[SourceList, Sequence[Source]]):
    sources = list(srclist.values()) if isinstance(srclist, dict) else list(srclist)
    for src in sources:
        yield src, src.components


def pack_source_list(srclist: Union[SourceList, Sequence[Source]]) -> PackedSourceList:
    sources = list(srclist.values()) if isinstance(srclist, dict) else list(srclist)
    comps = [c for s in sources for c in s.components]
    arr = (SmComponent * max(len(comps), 1))()
    offsets = (C.c_int32 * (len(sources) + 1))()
    keep = []
    k = 0
    for si_, s in enumerate(sources):
        offsets[si_] = k
        for c in s.components:
            sc = arr[k]
            sc.ra, sc.dec = c.ra, c.dec
            ct = c.comp_type
            if isinstance(ct, Point):
                sc.comp_type = POINT
            elif isinstance(ct, Gaussian):
                sc.comp_type = GAUSSIAN
                sc.maj, sc.min, sc.pa = ct.maj, ct.min, ct.pa
            elif isinstance(ct, Shapelet):
                sc.comp_type = SHAPELET
                sc.maj, sc.min, sc.pa = ct.maj, ct.min, ct.pa
                co = (ShapeletCoeff * max(len(ct.coeffs), 1))()
                for j, (n1, n2, val) in enumerate(ct.coeffs):
                    co[j].n1, co[j].n2, co[j].value = int(n1), int(n2), float(val)
                keep.append(co)
                sc.coeffs = C.cast(co, C.POINTER(ShapeletCoeff))
                sc.num_coeffs = len(ct.coeffs)
            else:
                raise TypeError(f"unknown component type {ct!r}")
            ft = c.flux_type
            if isinstance(ft, FluxList):
                fl = ft.fds
                sc.flux_type = LIST
            elif isinstance(ft, CurvedPowerLaw):
                fl = [ft.fd]
                sc.flux_type = CURVED_POWER_LAW
                sc.si, sc.q = ft.si, ft.q
            elif isinstance(ft, PowerLaw):
                fl = [ft.fd]
                sc.flux_type = POWER_LAW
                sc.si = ft.si
            else:
                raise TypeError(f"unknown flux type {ft!r}")
            if len(fl) == 0:
                raise ValueError("a flux-density list needs at least one entry")
            fa = (SmFluxDensity * len(fl))()
            for j, fd in enumerate(fl):
                fa[j].freq, fa[j].i, fa[j].q, fa[j].u, fa[j].v = fd.freq, fd.i, fd.q, fd.u, fd.v
            keep.append(fa)
            sc.fds = C.cast(fa, C.POINTER(SmFluxDensity))
            sc.num_fds = len(fl)
            k += 1
    offsets[len(sources)] = k
    return PackedSourceList(arr, offsets, keep)


def flat_components(srclist) -> List[SourceComponent]:
    sources = list(srclist.values()) if isinstance(srclist, dict) else list(srclist)
    return [c for s in sources for c in s.components]


def enh_to_xyz(e: np.ndarray, n: np.ndarray, h: np.ndarray, latitude: float) -> np.ndarray:
    """East-North-Height to geodetic XYZ (mdbook/src/defs/coords.md "East North Height")."""
    s, c = math.sin(latitude), math.cos(latitude)
    return np.stack([-n * s + h * c, e, n * c + h * s], axis=-1)
