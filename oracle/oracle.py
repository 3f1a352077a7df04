"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loaded by tests, `__graft_entry__.smoke()` and the CPU-baseline leg of bench.py as
the checker for the HIP product path. The product never imports this module.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

from mwa_hyperdrive_amd.srclist import SmComponent, pack_source_list

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "libskyoracle.so"
_lib = None


class OrcObs(C.Structure):
    _fields_ = [
        ("num_unflagged_tiles", C.c_int32),
        ("unflagged_xyzs", C.POINTER(C.c_double)),
        ("num_flagged_tiles", C.c_int32),
        ("flagged_tiles", C.POINTER(C.c_int32)),
        ("num_freqs", C.c_int32),
        ("freqs", C.POINTER(C.c_double)),
        ("phase_centre_ra", C.c_double),
        ("phase_centre_dec", C.c_double),
        ("lst", C.c_double),
        ("latitude", C.c_double),
        ("beam_type", C.c_int32),
        ("dipole_delays", C.POINTER(C.c_uint32)),
        ("dipole_gains", C.POINTER(C.c_double)),
        ("pols", C.c_int32),
        ("uvw_xyzs", C.POINTER(C.c_double)),
    ]


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        L = C.CDLL(str(_LIB_PATH))
        d, i32, p = C.c_double, C.c_int32, C.c_void_p
        L.orc_radec_to_lmn_rime.argtypes = [d, d, d, d, p]
        L.orc_xyzs_to_cross_uvws.argtypes = [i32, p, d, d, p]
        L.orc_hadec_to_azel.argtypes = [d, d, d, p, p]
        L.orc_estimate_at_freq.argtypes = [C.POINTER(SmComponent), d, p]
        L.orc_cpl_si_at_150mhz.argtypes = [C.POINTER(SmComponent)]
        L.orc_cpl_si_at_150mhz.restype = d
        L.orc_shapelet_basis_table.argtypes = [p]
        L.orc_analytic_beam_jones.argtypes = [d, d, d, p, p, p]
        L.orc_model_timestep.argtypes = [C.POINTER(OrcObs), p, p, i32, i32, p, p, i32]
        L.orc_model_timestep_slice.argtypes = [C.POINTER(OrcObs), p, p, i32, i32, i32, p, p, i32]
        L.orc_model_timestep_baselines.argtypes = [C.POINTER(OrcObs), p, p, i32, i32, i32, p, i32, p, p, i32]
        L.orc_model_timestep_autos.argtypes = [C.POINTER(OrcObs), p, p, i32, p, p, i32]
        L.orc_unique_tiles.argtypes = [i32, p, i32, p, p, p]
        L.orc_unique_tiles.restype = i32
        L.orc_baseline_to_tiles.argtypes = [i32, i32, p, p]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class Obs:
    """Observation parameters for the oracle (arguments of SkyModellerCpu::new)."""

    def __init__(self, xyzs, freqs, phase_centre, lst, latitude, flagged_tiles=(), beam="none",
                 delays=None, gains=None, pols=0, uvw_xyzs=None):
        self.xyzs = np.ascontiguousarray(np.asarray(xyzs, dtype=np.float64).reshape(-1, 3))
        self.freqs = np.ascontiguousarray(np.asarray(freqs, dtype=np.float64))
        self.flagged = np.ascontiguousarray(np.asarray(sorted(flagged_tiles), dtype=np.int32))
        self.phase_centre = phase_centre
        self.lst = float(lst)
        self.latitude = float(latitude)
        self.beam_type = {"none": 0, "analytic": 1}[beam]
        total = len(self.xyzs) + len(self.flagged)
        self.delays = None if delays is None else np.ascontiguousarray(np.asarray(delays, np.uint32).reshape(total, 16))
        self.gains = None if gains is None else np.ascontiguousarray(np.asarray(gains, np.float64).reshape(total, 32))
        self.pols = pols
        # Tile XYZs of the cross UVWs when they differ from xyzs (the precessed XYZs).
        self.uvw_xyzs = None if uvw_xyzs is None else np.ascontiguousarray(
            np.asarray(uvw_xyzs, np.float64).reshape(self.xyzs.shape))

    @property
    def num_baselines(self):
        n = len(self.xyzs)
        return n * (n - 1) // 2

    def c(self):
        o = OrcObs()
        o.num_unflagged_tiles = len(self.xyzs)
        o.unflagged_xyzs = self.xyzs.ctypes.data_as(C.POINTER(C.c_double))
        o.num_flagged_tiles = len(self.flagged)
        o.flagged_tiles = self.flagged.ctypes.data_as(C.POINTER(C.c_int32)) if len(self.flagged) else None
        o.num_freqs = len(self.freqs)
        o.freqs = self.freqs.ctypes.data_as(C.POINTER(C.c_double))
        o.phase_centre_ra, o.phase_centre_dec = self.phase_centre
        o.lst = self.lst
        o.latitude = self.latitude
        o.beam_type = self.beam_type
        o.dipole_delays = self.delays.ctypes.data_as(C.POINTER(C.c_uint32)) if self.delays is not None else None
        o.dipole_gains = self.gains.ctypes.data_as(C.POINTER(C.c_double)) if self.gains is not None else None
        o.pols = self.pols
        o.uvw_xyzs = self.uvw_xyzs.ctypes.data_as(C.POINTER(C.c_double)) if self.uvw_xyzs is not None else None
        return o


def model_timestep(obs: Obs, srclist, mask: int = 7, num_threads: int = 0, vis_f32=None, vis_f64=None):
    """Returns (vis_f32, vis_f64), shape [num_freqs][num_baselines][8]."""
    L = lib()
    packed = pack_source_list(srclist)
    shape = (len(obs.freqs), obs.num_baselines, 8)
    v32 = np.zeros(shape, np.float32) if vis_f32 is None else vis_f32
    v64 = np.zeros(shape, np.float64) if vis_f64 is None else vis_f64
    co = obs.c()
    nt = num_threads or min(16, os.cpu_count() or 1)
    rc = L.orc_model_timestep(C.byref(co), C.cast(packed.comps, C.c_void_p), C.cast(packed.offsets, C.c_void_p),
                              packed.num_sources, mask, _ptr(v32), _ptr(v64), nt)
    if rc != 0:
        raise RuntimeError(f"orc_model_timestep failed ({rc})")
    return v32, v64


def model_timestep_slice(obs: Obs, srclist, chan_start: int, chan_count: int, num_threads: int = 0):
    L = lib()
    packed = pack_source_list(srclist)
    shape = (chan_count, obs.num_baselines, 8)
    v32 = np.zeros(shape, np.float32)
    v64 = np.zeros(shape, np.float64)
    co = obs.c()
    nt = num_threads or min(16, os.cpu_count() or 1)
    rc = L.orc_model_timestep_slice(C.byref(co), C.cast(packed.comps, C.c_void_p), C.cast(packed.offsets, C.c_void_p),
                                    packed.num_sources, chan_start, chan_count, _ptr(v32), _ptr(v64), nt)
    if rc != 0:
        raise RuntimeError(f"orc_model_timestep_slice failed ({rc})")
    return v32, v64


def model_timestep_baselines(obs: Obs, srclist, chan_start: int, chan_count: int, baselines, num_threads: int = 0):
    """Channels [chan_start, chan_start + chan_count) of the listed baselines only:
    (vis_f32, vis_f64) of shape [chan_count][len(baselines)][8]."""
    L = lib()
    packed = pack_source_list(srclist)
    bls = np.ascontiguousarray(np.asarray(baselines, np.int32))
    shape = (chan_count, len(bls), 8)
    v32 = np.zeros(shape, np.float32)
    v64 = np.zeros(shape, np.float64)
    co = obs.c()
    nt = num_threads or min(16, os.cpu_count() or 1)
    rc = L.orc_model_timestep_baselines(C.byref(co), C.cast(packed.comps, C.c_void_p),
                                        C.cast(packed.offsets, C.c_void_p), packed.num_sources, chan_start, chan_count,
                                        _ptr(bls), len(bls), _ptr(v32), _ptr(v64), nt)
    if rc != 0:
        raise RuntimeError(f"orc_model_timestep_baselines failed ({rc})")
    return v32, v64


def model_autos(obs: Obs, srclist, num_threads: int = 1):
    L = lib()
    packed = pack_source_list(srclist)
    shape = (len(obs.freqs), len(obs.xyzs), 8)
    v32 = np.zeros(shape, np.float32)
    v64 = np.zeros(shape, np.float64)
    co = obs.c()
    rc = L.orc_model_timestep_autos(C.byref(co), C.cast(packed.comps, C.c_void_p), C.cast(packed.offsets, C.c_void_p),
                                    packed.num_sources, _ptr(v32), _ptr(v64), num_threads)
    if rc != 0:
        raise RuntimeError(f"orc_model_timestep_autos failed ({rc})")
    return v32, v64


def lmn_rime(ra, dec, pc_ra, pc_dec):
    out = np.zeros(3)
    lib().orc_radec_to_lmn_rime(ra, dec, pc_ra, pc_dec, _ptr(out))
    return out


def cross_uvws(xyzs, ha, dec):
    xyzs = np.ascontiguousarray(np.asarray(xyzs, np.float64).reshape(-1, 3))
    n = len(xyzs)
    out = np.zeros((n * (n - 1) // 2, 3))
    lib().orc_xyzs_to_cross_uvws(n, _ptr(xyzs), ha, dec, _ptr(out))
    return out


def azel(ha, dec, lat):
    az, el = C.c_double(), C.c_double()
    lib().orc_hadec_to_azel(ha, dec, lat, C.byref(az), C.byref(el))
    return az.value, el.value


def estimate_at_freq(component, freq):
    from mwa_hyperdrive_amd.srclist import Source
    packed = pack_source_list([Source([component])])
    out = np.zeros(5)
    lib().orc_estimate_at_freq(C.cast(packed.comps, C.POINTER(SmComponent)), freq, _ptr(out))
    return out  # freq, i, q, u, v


def cpl_si_at_150mhz(component):
    from mwa_hyperdrive_amd.srclist import Source
    packed = pack_source_list([Source([component])])
    return lib().orc_cpl_si_at_150mhz(C.cast(packed.comps, C.POINTER(SmComponent)))


def shapelet_basis_table():
    out = np.zeros((101, 10001))
    lib().orc_shapelet_basis_table(_ptr(out))
    return out


def analytic_beam(az, za, freq, delays=None, gains=None):
    d = None if delays is None else np.ascontiguousarray(np.asarray(delays, np.uint32))
    g = None if gains is None else np.ascontiguousarray(np.asarray(gains, np.float64))
    out = np.zeros(8)
    lib().orc_analytic_beam_jones(az, za, freq, _ptr(d), _ptr(g), _ptr(out))
    return out


def unique_tiles(total_tiles, flagged, delays, gains):
    fl = np.ascontiguousarray(np.asarray(flagged, np.int32))
    d = None if delays is None else np.ascontiguousarray(np.asarray(delays, np.uint32))
    g = None if gains is None else np.ascontiguousarray(np.asarray(gains, np.float64))
    tm = np.zeros(total_tiles, np.int32)
    n = lib().orc_unique_tiles(total_tiles, _ptr(fl) if len(fl) else None, len(fl), _ptr(d), _ptr(g), _ptr(tm))
    return n, tm
