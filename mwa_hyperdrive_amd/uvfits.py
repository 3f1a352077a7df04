"""uvfits output of modelled visibilities (the library's sm_uvfits_* entry points): the writer
the reference's vis-simulate feeds from its model thread (src/io/write/mod.rs:113-260,
src/params/vis_simulate.rs:241-300). The file layout is described in include/skymodel.h."""

from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _lib
from .precession import Epoch


def cross_baseline_tiles(num_tiles: int, autos: bool = False) -> np.ndarray:
    """(tile1, tile2) of each row, upper triangle in the reference's order (src/math/mod.rs:96-109);
    with autos, each tile's auto-correlation comes before its cross rows."""
    rows = [(i, j) for i in range(num_tiles) for j in range(i if autos else i + 1, num_tiles)]
    return np.ascontiguousarray(np.asarray(rows, np.int32).reshape(-1, 2))


class UvfitsWriter:
    """Writes num_timesteps x len(baseline_tiles) groups, then the AN table on finish()."""

    def __init__(self, path: str, num_timesteps: int, baseline_tiles, freqs: Sequence[float], phase_centre,
                 array_position, tile_names: Sequence[str], tile_xyzs, int_time_s: float, first_timestamp,
                 dut1: float = 0.0, apply_precession: bool = True, obs_name: Optional[str] = None):
        freqs = np.asarray(freqs, np.float64)
        self._pairs = np.ascontiguousarray(np.asarray(baseline_tiles, np.int32).reshape(-1, 2))
        self._xyz = np.ascontiguousarray(np.asarray(tile_xyzs, np.float64).reshape(-1, 3))
        self._names = [n.encode() for n in tile_names]
        self._cnames = (C.c_char_p * len(self._names))(*self._names)
        self.num_chans, self.num_baselines = int(freqs.size), int(self._pairs.shape[0])
        p = _lib.SmUvfitsParams()
        p.num_timesteps = num_timesteps
        p.num_baselines = self.num_baselines
        p.num_chans = self.num_chans
        p.num_tiles = len(self._names)
        p.tile_names = C.cast(self._cnames, C.c_void_p)
        p.tile_xyzs = self._xyz.ctypes.data
        p.baseline_tiles = self._pairs.ctypes.data
        p.first_freq_hz = float(freqs[0])
        p.freq_res_hz = float(freqs[1] - freqs[0]) if freqs.size > 1 else 0.0
        p.phase_centre_ra, p.phase_centre_dec = phase_centre
        p.array_longitude, p.array_latitude, p.array_height = array_position
        p.dut1 = float(dut1)
        p.apply_precession = 1 if apply_precession else 0
        p.int_time_s = float(int_time_s)
        first = first_timestamp.to_gpst_seconds() if isinstance(first_timestamp, Epoch) else float(first_timestamp)
        p.first_gps_seconds = first
        self._obs = obs_name.encode() if obs_name is not None else None
        p.obs_name = self._obs
        h = C.c_void_p()
        _lib.call("sm_uvfits_create", str(path).encode(), C.byref(p), C.byref(h))
        self._h = h

    def write_timestep(self, timestamp, vis_fb: np.ndarray, weights_fb: Optional[np.ndarray] = None) -> None:
        """vis_fb float32 [num_chans][num_baselines][8] (Jones XX, XY, YX, YY as re, im);
        weights_fb float32 [num_chans][num_baselines] or None (all 1)."""
        vis = np.ascontiguousarray(vis_fb, np.float32)
        assert vis.shape == (self.num_chans, self.num_baselines, 8)
        w = None if weights_fb is None else np.ascontiguousarray(weights_fb, np.float32)
        if w is not None:
            assert w.shape == (self.num_chans, self.num_baselines)
        gps = timestamp.to_gpst_seconds() if isinstance(timestamp, Epoch) else float(timestamp)
        _lib.call("sm_uvfits_write_timestep", self._h, gps, C.c_void_p(vis.ctypes.data),
                  C.c_void_p(w.ctypes.data) if w is not None else None)

    def finish(self) -> None:
        h, self._h = self._h, None
        if h is not None:
            _lib.call("sm_uvfits_finish", h)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.finish()
        return False

    def __del__(self):
        if getattr(self, "_h", None) is not None:
            try:
                self.finish()
            except Exception:
                pass
