"""Python mirror of hyperdrive's SkyModeller trait over the HIP library.

`new_sky_modeller` / `SkyModeller.model_timestep(_with)` / `update_with_a_source` /
`model_timestep_autos_with` follow src/model/mod.rs:123-261. The timestep methods take an
Epoch, as the reference's do: the modeller forms the LST, UVWs and latitude with the DUT1 and
precession setting it was made with (get_lst_uvws_latitude, src/model/cpu.rs:731-769). They
also take (LST, latitude) directly in place of the Epoch, for callers that form those
themselves. Visibilities are [freq][baseline] Jones, stored as float32/float64 arrays of shape
(num_freqs, num_baselines, 8) = [XX re, XX im, XY re, XY im, YX re, YX im, YY re, YY im].
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import _lib
from .precession import Epoch
from .srclist import Source, SourceList, pack_source_list

POLS = {"XX XY YX YY": 0, "XX": 1, "YY": 2, "XX YY": 3, "XX YY XY": 4}


@dataclass
class NoBeam:
    """Identity Jones everywhere (src/beam/mod.rs:255-330)."""


@dataclass
class AnalyticBeam:
    """The MWA-like analytic beam: per-tile dipole delays [total tiles][16] (32 = dead)
    and gains [total tiles][32] (X then Y). None means delays 0 / gains 1."""

    delays: Optional[np.ndarray] = None
    gains: Optional[np.ndarray] = None


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _dev_ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def analytic_beam_device(delays, gains, freqs, d_az, d_za):
    """The analytic beam on the device: torch float64 [tile][freq][dir][8] (sm_analytic_beam)."""
    import torch

    delays = np.ascontiguousarray(np.asarray(delays, np.uint32).reshape(-1, 16))
    gains = np.ascontiguousarray(np.asarray(gains, np.float64).reshape(-1, 32))
    freqs = np.ascontiguousarray(np.asarray(freqs, np.float64))
    out = torch.empty((delays.shape[0], freqs.size, d_az.shape[0], 8), dtype=torch.float64, device=d_az.device)
    _lib.call("sm_analytic_beam", delays.shape[0], C.c_void_p(delays.ctypes.data), C.c_void_p(gains.ctypes.data),
              freqs.size, C.c_void_p(freqs.ctypes.data), d_az.shape[0], _dev_ptr(d_az), _dev_ptr(d_za),
              _dev_ptr(out))
    return out


class SkyModeller:
    def __init__(self, beam, source_list, pols, unflagged_tile_xyzs, unflagged_fine_chan_freqs, flagged_tiles,
                 phase_centre, array_longitude_rad, array_latitude_rad, dut1: float = 0.0,
                 apply_precession: bool = False):
        L = _lib.lib()
        self._xyzs = np.ascontiguousarray(np.asarray(unflagged_tile_xyzs, np.float64).reshape(-1, 3))
        self._freqs = np.ascontiguousarray(np.asarray(unflagged_fine_chan_freqs, np.float64))
        self._flagged = np.ascontiguousarray(np.asarray(sorted(flagged_tiles), np.int32))
        total = len(self._xyzs) + len(self._flagged)
        self.num_tiles = len(self._xyzs)
        self.num_baselines = self.num_tiles * (self.num_tiles - 1) // 2
        self.num_freqs = len(self._freqs)
        self.pols = pols
        p = _lib.SmObsParams()
        p.num_unflagged_tiles = self.num_tiles
        p.unflagged_tile_xyzs = self._xyzs.ctypes.data
        p.num_flagged_tiles = len(self._flagged)
        p.flagged_tiles = self._flagged.ctypes.data if len(self._flagged) else None
        p.num_freqs = self.num_freqs
        p.freqs = self._freqs.ctypes.data
        p.phase_centre_ra, p.phase_centre_dec = phase_centre
        p.array_longitude = array_longitude_rad
        p.array_latitude = array_latitude_rad
        p.dut1 = float(dut1)
        p.apply_precession = 1 if apply_precession else 0
        p.pols = POLS[pols] if isinstance(pols, str) else int(pols)
        self._delays = self._gains = None
        if isinstance(beam, AnalyticBeam):
            p.beam_type = 1
            if beam.delays is not None:
                self._delays = np.ascontiguousarray(np.asarray(beam.delays, np.uint32).reshape(total, 16))
                p.dipole_delays = self._delays.ctypes.data
            if beam.gains is not None:
                self._gains = np.ascontiguousarray(np.asarray(beam.gains, np.float64).reshape(total, 32))
                p.dipole_gains = self._gains.ctypes.data
        elif isinstance(beam, NoBeam) or beam is None:
            p.beam_type = 0
        else:
            raise TypeError(f"unsupported beam {beam!r}")
        self._params = p
        packed = pack_source_list(source_list)
        h = C.c_void_p()
        _lib.check(L.sm_new(C.byref(p), C.cast(packed.comps, C.c_void_p), C.cast(packed.offsets, C.c_void_p),
                            packed.num_sources, C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().sm_free(h)
            except Exception:
                pass
            self._h = None

    # ---------------------------------------------------------------- trait surface
    def _shape_check(self, vis_fb, vis_fb_f64):
        assert vis_fb.dtype == np.float32 and vis_fb.flags.c_contiguous
        assert vis_fb.shape == (self.num_freqs, self.num_baselines, 8)
        if vis_fb_f64 is not None:
            assert vis_fb_f64.dtype == np.float64 and vis_fb_f64.shape == vis_fb.shape

    def model_timestep(self, timestamp, latitude: Optional[float] = None, want_f64: bool = False):
        """Visibilities and UVWs of one timestep: `timestamp` is an Epoch, or an LST with
        `latitude`."""
        vis = np.zeros((self.num_freqs, self.num_baselines, 8), np.float32)
        vis64 = np.zeros((self.num_freqs, self.num_baselines, 8), np.float64) if want_f64 else None
        if isinstance(timestamp, Epoch):
            if latitude is not None:
                raise TypeError("an Epoch carries its own LST and latitude")
            uvws = self.model_timestep_with(timestamp, vis, vis64)
        else:
            uvws = self.model_timestep_with(timestamp, latitude, vis, vis64)
        return (vis, uvws, vis64) if want_f64 else (vis, uvws)

    def model_timestep_with(self, timestamp, *args) -> np.ndarray:
        """Accumulates into vis_fb (float32 [freq][baseline][8]); returns the baseline UVWs.
        Arguments: (epoch, vis_fb[, vis_fb_f64]) or (lst, latitude, vis_fb[, vis_fb_f64])."""
        uvws = np.zeros((self.num_baselines, 3))
        if isinstance(timestamp, Epoch):
            vis_fb, vis_fb_f64 = (tuple(args) + (None,))[:2]
            self._shape_check(vis_fb, vis_fb_f64)
            _lib.call("sm_model_timestep_epoch_with", self._h, timestamp.to_gpst_seconds(), _ptr(uvws), _ptr(vis_fb),
                      _ptr(vis_fb_f64))
        else:
            latitude, vis_fb, vis_fb_f64 = (tuple(args) + (None,))[:3]
            self._shape_check(vis_fb, vis_fb_f64)
            _lib.call("sm_model_timestep_with", self._h, timestamp, latitude, _ptr(uvws), _ptr(vis_fb),
                      _ptr(vis_fb_f64))
        return uvws

    def model_timestep_with_device(self, timestamp, *args, d_uvws=None):
        """Device-buffer form (torch tensors on the GPU); enqueues and returns. Arguments:
        (epoch, d_vis_fb[, d_vis_fb_f64]) or (lst, latitude, d_vis_fb[, d_vis_fb_f64[, d_uvws]]);
        d_uvws (LST form only) overrides the UVWs."""
        if isinstance(timestamp, Epoch):
            d_vis_fb, d_vis_fb_f64 = (tuple(args) + (None,))[:2]
            assert d_uvws is None, "the Epoch form computes its own UVWs"
            assert d_vis_fb.is_cuda and d_vis_fb.numel() == self.num_freqs * self.num_baselines * 8
            _lib.call("sm_model_timestep_epoch_with_device", self._h, timestamp.to_gpst_seconds(),
                      _dev_ptr(d_vis_fb), _dev_ptr(d_vis_fb_f64))
            return
        latitude, d_vis_fb, d_vis_fb_f64, pos_uvws = (tuple(args) + (None, None))[:4]
        if d_uvws is None:
            d_uvws = pos_uvws
        assert d_vis_fb.is_cuda and d_vis_fb.numel() == self.num_freqs * self.num_baselines * 8
        _lib.call("sm_model_timestep_with_device", self._h, timestamp, latitude, _dev_ptr(d_uvws), _dev_ptr(d_vis_fb),
                  _dev_ptr(d_vis_fb_f64))

    def lst_uvws_latitude(self, epoch: Epoch):
        """(LST, cross UVWs [baseline][3], latitude) of a timestamp (get_lst_uvws_latitude)."""
        lst, lat = C.c_double(), C.c_double()
        uvws = np.zeros((self.num_baselines, 3))
        _lib.call("sm_lst_uvws_latitude", self._h, epoch.to_gpst_seconds(), C.byref(lst), C.byref(lat), _ptr(uvws))
        return lst.value, uvws, lat.value

    def update_with_a_source(self, source: Source, phase_centre) -> None:
        packed = pack_source_list([source])
        _lib.call("sm_update_with_a_source", self._h, C.cast(packed.comps, C.c_void_p), packed.num_components,
                  phase_centre[0], phase_centre[1])
        self._keep = packed

    def model_timestep_autos_with(self, timestamp, *args):
        """Autos [freq][tile][8]: (epoch[, vis_fb[, vis_fb_f64]]) or (lst, latitude[, vis_fb[,
        vis_fb_f64]])."""
        if isinstance(timestamp, Epoch):
            vis_fb, vis_fb_f64 = (tuple(args) + (None, None))[:2]
        else:
            latitude, vis_fb, vis_fb_f64 = (tuple(args) + (None, None))[:3]
        if vis_fb is None:
            vis_fb = np.zeros((self.num_freqs, self.num_tiles, 8), np.float32)
        if isinstance(timestamp, Epoch):
            _lib.call("sm_model_timestep_autos_epoch_with", self._h, timestamp.to_gpst_seconds(), _ptr(vis_fb),
                      _ptr(vis_fb_f64))
        else:
            _lib.call("sm_model_timestep_autos_with", self._h, timestamp, latitude, _ptr(vis_fb), _ptr(vis_fb_f64))
        return vis_fb

    def set_beam_surrogate(self, surrogate, threshold: float) -> int:
        """Beam Jones from the MLP surrogate with exact per-tile fallback above `threshold`
        (None restores the analytic beam). Returns the fallbacks counted since the last call."""
        n = C.c_int32()
        _lib.call("sm_set_beam_surrogate", self._h, surrogate._h if surrogate is not None else None,
                  float(threshold), C.byref(n))
        self._surrogate = surrogate  # keep it alive while in use
        return n.value

    # ---------------------------------------------------------------- measurement
    def set_timing(self, enabled: bool) -> None:
        _lib.call("sm_set_timing", self._h, 1 if enabled else 0)

    def info(self) -> dict:
        i = _lib.SmInfo()
        _lib.call("sm_get_info", self._h, C.byref(i))
        return {k: getattr(i, k) for k, _ in _lib.SmInfo._fields_}


def new_sky_modeller(beam, source_list, pols, unflagged_tile_xyzs, unflagged_fine_chan_freqs, flagged_tiles,
                     phase_centre, array_longitude_rad, array_latitude_rad, dut1: float = 0.0,
                     apply_precession: bool = False) -> SkyModeller:
    """new_sky_modeller (src/model/mod.rs:215-261); the device is always the MI355X. dut1 in
    seconds."""
    return SkyModeller(beam, source_list, pols, unflagged_tile_xyzs, unflagged_fine_chan_freqs, flagged_tiles,
                       phase_centre, array_longitude_rad, array_latitude_rad, dut1, apply_precession)
