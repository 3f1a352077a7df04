"""Timestamps for the SkyModeller: an Epoch (GPS seconds) and marlu::precession's
`get_lmst`, `precess_time` and `PrecessionInfo::precess_xyz`, which the reference calls in
get_lst_uvws_latitude (src/model/cpu.rs:731-769). They run in the library (host code, no GPU).
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib


@dataclass(frozen=True)
class Epoch:
    """A timestamp on the GPS scale (hifitime::Epoch's role at the modeller's boundary)."""

    gps_seconds: float

    @classmethod
    def from_gpst_seconds(cls, s: float) -> "Epoch":
        return cls(float(s))

    def to_gpst_seconds(self) -> float:
        return self.gps_seconds


@dataclass
class PrecessionInfo:
    """marlu::precession::PrecessionInfo: rotation_matrix takes true-of-date vectors to
    J2000; hadec_j2000 is the aberrated phase centre at lmst_j2000."""

    rotation_matrix: np.ndarray
    hadec_j2000: tuple
    lmst: float
    lmst_j2000: float
    array_latitude_j2000: float
    _c: _lib.SmPrecessionInfo

    def precess_xyz(self, xyzs) -> np.ndarray:
        """Tile XYZs [n][3] rotated into the J2000 frame (PrecessionInfo::precess_xyz)."""
        x = np.ascontiguousarray(np.asarray(xyzs, np.float64).reshape(-1, 3))
        out = np.empty_like(x)
        _lib.call("sm_precess_xyz", C.byref(self._c), C.c_void_p(x.ctypes.data), C.c_void_p(out.ctypes.data),
                  len(x))
        return out


def _gps(epoch) -> float:
    return epoch.to_gpst_seconds() if isinstance(epoch, Epoch) else float(epoch)


def get_lmst(array_longitude_rad: float, epoch, dut1_seconds: float = 0.0) -> float:
    """Local mean sidereal time [rad] (marlu::precession::get_lmst)."""
    return float(_lib.lib().sm_get_lmst(array_longitude_rad, _gps(epoch), dut1_seconds))


def precess_time(array_longitude_rad: float, array_latitude_rad: float, phase_centre, epoch,
                 dut1_seconds: float = 0.0) -> PrecessionInfo:
    """marlu::precession::precess_time (src/model/cpu.rs:733-739)."""
    c = _lib.SmPrecessionInfo()
    _lib.call("sm_precess_time", array_longitude_rad, array_latitude_rad, phase_centre[0], phase_centre[1],
              _gps(epoch), dut1_seconds, C.byref(c))
    rot = np.array([[c.rotation_matrix[i][j] for j in range(3)] for i in range(3)])
    return PrecessionInfo(rot, (c.hadec_j2000.ha, c.hadec_j2000.dec), c.lmst, c.lmst_j2000, c.array_latitude_j2000, c)
