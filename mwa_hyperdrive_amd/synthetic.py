"""Deterministic synthetic observations and skies for the benchmark configurations.

One generator (numpy PCG64, seed = 1000 + config index) serves the benchmark, the
parity tests and the CPU baseline, so all of them model the same inputs:

  C1  128 tiles (E/N uniform in a 1.5 km disc, H ~ N(0, 1 m)); 32 x 40 kHz from 180 MHz;
      1 timestep at LST 0; 100 Stokes-I power-law points within 10 deg; no beam.
  C2  128 tiles; 768 x 40 kHz from 167 MHz; 14 timesteps of 8 s; 8000 components within
      20 deg (70% points, 25% Gaussians, 5% shapelets); fluxes 60% power law, 30% curved
      power law (q ~ N(0, 0.2)), 10% five-entry lists; analytic beam, delays 0.
  C3  128 tiles; 768 ch; 14 timesteps; 20 000 full-Stokes points within 25 deg.
  C5  256 tiles in a 3 km disc; 384 x 80 kHz; 28 timesteps; 50 000 points, 10 000
      Gaussians, 500 shapelets; analytic beam with one dead dipole on 10% of tiles.

S_I(150 MHz) ~ lognormal(0, 1) Jy and spectral index ~ N(-0.8, 0.2) throughout.
Phase centre (RA 0, Dec -27 deg); LST_k = k * 8 s * 2 pi / 86164.0905 s.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .srclist import (
    CurvedPowerLaw,
    FluxDensity,
    FluxList,
    Gaussian,
    Point,
    PowerLaw,
    Shapelet,
    Source,
    SourceComponent,
    SourceList,
    enh_to_xyz,
    radec_from_degrees,
)

MWA_LAT_RAD = -0.4660608448386394
MWA_LONG_RAD = 2.0362898668561042
SIDEREAL_DAY_S = 86164.0905
PHASE_CENTRE = radec_from_degrees(0.0, -27.0)


@dataclass
class SyntheticObs:
    name: str
    xyzs: np.ndarray                   # [tiles][3] geodetic metres
    freqs: np.ndarray                  # Hz
    lsts: np.ndarray                   # radians, one per timestep
    phase_centre: tuple
    latitude: float
    srclist: SourceList
    beam: str                          # "none" | "analytic"
    delays: Optional[np.ndarray] = None  # [tiles][16]
    gains: Optional[np.ndarray] = None   # [tiles][32]
    counts: dict = field(default_factory=dict)

    @property
    def num_tiles(self):
        return len(self.xyzs)

    @property
    def num_baselines(self):
        return self.num_tiles * (self.num_tiles - 1) // 2

    @property
    def num_components(self):
        return sum(len(s.components) for s in self.srclist.values())

    def comp_vis(self, timesteps: Optional[int] = None, chans: Optional[int] = None) -> int:
        t = len(self.lsts) if timesteps is None else timesteps
        c = len(self.freqs) if chans is None else chans
        return self.num_components * self.num_baselines * c * t


def tiles(rng, n: int, radius_m: float) -> np.ndarray:
    r = radius_m * np.sqrt(rng.uniform(size=n))
    th = 2 * np.pi * rng.uniform(size=n)
    e, nn = r * np.cos(th), r * np.sin(th)
    h = rng.normal(0.0, 1.0, size=n)
    return np.ascontiguousarray(enh_to_xyz(e, nn, h, MWA_LAT_RAD))


def cap_positions(rng, n: int, radius_deg: float, centre=PHASE_CENTRE):
    """Uniform on the spherical cap of the given radius about the phase centre."""
    cr = math.cos(math.radians(radius_deg))
    rho = np.arccos(rng.uniform(cr, 1.0, size=n))
    th = rng.uniform(0.0, 2 * np.pi, size=n)
    ra0, dec0 = centre
    sd = np.sin(dec0) * np.cos(rho) + np.cos(dec0) * np.sin(rho) * np.cos(th)
    dec = np.arcsin(np.clip(sd, -1, 1))
    ra = ra0 + np.arctan2(np.sin(th) * np.sin(rho) * np.cos(dec0), np.cos(rho) - np.sin(dec0) * sd)
    return np.mod(ra, 2 * np.pi), dec


def _flux(rng, kind: str, full_stokes: bool = False):
    i150 = float(rng.lognormal(0.0, 1.0))
    si = float(rng.normal(-0.8, 0.2))
    q = u = v = 0.0
    if full_stokes:
        q, u, v = (0.05 * i150 * rng.normal(size=3)).tolist()
    if kind == "pl":
        return PowerLaw(si, FluxDensity(150e6, i150, q, u, v))
    if kind == "cpl":
        return CurvedPowerLaw(si, FluxDensity(150e6, i150, q, u, v), float(rng.normal(0.0, 0.2)))
    fds = []
    for f in (160e6, 170e6, 180e6, 190e6, 200e6):
        r = (f / 150e6) ** si
        fds.append(FluxDensity(f, i150 * r, q * r, u * r, v * r))
    return FluxList(fds)


def _sky(rng, n_point, n_gauss, n_shape, radius_deg, flux_mix=(0.6, 0.3, 0.1), full_stokes=False) -> SourceList:
    n = n_point + n_gauss + n_shape
    ra, dec = cap_positions(rng, n, radius_deg)
    types = np.array([0] * n_point + [1] * n_gauss + [2] * n_shape)
    rng.shuffle(types)
    kinds = rng.choice(["pl", "cpl", "list"], size=n, p=list(flux_mix))
    arcmin = math.radians(1.0 / 60.0)
    sl = SourceList()
    for k in range(n):
        flux = _flux(rng, str(kinds[k]), full_stokes)
        if types[k] == 0:
            ct = Point()
        else:
            a, b = sorted(rng.uniform(0.5 * arcmin, 5 * arcmin, size=2), reverse=True)
            pa = float(rng.uniform(0.0, np.pi))
            if types[k] == 1:
                ct = Gaussian(float(a), float(b), pa)
            else:
                nc = int(rng.integers(10, 31))
                coeffs = [(int(rng.integers(0, 11)), int(rng.integers(0, 11)), float(rng.normal())) for _ in range(nc)]
                ct = Shapelet(float(a), float(b), pa, coeffs)
        sl[f"s{k:06d}"] = Source([SourceComponent(float(ra[k]), float(dec[k]), ct, flux)])
    return sl


def lsts(n: int, step_s: float = 8.0) -> np.ndarray:
    return np.array([k * step_s * 2 * math.pi / SIDEREAL_DAY_S for k in range(n)])


def config(index: int, scale: float = 1.0) -> SyntheticObs:
    """Configuration C{index} (1, 2, 3 or 5). `scale` < 1 shrinks the component count
    (used only by tests that need an oracle-sized sky with the same distributions)."""
    rng = np.random.default_rng(1000 + index)
    if index == 1:
        xyz = tiles(rng, 128, 1500.0)
        sky = _sky(rng, max(1, int(100 * scale)), 0, 0, 10.0, flux_mix=(1.0, 0.0, 0.0))
        return SyntheticObs("C1", xyz, 180e6 + 40e3 * np.arange(32), lsts(1), PHASE_CENTRE, MWA_LAT_RAD, sky, "none",
                            counts={"points": len(sky)})
    if index == 2:
        xyz = tiles(rng, 128, 1500.0)
        n = max(20, int(8000 * scale))
        npnt, ngau = int(round(0.70 * n)), int(round(0.25 * n))
        nsh = n - npnt - ngau
        sky = _sky(rng, npnt, ngau, nsh, 20.0)
        return SyntheticObs("C2", xyz, 167e6 + 40e3 * np.arange(768), lsts(14), PHASE_CENTRE, MWA_LAT_RAD, sky,
                            "analytic", np.zeros((128, 16), np.uint32), np.ones((128, 32)),
                            counts={"points": npnt, "gaussians": ngau, "shapelets": nsh})
    if index == 3:
        xyz = tiles(rng, 128, 1500.0)
        n = max(1, int(20000 * scale))
        sky = _sky(rng, n, 0, 0, 25.0, full_stokes=True)
        return SyntheticObs("C3", xyz, 167e6 + 40e3 * np.arange(768), lsts(14), PHASE_CENTRE, MWA_LAT_RAD, sky,
                            "none", counts={"points": n})
    if index == 5:
        xyz = tiles(rng, 256, 3000.0)
        npnt, ngau, nsh = (max(1, int(x * scale)) for x in (50000, 10000, 500))
        sky = _sky(rng, npnt, ngau, nsh, 20.0)
        delays = np.zeros((256, 16), np.uint32)
        gains = np.ones((256, 32))
        dead = rng.choice(256, size=26, replace=False)
        for t in dead:
            delays[t, int(rng.integers(0, 16))] = 32
        return SyntheticObs("C5", xyz, 167e6 + 80e3 * np.arange(384), lsts(28), PHASE_CENTRE, MWA_LAT_RAD, sky,
                            "analytic", delays, gains, counts={"points": npnt, "gaussians": ngau, "shapelets": nsh})
    raise ValueError(f"no configuration C{index}")


def shrink(obs: SyntheticObs, tiles_n: Optional[int] = None, chans: Optional[int] = None,
           timesteps: Optional[int] = None, comps: Optional[int] = None) -> SyntheticObs:
    """A smaller slice of a configuration (first tiles / channels / timesteps / sources)."""
    sl = obs.srclist
    if comps is not None:
        sl = SourceList(list(obs.srclist.items())[:comps])
    nt = obs.num_tiles if tiles_n is None else tiles_n
    return SyntheticObs(obs.name + "-slice", obs.xyzs[:nt].copy(),
                        obs.freqs[: chans or len(obs.freqs)].copy(), obs.lsts[: timesteps or len(obs.lsts)].copy(),
                        obs.phase_centre, obs.latitude, sl, obs.beam,
                        None if obs.delays is None else obs.delays[:nt].copy(),
                        None if obs.gains is None else obs.gains[:nt].copy(), dict(obs.counts))
