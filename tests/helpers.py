"""Helpers that run one synthetic observation through the HIP product and the CPU oracle."""

import numpy as np

from mwa_hyperdrive_amd.modeller import AnalyticBeam, NoBeam, SkyModeller


def gpu_model(obs, t=0, want_f64=True, pols=0):
    beam = AnalyticBeam(obs.delays, obs.gains) if obs.beam == "analytic" else NoBeam()
    m = SkyModeller(beam, obs.srclist, pols, obs.xyzs, obs.freqs, [], obs.phase_centre, 0.0, obs.latitude)
    vis, uvws, vis64 = m.model_timestep(float(obs.lsts[t]), obs.latitude, want_f64=True)
    return vis, vis64, uvws, m


def oracle_obs(obs, t=0, pols=0):
    from oracle import oracle as orc
    return orc.Obs(obs.xyzs, obs.freqs, obs.phase_centre, float(obs.lsts[t]), obs.latitude, beam=obs.beam,
                   delays=obs.delays, gains=obs.gains, pols=pols)


def oracle_model(obs, t=0, mask=7, pols=0):
    from oracle import oracle as orc
    return orc.model_timestep(oracle_obs(obs, t, pols), obs.srclist, mask=mask)


def rel_err(a, b):
    """max |a - b| over complex entries, relative to max |b| (the parity metric)."""
    a = np.asarray(a, np.float64).reshape(-1, 2)
    b = np.asarray(b, np.float64).reshape(-1, 2)
    scale = np.sqrt((b ** 2).sum(1)).max()
    diff = np.sqrt(((a - b) ** 2).sum(1)).max()
    return diff / scale if scale > 0 else diff


def f32_ulps(a, b):
    """Largest difference of two float32 arrays in units of float32 spacing at b."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    sp = np.spacing(np.abs(b)).astype(np.float64)
    sp = np.maximum(sp, np.spacing(np.float32(1e-30)))
    return float((np.abs(a.astype(np.float64) - b.astype(np.float64)) / sp).max())
