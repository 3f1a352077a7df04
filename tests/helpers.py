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


def vis_simulate_case():
    """The reference's vis-simulate check (src/cli/vis_simulate/tests.rs:103-316) as fixtures:
    (SourceList, tile XYZs, channel frequencies, phase centre, longitude, latitude, GPS times,
    expected first / last cross rows' XX). The list is converted with the reference reader's
    units (src/srclist/types/components/mod.rs:99-115): degrees * pi / 180 for RA, Dec and
    position angle, arcseconds * pi / 180 / 3600 for Gaussian axes. The channels are the
    metafits centre frequency and 80 kHz below it (src/cli/vis_simulate/mod.rs:463-480)."""
    import json
    import math
    from pathlib import Path

    from mwa_hyperdrive_amd.srclist import (FluxDensity, FluxList, Gaussian, Point, Source, SourceComponent,
                                            SourceList, radec_from_degrees)

    here = Path(__file__).parent / "golden"
    g = json.loads((here / "vis_simulate_goldens.json").read_text())
    pg = json.loads((here / "precession_goldens.json").read_text())
    sl = SourceList()
    for c in g["components"]:
        ra, dec = radec_from_degrees(c["ra_deg"], c["dec_deg"])
        if c["type"] == "point":
            ct = Point()
        else:
            ct = Gaussian(math.radians(c["maj_arcsec"]) / 3600.0, math.radians(c["min_arcsec"]) / 3600.0,
                          math.radians(c["pa_deg"]))
        sl.setdefault(c["source"], Source()).components.append(
            SourceComponent(ra, dec, ct, FluxList([FluxDensity(*f) for f in c["list"]])))
    middle = g["metafits_freqcent_mhz"] * 1e6
    n = g["num_fine_channels"]
    freqs = np.array([middle - n / 2.0 * g["freq_res_hz"] + g["freq_res_hz"] * i for i in range(n)])
    pc = tuple(math.radians(x) for x in pg["phase_centre_deg"])
    return {"srclist": sl, "xyzs": np.asarray(pg["tile_xyzs"]), "freqs": freqs, "phase_centre": pc,
            "longitude": math.radians(pg["array_longitude_deg"]), "latitude": math.radians(pg["array_latitude_deg"]),
            "gps": pg["timestamps_gps"], "first": g["first_cross_xx_re_im_weight"][:2],
            "last": g["last_cross_xx_re_im_weight"][:2]}
