"""Writes tests/golden/vis_simulate_goldens.json: the sky and the expected visibilities of the
reference's vis-simulate uvfits check (src/cli/vis_simulate/tests.rs:103-316, no beam, two
fine channels, two timesteps, cross-correlations only).

Inputs, read as data from the reference tree:
- the 100-source list test_files/1090008640/srclist_pumav3_EoR0aegean_EoR1pietro+ForA_1090008640_100.yaml
  (the test's --source-list), kept in file order, with the reader's units: RA/Dec in degrees,
  Gaussian axes in arcseconds, position angles in degrees
  (src/srclist/types/components/mod.rs:60-115);
- the metafits centre frequency (FREQCENT) and coarse channels (CHANNELS) of
  test_files/1090008640/1090008640.metafits.

Expected outputs, parsed from the test's text: the first three floats of the first and last
cross-correlation rows (XX re, XX im, weight of the first channel). The tile positions and
timestamps are those of precession_goldens.json.

    python tests/golden/extract_vis_simulate_goldens.py
"""

from __future__ import annotations

import json
import re
from pathlib import Path

import yaml

REF = Path("/root/reference")
SRCLIST = REF / "test_files/1090008640/srclist_pumav3_EoR0aegean_EoR1pietro+ForA_1090008640_100.yaml"
METAFITS = REF / "test_files/1090008640/1090008640.metafits"
TESTS_RS = REF / "src/cli/vis_simulate/tests.rs"
OUT = Path(__file__).resolve().parent / "vis_simulate_goldens.json"


def components(doc):
    out = []
    for name, comps in doc.items():
        for c in comps:
            e = {"source": name, "ra_deg": float(c["ra"]), "dec_deg": float(c["dec"])}
            ct = c["comp_type"]
            if ct == "point":
                e["type"] = "point"
            elif isinstance(ct, dict) and "gaussian" in ct:
                g = ct["gaussian"]
                e.update(type="gaussian", maj_arcsec=float(g["maj"]), min_arcsec=float(g["min"]),
                         pa_deg=float(g["pa"]))
            else:
                raise ValueError(f"component type {ct!r} is not in this list")
            ft = c["flux_type"]
            if "list" not in ft:
                raise ValueError(f"flux type {list(ft)} is not in this list")
            e["list"] = [[float(f["freq"]), float(f["i"]), float(f.get("q", 0.0)), float(f.get("u", 0.0)),
                          float(f.get("v", 0.0))] for f in ft["list"]]
            out.append(e)
    return out


def metafits_keys(keys):
    raw = METAFITS.read_bytes()[:2880 * 20]
    out = {}
    for k in keys:
        i = raw.find(k.ljust(8).encode())
        card = raw[i:i + 80].decode()
        out[k] = card[10:].split("/")[0].strip().strip("'")
    return out


def expected_vis(text: str):
    blocks = re.findall(r"assert_abs_diff_eq!\(vis\[\.\.3\], \[([^\]]*)\]", text)
    vals = [[float(x) for x in b.split(",")] for b in blocks]
    cross = [v for v in vals if v[1] != 0.0]  # autos have zero imaginary parts
    assert len(cross) == 2, vals
    return cross


def main():
    doc = yaml.safe_load(SRCLIST.read_text())
    comps = components(doc)
    mk = metafits_keys(["FREQCENT", "CHANNELS", "FINECHAN"])
    first, last = expected_vis(TESTS_RS.read_text())
    out = {
        "source": "src/cli/vis_simulate/tests.rs:103-316 with the 100-source list and 1090008640.metafits",
        "num_fine_channels": 2,
        "freq_res_hz": 80e3,
        "metafits_freqcent_mhz": float(mk["FREQCENT"]),
        "metafits_channels": mk["CHANNELS"],
        "components": comps,
        "first_cross_xx_re_im_weight": first,
        "last_cross_xx_re_im_weight": last,
        "epsilon": 1e-3,
    }
    OUT.write_text(json.dumps(out, indent=0) + "\n")
    print(f"wrote {OUT}: {len(comps)} components, first {first}, last {last}")


if __name__ == "__main__":
    main()
