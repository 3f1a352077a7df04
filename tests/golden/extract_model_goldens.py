"""Write the reference's sky-model golden vectors as a JSON fixture.

Run in a container where the reference source tree is mounted (it reads the
test file as text; nothing of the reference is imported or executed):

    python tests/golden/extract_model_goldens.py /root/reference

The expected-value checker functions of src/model/tests/mod.rs:350-1600 hold
3 x 3 arrays of f32 Jones matrices (3 frequencies x 3 baselines, or 3 tiles for
the autos) plus the shapelet UVW goldens. This script pulls those numbers out
and writes tests/golden/model_goldens.json, together with the observation and
source-list parameters of the fixture builder (mod.rs:125-348), which the tests
restate.
"""

import json
import re
import sys
from pathlib import Path

NUM = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"
COMPLEX = rf"Complex::new\(\s*({NUM})\s*,\s*({NUM})\s*\)"
JONES_FROM = rf"Jones::from\(\[\s*{COMPLEX}\s*,\s*{COMPLEX}\s*,\s*{COMPLEX}\s*,\s*{COMPLEX}\s*,?\s*\]\)"
JONES_IDENT = rf"Jones::identity\(\)(?:\s*\*\s*({NUM}))?"
FROM_ELEM = rf"Array1::from_elem\(\s*vis\.dim\(\)\.0\s*,\s*{JONES_IDENT}\s*\)"
UVW = rf"UVW\s*\{{\s*u:\s*({NUM})\s*,\s*v:\s*({NUM})\s*,\s*w:\s*({NUM})\s*,?\s*\}}"


def split_functions(text):
    out = {}
    for m in re.finditer(r"\nfn (test_\w+)\(", text):
        start = m.start()
        depth = 0
        i = text.index("{", start)
        j = i
        while True:
            if text[j] == "{":
                depth += 1
            elif text[j] == "}":
                depth -= 1
                if depth == 0:
                    break
            j += 1
        out[m.group(1)] = text[i : j + 1]
    return out


def jones_sequence(body):
    token = re.compile(f"(?P<elem>{FROM_ELEM})|(?P<from>{JONES_FROM})|(?P<ident>{JONES_IDENT})")
    seq = []
    for m in token.finditer(body):
        if m.group("elem"):
            scale = re.search(JONES_IDENT, m.group("elem")).group(1)
            s = float(scale) if scale else 1.0
            seq.extend([[s, 0.0, 0.0, 0.0, 0.0, 0.0, s, 0.0]] * 3)
        elif m.group("from"):
            seq.append([float(x) for x in re.match(JONES_FROM, m.group("from")).groups()])
        else:
            scale = re.match(JONES_IDENT, m.group("ident")).group(1)
            s = float(scale) if scale else 1.0
            seq.append([s, 0.0, 0.0, 0.0, 0.0, 0.0, s, 0.0])
    return seq


def main():
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    text = (ref / "src/model/tests/mod.rs").read_text()
    funcs = split_functions(text)
    goldens = {}
    for name, body in funcs.items():
        seq = jones_sequence(body)
        if name == "test_multiple_shapelet_components":
            seq = seq[:9]
            uvws = [[float(a), float(b), float(c)] for a, b, c in re.findall(UVW, body)]
            goldens["shapelet_uvws_multiple"] = uvws
        if name in ("test_model_timestep_autos_with_gaussian", "test_model_timestep_autos_with_shapelet"):
            continue
        if len(seq) != 9:
            raise SystemExit(f"{name}: expected 9 Jones, found {len(seq)}")
        goldens[name] = [seq[0:3], seq[3:6], seq[6:9]]
    fixture = {
        "source": "src/model/tests/mod.rs:350-1600 (mwa_hyperdrive 0.8.0)",
        "layout": "[freq][baseline or tile][XX_re, XX_im, XY_re, XY_im, YX_re, YX_im, YY_re, YY_im] (f32 values)",
        "obs": {
            "phase_centre_deg": [0.0, -27.0],
            "off_phase_centre_deg": [1.0, -27.0],
            "freqs_hz": [150e6, 175e6, 200e6],
            "lst": 0.0,
            "xyzs": [[0.0, 1.0, 0.0], [0.0, -1.0, 0.0], [1.0, 0.0, 0.0]],
            "latitude_rad": -0.4660608448386394,
        },
        "fluxes": {
            "list": [[150e6, 1.0], [175e6, 3.0], [200e6, 2.0]],
            "power_law": {"si": -0.8, "freq": 150e6, "i": 1.0},
            "curved_power_law": {"si": -0.8, "freq": 150e6, "i": 1.0, "q": 0.03},
            "gaussian2": {"maj": 1.0, "min": 0.5, "pa": 0.25},
        },
        "goldens": goldens,
    }
    out = Path(__file__).with_name("model_goldens.json")
    out.write_text(json.dumps(fixture, indent=1))
    print(f"wrote {out} with {len(goldens)} golden sets")


if __name__ == "__main__":
    main()
