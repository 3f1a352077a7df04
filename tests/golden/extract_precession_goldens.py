"""Writes tests/golden/precession_goldens.json: the inputs and expected outputs of the
reference's vis-simulate uvfits check (src/cli/vis_simulate/tests.rs:225-316) that pin the
epoch and precession handling.

Inputs, read as data from the reference tree:
- tile names and positions of observation 1090008640 from test_files/1090008640/1090008640.metafits
  (binary table columns Antenna, Pol, TileName, East, North, Height), positions turned into geodetic XYZ at the
  MWA latitude as XyzGeodetic's ENH conversion does: x = -n sin(lat) + h cos(lat), y = e,
  z = n cos(lat) + h sin(lat);
- the scheduled start (GPSTIME) and the phase centre (RAPHASE, DECPHASE);
- the timestamps vis-simulate forms: start + time_res / 2 + i time_res with the default 8 s
  resolution (src/cli/vis_simulate/mod.rs:390-405), two timesteps, DUT1 0 (the metafits has
  no DUT1 key).

Expected outputs, parsed from the test's text: the first and last cross-correlation group
parameters (u, v, w in seconds, baseline, date) and the first row's Julian date.

    python tests/golden/extract_precession_goldens.py
"""

from __future__ import annotations

import json
import math
import re
import struct
from pathlib import Path

REF = Path("/root/reference")
METAFITS = REF / "test_files/1090008640/1090008640.metafits"
TESTS_RS = REF / "src/cli/vis_simulate/tests.rs"
OUT = Path(__file__).resolve().parent / "precession_goldens.json"

MWA_LAT_DEG = -26.703319405555554
MWA_LONG_DEG = 116.67081523611111


def fits_header(data: bytes, off: int):
    cards = {}
    while True:
        block = data[off:off + 2880]
        off += 2880
        for i in range(36):
            card = block[i * 80:(i + 1) * 80].decode("ascii")
            key = card[:8].strip()
            if key == "END":
                return cards, off
            if card[8:10] == "= ":
                cards[key] = card[10:].split("/")[0].strip().strip("'").strip()


def tile_xyzs(data: bytes):
    h0, off = fits_header(data, 0)
    h1, off = fits_header(data, off)
    nrow, rowlen, nf = int(h1["NAXIS2"]), int(h1["NAXIS1"]), int(h1["TFIELDS"])
    sizes = {"I": 2, "E": 4, "A": 1, "J": 4, "D": 8}
    cols, o = {}, 0
    for i in range(1, nf + 1):
        form = h1[f"TFORM{i}"]
        n = int(form[:-1]) if len(form) > 1 else 1
        cols[h1[f"TTYPE{i}"]] = (o, n, form[-1])
        o += n * sizes[form[-1]]
    assert o == rowlen
    rows = []
    for r in range(nrow):
        rec = data[off + r * rowlen:off + (r + 1) * rowlen]

        def get(name):
            o, n, t = cols[name]
            b = rec[o:o + n * sizes[t]]
            if t == "A":
                return b.decode().rstrip("\x00 ").strip()
            return struct.unpack({"I": ">h", "E": ">f"}[t], b)[0]

        if get("Pol") == "X":
            rows.append((get("Antenna"), get("East"), get("North"), get("Height"), get("TileName")))
    rows.sort()
    lat = math.radians(MWA_LAT_DEG)
    s, c = math.sin(lat), math.cos(lat)
    return [[-n * s + h * c, e, n * c + h * s] for _, e, n, h, _ in rows], [r[4] for r in rows], h0


def expected_group_params(text: str):
    blocks = re.findall(r"group_params\[\.\.\],\s*\[([^\]]*)\]", text)
    arrays = [[float(x) for x in re.findall(r"-?\d+\.\d+(?:e-?\d+)?", b)] for b in blocks]
    cross = [a for a in arrays if a[0] != 0.0]
    assert len(cross) == 2, arrays
    jd = re.search(r"group_params\[4\] as f64 \+ jd_zero,\s*([0-9.]+)", text)
    return cross, float(jd.group(1))


def main():
    data = METAFITS.read_bytes()
    xyzs, names, h0 = tile_xyzs(data)
    cross, first_jd = expected_group_params(TESTS_RS.read_text())
    start = float(h0["GPSTIME"])
    out = {
        "source": "src/cli/vis_simulate/tests.rs:225-316 with test_files/1090008640/1090008640.metafits",
        "array_longitude_deg": MWA_LONG_DEG,
        "array_latitude_deg": MWA_LAT_DEG,
        "phase_centre_deg": [float(h0["RAPHASE"]), float(h0["DECPHASE"])],
        "dut1_seconds": 0.0,
        "timestamps_gps": [start + 4.0, start + 12.0],
        "tile_xyzs": xyzs,
        "tile_names": names,
        "first_cross": {"tiles": [0, 1], "timestep": 0, "uvw_seconds": cross[0][:3], "baseline": cross[0][3]},
        "last_cross": {"tiles": [len(xyzs) - 2, len(xyzs) - 1], "timestep": 1, "uvw_seconds": cross[1][:3],
                       "baseline": cross[1][3]},
        "first_row_jd": first_jd,
    }
    OUT.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {OUT}: {len(xyzs)} tiles")


if __name__ == "__main__":
    main()
