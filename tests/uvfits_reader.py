"""A minimal reader of random-groups uvfits files for the tests: the primary header, the group
parameters and data, and the columns of binary-table extensions."""

from __future__ import annotations

import numpy as np


def _header(data: bytes, off: int):
    cards = {}
    order = []
    while True:
        block = data[off:off + 2880]
        if len(block) < 2880:
            raise ValueError("truncated header")
        off += 2880
        for i in range(36):
            card = block[i * 80:(i + 1) * 80].decode("ascii")
            key = card[:8].strip()
            if key == "END":
                return cards, order, off
            if card[8:10] == "= ":
                raw = card[10:]
                if raw.lstrip().startswith("'"):
                    body = raw.lstrip()[1:]
                    value, j = "", 0
                    while j < len(body):
                        if body[j] == "'":
                            if j + 1 < len(body) and body[j + 1] == "'":
                                value += "'"
                                j += 2
                                continue
                            break
                        value += body[j]
                        j += 1
                    cards[key] = value.rstrip()
                else:
                    cards[key] = raw.split("/")[0].strip()
                order.append(key)


def read_uvfits(path):
    """Returns (primary header, group parameters [groups][pcount] float32, data [groups][...]
    float32, {extname: (header, {column: values})})."""
    data = open(path, "rb").read()
    h0, _, off = _header(data, 0)
    naxis = int(h0["NAXIS"])
    shape = [int(h0[f"NAXIS{i}"]) for i in range(2, naxis + 1)]
    per = int(np.prod(shape))
    pcount, gcount = int(h0["PCOUNT"]), int(h0["GCOUNT"])
    n = gcount * (pcount + per)
    arr = np.frombuffer(data, ">f4", count=n, offset=off).astype(np.float32).reshape(gcount, pcount + per)
    off += (n * 4 + 2879) // 2880 * 2880
    params, vis = arr[:, :pcount], arr[:, pcount:].reshape([gcount] + shape[::-1])
    tables = {}
    sizes = {"A": 1, "J": 4, "E": 4, "D": 8, "L": 1, "I": 2}
    fmts = {"J": ">i4", "E": ">f4", "D": ">f8", "I": ">i2"}
    while off < len(data):
        h, _, off = _header(data, off)
        nrow, rowlen = int(h["NAXIS2"]), int(h["NAXIS1"])
        cols, o = {}, 0
        for i in range(1, int(h["TFIELDS"]) + 1):
            form = h[f"TFORM{i}"]
            count = int(form[:-1]) if len(form) > 1 else 1
            t = form[-1]
            cols[h[f"TTYPE{i}"]] = (o, count, t)
            o += count * sizes[t]
        assert o == rowlen, (o, rowlen)
        values = {}
        for name, (o, count, t) in cols.items():
            col = []
            for r in range(nrow):
                raw = data[off + r * rowlen + o:off + r * rowlen + o + count * sizes[t]]
                if t == "A":
                    col.append(raw.decode().rstrip("\x00 "))
                elif count > 0:
                    v = np.frombuffer(raw, fmts[t])
                    col.append(v.astype(v.dtype.newbyteorder("=")) if count > 1 else v[0].item())
                else:
                    col.append(None)
            values[name] = col
        tables[h.get("EXTNAME", "")] = (h, values)
        off += (nrow * rowlen + 2879) // 2880 * 2880
    return h0, params, vis, tables


def decode_baseline(b: float):
    """(tile1, tile2), 1-based, of an AIPS baseline number."""
    b = int(round(b))
    if b > 65535:
        b -= 65536
        return b // 2048, b % 2048
    return b // 256, b % 256
