"""CPU oracle of the beam-surrogate MLP (test infrastructure only).

Only the tests, the smoke check and the CPU-baseline leg of bench.py use this module, and only
as the checker. The reference has no surrogate, so this restates our own definition (the C ABI
in include/skymodel.h, section "beam surrogate"):

* features: sin az, cos az, sin za, cos za, (f - 182 MHz) / 15 MHz, delays / 31 (a delay of
  32 marks a dead dipole: feature 0 and both its gains 0), 32 gains;
* network: 53 -> 256 -> 256 -> 256 -> 256 -> 8, SiLU on the hidden layers (other widths and
  depths through `shape=dims(hidden, depth)`);
* loss: mean of squared errors over the batch's n x 8 outputs;
* Adam with beta1 0.9, beta2 0.999, eps 1e-8 and bias correction.

The targets come from the analytic beam, orc_analytic_beam_jones in sky_oracle.c.

Everything is float64 numpy. The gradients are pinned by central finite differences in
tests/test_surrogate_oracle.py.
"""

from __future__ import annotations

import numpy as np

N_IN, N_HID, N_OUT = 53, 256, 8


def dims(hidden: int = N_HID, depth: int = 4):
    """(out, in) of each layer: `depth` hidden layers of `hidden` units, then the output."""
    return [(hidden, N_IN)] + [(hidden, hidden)] * (depth - 1) + [(N_OUT, hidden)]


DIMS = dims()


def num_params(shape=DIMS) -> int:
    return sum(o * i + o for o, i in shape)


def unpack(flat, shape=DIMS):
    """Flat parameters (W0, b0, W1, b1, ...) -> list of (W [out][in], b [out])."""
    flat = np.asarray(flat, dtype=np.float64)
    out, o = [], 0
    for no, ni in shape:
        w = flat[o:o + no * ni].reshape(no, ni)
        o += no * ni
        b = flat[o:o + no]
        o += no
        out.append((w, b))
    return out


def pack(layers) -> np.ndarray:
    return np.concatenate([np.concatenate([w.ravel(), b.ravel()]) for w, b in layers])


def features(raw) -> np.ndarray:
    """raw [n][51] = az, za, freq, delays[16], gains[32] -> features [n][53]."""
    raw = np.asarray(raw, dtype=np.float64)
    az, za, f = raw[:, 0], raw[:, 1], raw[:, 2]
    delays = raw[:, 3:19]
    gains = raw[:, 19:51].copy()
    dead = delays == 32
    gains[:, :16][dead] = 0.0
    gains[:, 16:][dead] = 0.0
    x = np.empty((raw.shape[0], N_IN))
    x[:, 0], x[:, 1], x[:, 2], x[:, 3] = np.sin(az), np.cos(az), np.sin(za), np.cos(za)
    x[:, 4] = (f - 182e6) / 15e6
    x[:, 5:21] = np.where(dead, 0.0, delays) / 31.0
    x[:, 21:53] = gains
    return x


def _silu(z):
    return z / (1.0 + np.exp(-z))


def forward(flat, x, shape=DIMS):
    """Returns (y [n][8], cache of (input, pre-activation) per layer)."""
    h = np.asarray(x, dtype=np.float64)
    cache = []
    layers = unpack(flat, shape)
    for li, (w, b) in enumerate(layers):
        z = h @ w.T + b
        cache.append((h, z))
        h = _silu(z) if li < len(layers) - 1 else z
    return h, cache


def loss_and_grads(flat, x, y, shape=DIMS):
    """MSE over all n x 8 outputs and its gradient (flat, same order as the parameters)."""
    out, cache = forward(flat, x, shape)
    y = np.asarray(y, dtype=np.float64)
    d = out - y
    loss = float(np.mean(d * d))
    dz = 2.0 * d / d.size
    layers = unpack(flat, shape)
    grads = [None] * len(layers)
    for li in range(len(layers) - 1, -1, -1):
        h, _ = cache[li]
        w, _ = layers[li]
        grads[li] = (dz.T @ h, dz.sum(axis=0))
        if li == 0:
            break
        dh = dz @ w
        z = cache[li - 1][1]
        s = 1.0 / (1.0 + np.exp(-z))
        dz = dh * s * (1.0 + z * (1.0 - s))
    return loss, pack(grads)


def adam_step(flat, grads, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (t counts from 1); returns new (flat, m, v)."""
    m = b1 * m + (1.0 - b1) * grads
    v = b2 * v + (1.0 - b2) * grads * grads
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    return flat - lr * mh / (np.sqrt(vh) + eps), m, v
