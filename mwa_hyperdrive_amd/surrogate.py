"""MLP surrogate of the beam Jones on the device (include/skymodel.h, "beam surrogate").

The network is trained and evaluated by the library's own HIP kernels. torch only supplies
device memory here.
"""

from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib

N_IN, N_OUT = 53, 8


def _ptr(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(None)


POINTING, INDEPENDENT = 1, 0


def samples(seed: int, first: int, n: int, want_raw: bool = False, mode: int = POINTING):
    """Seeded training samples from the exact (analytic) beam, as device tensors.

    mode POINTING: tile delays that steer the beam within 60 deg of zenith (the delays MWA
    tiles use); INDEPENDENT: each dipole delay uniform in 0..31.
    Returns (x [n][53] f32, y [n][8] f32) or, with want_raw, also raw [n][51] f64."""
    import torch

    x = torch.empty((n, N_IN), dtype=torch.float32, device="cuda")
    y = torch.empty((n, N_OUT), dtype=torch.float32, device="cuda")
    raw = torch.empty((n, 51), dtype=torch.float64, device="cuda") if want_raw else None
    _lib.call("sm_surrogate_samples", C.c_uint64(seed), C.c_int64(first), n, mode, _ptr(x), _ptr(y), _ptr(raw))
    return (x, y, raw) if want_raw else (x, y)


class BeamSurrogate:
    """53 -> depth x hidden SiLU -> 8 MLP (4 x 256 by default); parameters live on the device.
    Shapes other than 4 x 256 train and evaluate in FP32."""

    def __init__(self, seed: int = 1, hidden: int = 256, depth: int = 4):
        h = C.c_void_p()
        _lib.call("sm_surrogate_new_shaped", C.c_uint64(seed), hidden, depth, C.byref(h))
        self._h = h
        n = C.c_int64()
        _lib.call("sm_surrogate_shape", h, None, None, C.byref(n))
        self.hidden, self.depth, self._nparams = hidden, depth, int(n.value)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().sm_surrogate_free(h)
            except Exception:  # interpreter shutdown: the library may already be gone
                pass
            self._h = None

    def num_params(self) -> int:
        return self._nparams

    def get_params(self) -> np.ndarray:
        p = np.empty(self.num_params(), dtype=np.float32)
        _lib.call("sm_surrogate_get_params", self._h, p.ctypes.data_as(C.c_void_p))
        return p

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.num_params()
        _lib.call("sm_surrogate_set_params", self._h, p.ctypes.data_as(C.c_void_p))

    def gradients(self, x, y):
        """(loss, flat gradients) on a device batch, without an update."""
        loss = C.c_double()
        g = np.empty(self.num_params(), dtype=np.float32)
        _lib.call("sm_surrogate_gradients", self._h, x.shape[0], _ptr(x), _ptr(y), C.byref(loss),
                  g.ctypes.data_as(C.c_void_p))
        return loss.value, g

    def train_step(self, x, y, lr: float) -> float:
        loss = C.c_double()
        _lib.call("sm_surrogate_train_step", self._h, x.shape[0], _ptr(x), _ptr(y), C.c_float(lr), C.byref(loss))
        return loss.value

    def forward(self, x, fp16: bool = True):
        """Outputs [n][8]; fp16 (the fused kernel) only for the 4 x 256 network."""
        import torch

        y = torch.empty((x.shape[0], N_OUT), dtype=torch.float32, device=x.device)
        _lib.call("sm_surrogate_forward", self._h, x.shape[0], _ptr(x), 1 if fp16 else 0, _ptr(y))
        return y

    def train(self, n_samples: int, epochs: int, batch: int = 8192, lr: float = 2e-3, seed: int = 7,
              lr_final: float = 1e-5, log=None, mode: int = POINTING) -> dict:
        """Adam over `epochs` passes of a fixed seeded sample set, cosine-decayed learning rate.

        The samples are generated once on the device (n_samples x 61 floats)."""
        import math

        import torch

        x, y = samples(seed, 0, n_samples, mode=mode)
        steps_per_epoch = max(1, n_samples // batch)
        total = steps_per_epoch * epochs
        t0 = time.perf_counter()
        step, last = 0, float("nan")
        g = torch.Generator(device="cuda")
        g.manual_seed(seed)
        for ep in range(epochs):
            perm = torch.randperm(n_samples, device="cuda", generator=g)
            for s in range(steps_per_epoch):
                idx = perm[s * batch:(s + 1) * batch]
                xb, yb = x[idx].contiguous(), y[idx].contiguous()
                frac = step / max(1, total - 1)
                cur = lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * frac))
                last = self.train_step(xb, yb, cur)
                step += 1
            if log is not None:
                log(f"epoch {ep + 1}/{epochs}: batch loss {last:.3e}")
        torch.cuda.synchronize()
        return {"steps": step, "seconds": time.perf_counter() - t0, "final_batch_loss": last}

    def beam(self, delays, gains, freqs, d_az, d_za, threshold: float):
        """Jones [tile][freq][dir] (torch float64 [..., 8]) with exact fallback; returns (jones, n_fallback)."""
        import torch

        delays = np.ascontiguousarray(delays, dtype=np.uint32).reshape(-1, 16)
        gains = np.ascontiguousarray(gains, dtype=np.float64).reshape(-1, 32)
        freqs = np.ascontiguousarray(freqs, dtype=np.float64)
        nt, nf, nd = delays.shape[0], freqs.size, d_az.shape[0]
        out = torch.empty((nt, nf, nd, 8), dtype=torch.float64, device="cuda")
        nfb = C.c_int32()
        _lib.call("sm_surrogate_beam", self._h, nt, delays.ctypes.data_as(C.c_void_p),
                  gains.ctypes.data_as(C.c_void_p), nf, freqs.ctypes.data_as(C.c_void_p), nd, _ptr(d_az),
                  _ptr(d_za), float(threshold), _ptr(out), C.byref(nfb))
        return out, nfb.value


def accuracy_sweep(runs, n_samples: int = 1 << 21, batch: int = 4096, lr: float = 2e-3, heldout: int = 100_000,
                   log=None) -> list:
    """What accuracy a network size and training length reach: for each (hidden, depth,
    epochs) in `runs`, a network trained from the same seed on the same pointing-delay samples,
    then its held-out Jones error (FP32 forward, 100 000 samples of another seed). Returns one
    dict per run: shape, parameters, epochs, training seconds, held-out RMS and max |dJ|."""
    xh, yh = samples(99, 0, heldout)
    rows = []
    for hidden, depth, epochs in runs:
        s = BeamSurrogate(seed=1, hidden=hidden, depth=depth)
        tr = s.train(n_samples=n_samples, epochs=epochs, batch=batch, lr=lr, log=log)
        d = s.forward(xh, fp16=False) - yh
        row = {"hidden": hidden, "depth": depth, "params": s.num_params(), "epochs": epochs,
               "train_seconds": tr["seconds"], "heldout_rms": d.pow(2).mean().sqrt().item(),
               "heldout_max": d.abs().max().item()}
        rows.append(row)
        if log is not None:
            log(f"{depth} x {hidden}, {epochs} epochs: rms {row['heldout_rms']:.3e}, max {row['heldout_max']:.3e} "
                f"({row['train_seconds']:.1f} s)")
        del s, d
    return rows
