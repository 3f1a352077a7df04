"""CPU restatement of the peel-loop neighbours (TEST INFRASTRUCTURE ONLY: the checker for the
HIP entry points xyzs_to_uvws / rotate / add_model; nothing in the product imports it).

Each function follows the reference kernel it names, in numpy float64 / float32 arithmetic:
- xyzs_to_uvws: src/gpu/peel.cu:17-49 (baseline XYZ difference, then the rotation to the
  pointing centre's HA/Dec). Pinned by tests/test_oracle_peel.py: against the modeller's
  xyzs_to_cross_uvws form (tile UVWs, then differences; the reference's own check of the two,
  src/params/peel/tests.rs:2255-2378) and the closed form at HA = 0, Dec = 90 deg.
- rotate: src/gpu/peel.cu:56-103 (FP64 product, rounded to FP32). Pinned by the identity at
  w_to = w_from and the round trip to and back.
- add_model: src/gpu/peel.cu:559-592 (factor formed in FP64, rounded to FP32; FP32 complex
  product and sum). Pinned by the reference's KAT (src/params/peel/tests.rs:2524-2600:
  |residual| = |model| gain) and the alpha = beta = 0 case.
"""

import numpy as np

TAU = 2 * np.pi


def baseline_tiles(n):
    """Upper-triangle baseline order (src/math/mod.rs:96-109)."""
    return np.array([(i, j) for i in range(n) for j in range(i + 1, n)], dtype=np.int64).reshape(-1, 2)


def xyzs_to_uvws(xyzs_t, lmsts, pc_ra, pc_dec):
    """xyzs_t [time][tile][3] -> uvws [time][baseline][3]."""
    xyzs_t = np.asarray(xyzs_t, np.float64)
    ntime, nt, _ = xyzs_t.shape
    tp = baseline_tiles(nt)
    d = xyzs_t[:, tp[:, 0]] - xyzs_t[:, tp[:, 1]]
    s_dec, c_dec = np.sin(pc_dec), np.cos(pc_dec)
    out = np.empty_like(d)
    for t in range(ntime):
        ha = lmsts[t] - pc_ra
        s_ha, c_ha = np.sin(ha), np.cos(ha)
        x, y, z = d[t, :, 0], d[t, :, 1], d[t, :, 2]
        out[t, :, 0] = s_ha * x + c_ha * y
        out[t, :, 1] = -s_dec * c_ha * x + s_dec * s_ha * y + c_dec * z
        out[t, :, 2] = c_dec * c_ha * x - c_dec * s_ha * y + s_dec * z
    return out


def rotate(vis_tfb, uvws_from, uvws_to, lambdas):
    """vis [time][freq][baseline][8] float32 -> rotated copy."""
    vis = np.asarray(vis_tfb, np.float32)
    arg = -TAU * (uvws_to[..., 2] - uvws_from[..., 2])             # [time][baseline]
    ph = arg[:, None, :] / np.asarray(lambdas)[None, :, None]       # [time][freq][baseline]
    c, s = np.cos(ph), np.sin(ph)
    re = vis[..., 0::2].astype(np.float64)
    im = vis[..., 1::2].astype(np.float64)
    out = np.empty_like(vis)
    out[..., 0::2] = (re * c[..., None] - im * s[..., None]).astype(np.float32)
    out[..., 1::2] = (re * s[..., None] + im * c[..., None]).astype(np.float32)
    return out


def add_model(resid_tfb, model_tfb, alpha, beta, gain, lambdas, uvws):
    """resid + model * gain * exp(-i tau (u alpha + v beta) lambda), float32 result."""
    resid = np.asarray(resid_tfb, np.float32)
    model = np.asarray(model_tfb, np.float32)
    nt, nf, nb, _ = model.shape
    if alpha == 0.0 and beta == 0.0:
        c = np.ones((nt, nf, nb))
        s = np.zeros((nt, nf, nb))
    else:
        arg = -TAU * (uvws[..., 0] * alpha + uvws[..., 1] * beta)
        ph = arg[:, None, :] * np.asarray(lambdas)[None, :, None]
        c, s = np.cos(ph), np.sin(ph)
    cr = (c * gain).astype(np.float32)[..., None]
    ci = (s * gain).astype(np.float32)[..., None]
    mr, mi = model[..., 0::2], model[..., 1::2]
    out = resid.copy()
    out[..., 0::2] = resid[..., 0::2] + (mr * cr - mi * ci)
    out[..., 1::2] = resid[..., 1::2] + (mr * ci + mi * cr)
    return out
