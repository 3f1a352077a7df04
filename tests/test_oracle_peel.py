"""Pins of the peel-loop restatement (oracle/peel_oracle.py) on the CPU."""

import math

import numpy as np

from oracle import oracle as orc
from oracle import peel_oracle as po


def test_xyzs_to_uvws_equals_cross_uvws_form():
    """src/params/peel/tests.rs:2255-2378: the peel kernel's UVWs (difference, then rotation)
    equal xyzs_to_cross_uvws (rotation, then difference) to rounding."""
    rng = np.random.default_rng(1)
    xyz = rng.uniform(-1500, 1500, size=(2, 9, 3))
    lmsts = [0.3, 0.35]
    got = po.xyzs_to_uvws(xyz, lmsts, 0.1, -0.47)
    for t in range(2):
        want = orc.cross_uvws(xyz[t], lmsts[t] - 0.1, -0.47)
        np.testing.assert_allclose(got[t], want, rtol=0, atol=1e-9)


def test_xyzs_to_uvws_at_the_pole():
    """Pointing at Dec = 90 deg, HA = 0: u = y, v = -x, w = z for the baseline XYZ difference."""
    xyz = np.array([[[1.0, 2.0, 3.0], [0.5, -1.0, 4.0]]])
    got = po.xyzs_to_uvws(xyz, [0.7], 0.7, math.pi / 2)[0, 0]
    d = xyz[0, 0] - xyz[0, 1]
    np.testing.assert_allclose(got, [d[1], -d[0], d[2]], atol=1e-15)


def test_rotate_identity_and_round_trip():
    rng = np.random.default_rng(2)
    vis = rng.normal(size=(2, 3, 5, 8)).astype(np.float32)
    u = rng.normal(size=(2, 5, 3)) * 100
    np.testing.assert_array_equal(po.rotate(vis, u, u, [1.0, 2.0, 3.0]), vis)
    u2 = u.copy()
    u2[..., 2] += rng.normal(size=(2, 5)) * 50
    back = po.rotate(po.rotate(vis, u, u2, [1.0, 2.0, 3.0]), u2, u, [1.0, 2.0, 3.0])
    np.testing.assert_allclose(back, vis, rtol=0, atol=2e-6 * np.abs(vis).max())


def test_add_model_reference_kat():
    """src/params/peel/tests.rs:2524-2600: model Jones [0, i, 1, 1+i] everywhere, uvw (0, 1, 0),
    alpha 0.1, beta 0.2, gain 0.3, lambdas 1..20: every entry's norm is the model's times gain
    (epsilon 1e-7)."""
    nt, nc, nb = 10, 20, 30 * 29 // 2
    lambdas = 1.0 + np.arange(nc)
    model = np.zeros((nt, nc, nb, 8), np.float32)
    model[...] = np.array([0, 0, 0, 1, 1, 0, 1, 1], np.float32)
    uvws = np.zeros((nt, nb, 3))
    uvws[..., 1] = 1.0
    res = po.add_model(np.zeros_like(model), model, 0.1, 0.2, 0.3, lambdas, uvws)
    norms = np.hypot(res[..., 0::2], res[..., 1::2])
    want = np.array([0.0, 1.0, 1.0, math.sqrt(2.0)]) * np.float32(0.3)
    np.testing.assert_allclose(norms, np.broadcast_to(want, norms.shape), rtol=0, atol=1e-7)


def test_add_model_without_offset_adds_gain_times_model():
    rng = np.random.default_rng(3)
    resid = rng.normal(size=(1, 2, 3, 8)).astype(np.float32)
    model = rng.normal(size=(1, 2, 3, 8)).astype(np.float32)
    got = po.add_model(resid, model, 0.0, 0.0, 1.0, [1.0, 2.0], np.zeros((1, 3, 3)))
    np.testing.assert_array_equal(got, resid + model)
