"""CPU checks of the surrogate oracle (oracle/surrogate_oracle.py): its gradients against
central finite differences, Adam's first step against its closed form, and the features
against their definition. These pin the checker that the GPU surrogate tests rely on."""

import numpy as np

from oracle import surrogate_oracle as so


def _params(seed=3, scale=0.2):
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, scale, so.num_params())


def test_num_params():
    assert so.num_params() == 256 * 53 + 256 + 3 * (256 * 256 + 256) + 8 * 256 + 8


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    p = _params()
    x = rng.normal(size=(6, so.N_IN))
    y = rng.normal(size=(6, so.N_OUT))
    loss, g = so.loss_and_grads(p, x, y)
    # A mix of indices in every layer's weights and biases.
    offs = np.cumsum([0] + [o * i + o for o, i in so.DIMS])
    idx = []
    for li, (o, i) in enumerate(so.DIMS):
        base = offs[li]
        idx += list(base + rng.choice(o * i, 6, replace=False)) + list(base + o * i + rng.choice(o, 2, replace=False))
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (so.loss_and_grads(pp, x, y)[0] - so.loss_and_grads(pm, x, y)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (k, fd, g[k])


def test_loss_is_mean_squared_error():
    rng = np.random.default_rng(1)
    p = _params()
    x = rng.normal(size=(5, so.N_IN))
    out, _ = so.forward(p, x)
    y = out + 0.5
    loss, _ = so.loss_and_grads(p, x, y)
    assert np.isclose(loss, 0.25)


def test_adam_first_step_closed_form():
    # At t = 1 the bias-corrected moments are g and g^2, so the step is lr g / (|g| + eps).
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 1e-3])
    new, m, v = so.adam_step(p, g, np.zeros(3), np.zeros(3), 1, 0.01)
    np.testing.assert_allclose(new, p - 0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12)
    np.testing.assert_allclose(m, 0.1 * g)
    np.testing.assert_allclose(v, 0.001 * g * g)


def test_features_definition():
    raw = np.zeros((2, 51))
    raw[:, 0] = [0.3, 2.0]
    raw[:, 1] = [0.1, 1.2]
    raw[:, 2] = [182e6, 197e6]
    raw[0, 3:19] = np.arange(16)
    raw[1, 3:19] = 31
    raw[1, 3 + 5] = 32          # dead dipole 5: delay feature 0, both its gains 0
    raw[:, 19:51] = 1.0
    raw[0, 19 + 7] = 0.0
    x = so.features(raw)
    np.testing.assert_allclose(x[:, :4], [[np.sin(0.3), np.cos(0.3), np.sin(0.1), np.cos(0.1)],
                                          [np.sin(2.0), np.cos(2.0), np.sin(1.2), np.cos(1.2)]])
    np.testing.assert_allclose(x[:, 4], [0.0, 1.0])
    np.testing.assert_allclose(x[0, 5:21], np.arange(16) / 31)
    assert x[1, 5 + 5] == 0.0 and x[1, 5 + 4] == 1.0
    assert x[0, 21 + 7] == 0.0 and x[1, 21 + 5] == 0.0 and x[1, 21 + 16 + 5] == 0.0
    assert x[1, 21 + 4] == 1.0


def test_other_shapes_gradients_match_finite_differences():
    """The same code over other widths and depths (the shaped networks' checker)."""
    rng = np.random.default_rng(3)
    for hidden, depth in ((12, 1), (9, 3)):
        shape = so.dims(hidden, depth)
        assert len(shape) == depth + 1 and shape[0] == (hidden, so.N_IN) and shape[-1] == (so.N_OUT, hidden)
        p = rng.normal(0.0, 0.3, so.num_params(shape))
        x = rng.normal(size=(5, so.N_IN))
        y = rng.normal(size=(5, so.N_OUT))
        _, g = so.loss_and_grads(p, x, y, shape)
        h = 1e-6
        for k in rng.choice(p.size, 25, replace=False):
            pp, pm = p.copy(), p.copy()
            pp[k] += h
            pm[k] -= h
            fd = (so.loss_and_grads(pp, x, y, shape)[0] - so.loss_and_grads(pm, x, y, shape)[0]) / (2 * h)
            assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (hidden, depth, k)
