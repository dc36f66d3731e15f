"""Pins the surrogate-MLP oracle (oracle/mlp.py) without a GPU: the backward
pass against central finite differences, a closed form for a collapsed
network, and Adam's first step (lr * g / (|g| + eps))."""
import numpy as np

from oracle import mlp as om

IN, H = 7, 16


def _rand(seed, B=12):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.5, om.count(IN, H))
    X = rng.uniform(0, 1, (B, IN))
    T = rng.normal(0, 1, B)
    return p, X, T


def test_backward_matches_finite_differences():
    p, X, T = _rand(1)
    _, g = om.loss_and_grad(p, X, T, IN, H)
    rng = np.random.default_rng(2)
    idx = rng.choice(p.size, 40, replace=False)
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (om.loss_and_grad(pp, X, T, IN, H)[0] - om.loss_and_grad(pm, X, T, IN, H)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_collapsed_network_closed_form():
    """W1 = W2 = 0: y = W3 . tanh(b2) + b3 for every sample, so dL/db3 = 2 mean(y - T)
    and dL/dW3 = dL/db3 * tanh(b2)."""
    p, X, T = _rand(3)
    w = om.unpack(p, IN, H)
    p = p.copy()
    off = 0
    for name, shp in om.shapes(IN, H):
        n = int(np.prod(shp))
        if name in ("W1", "W2"):
            p[off:off + n] = 0.0
        off += n
    w = om.unpack(p, IN, H)
    y0 = float(w["W3"] @ np.tanh(w["b2"]) + w["b3"][0])
    loss, g = om.loss_and_grad(p, X, T, IN, H)
    assert np.isclose(loss, np.mean((y0 - T) ** 2))
    gb3 = 2 * np.mean(y0 - T)
    assert np.isclose(g[-1], gb3)
    assert np.allclose(g[-1 - H:-1], gb3 * np.tanh(w["b2"]))


def test_adam_first_step_is_lr_sign():
    """After bias correction the first moment is g and the second g^2, so the
    first step is lr * g / (|g| + eps): lr * sign(g) up to eps / |g|."""
    rng = np.random.default_rng(4)
    p = rng.normal(size=50)
    g = rng.normal(size=50)
    p1, m, v = om.adam(p, g, np.zeros(50), np.zeros(50), 1, 1e-3)
    assert np.allclose(p - p1, 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=0)
    assert np.allclose(p - p1, 1e-3 * np.sign(g), rtol=1e-5, atol=0)
