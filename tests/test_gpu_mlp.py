"""GPU surrogate MLP (HIP forward, backward and Adam kernels behind
spires_ext_mlp_*) against the CPU oracle oracle/mlp.py.  FP32 on the device,
float64 in the oracle: the tolerances below are FP32 rounding bounds."""
import numpy as np
import pytest

import spires_amd as sa
from oracle import mlp as om

pytestmark = pytest.mark.gpu
IN, H = 7, 64


def _data(seed, B):
    rng = np.random.default_rng(seed)
    X = rng.uniform(0, 1, (B, IN)).astype(np.float32)
    T = (np.sin(3 * X[:, 0]) + X[:, 1] * X[:, 2] - 0.5 * X[:, 6]).astype(np.float32)
    return X, T


@pytest.mark.parametrize("B", [1, 31, 200])
def test_forward_and_gradient_match_oracle(B):
    net = sa.Surrogate(IN, H, seed=7)
    p = net.params().astype(np.float64)
    X, T = _data(B, B)
    y = net.predict(X)
    yo, _ = om.forward(p, X, IN, H)
    assert np.allclose(y, yo, rtol=1e-5, atol=1e-5)
    g, loss = net.gradient(X, T)
    lo, go = om.loss_and_grad(p, X, T, IN, H)
    assert np.isclose(loss, lo, rtol=1e-5, atol=1e-6)
    assert np.max(np.abs(g - go)) <= 2e-5 * max(1.0, np.max(np.abs(go)))
    g2, _ = net.gradient(X, T)
    assert np.array_equal(g, g2)          # one workgroup, fixed order: bit-reproducible
    net.close()


def test_adam_steps_match_oracle():
    net = sa.Surrogate(IN, H, seed=11)
    X, T = _data(5, 64)
    p = net.params().astype(np.float64)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for step in range(1, 4):
        _, g = om.loss_and_grad(p, X, T, IN, H)
        p, m, v = om.adam(p, g, m, v, step, 1e-3)
    net.train(X, T, epochs=3, lr=1e-3)
    assert np.max(np.abs(net.params() - p)) <= 1e-5
    net.close()


def test_training_fits_a_smooth_target():
    net = sa.Surrogate(IN, H, seed=3)
    X, T = _data(9, 256)
    _, l0 = net.gradient(X, T)
    l1 = net.train(X, T, epochs=1500, lr=1e-2)
    assert l1 < 0.05 * l0, (l0, l1)
    Xv, Tv = _data(10, 256)
    assert np.mean((net.predict(Xv) - Tv) ** 2) < 0.1 * np.var(Tv)
    net.close()


def test_invalid_shapes_raise():
    with pytest.raises(RuntimeError):
        sa.Surrogate(32, 64)
    with pytest.raises(RuntimeError):
        sa.Surrogate(7, 128)
