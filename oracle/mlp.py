"""CPU oracle of the AGILE surrogate MLP (in -> H -> H -> 1, tanh, linear output).

TEST INFRASTRUCTURE ONLY: tests/ use it to check the GPU forward, backward and
Adam kernels (spires_ext_mlp_*); the product never imports it.  It has no
counterpart in the reference (the reference has no surrogate model).  All
arithmetic is float64; the backward pass is pinned against central finite
differences in tests/test_oracle_mlp.py.

Parameter vector layout (same as the device buffer):
  W1 [H x in] row-major, b1 [H], W2 [H x H], b2 [H], W3 [H], b3 [1].
"""
import numpy as np


def shapes(in_dim, H):
    return [("W1", (H, in_dim)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H,)), ("b3", (1,))]


def count(in_dim, H):
    return H * in_dim + H + H * H + H + H + 1


def unpack(p, in_dim, H):
    p = np.asarray(p, dtype=np.float64)
    out, o = {}, 0
    for name, shp in shapes(in_dim, H):
        n = int(np.prod(shp))
        out[name] = p[o:o + n].reshape(shp)
        o += n
    return out


def forward(p, X, in_dim, H):
    w = unpack(p, in_dim, H)
    X = np.asarray(X, dtype=np.float64).reshape(-1, in_dim)
    h1 = np.tanh(X @ w["W1"].T + w["b1"])
    h2 = np.tanh(h1 @ w["W2"].T + w["b2"])
    y = h2 @ w["W3"] + w["b3"][0]
    return y, (X, h1, h2)


def loss_and_grad(p, X, T, in_dim, H):
    """L = mean((y - T)^2) and dL/dp (float64, same layout as p)."""
    w = unpack(p, in_dim, H)
    y, (X, h1, h2) = forward(p, X, in_dim, H)
    T = np.asarray(T, dtype=np.float64).reshape(-1)
    B = X.shape[0]
    d = y - T
    loss = float(np.mean(d * d))
    dy = 2.0 * d / B
    gW3 = h2.T @ dy
    gb3 = np.array([dy.sum()])
    dz2 = np.outer(dy, w["W3"]) * (1.0 - h2 * h2)
    gW2 = dz2.T @ h1
    gb2 = dz2.sum(axis=0)
    dz1 = (dz2 @ w["W2"]) * (1.0 - h1 * h1)
    gW1 = dz1.T @ X
    gb1 = dz1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, gb3])
    return loss, g


def adam(p, g, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, bias-corrected); returns (p, m, v)."""
    m = beta1 * m + (1.0 - beta1) * g
    v = beta2 * v + (1.0 - beta2) * g * g
    mh = m / (1.0 - beta1 ** step)
    vh = v / (1.0 - beta2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
