"""Harness data formats (host): the Mackey-Glass RK4 generator, the AR(1)
class sequences of the batched workload and the delayed-recall targets."""
import math

import numpy as np

from spires_amd import workloads as w


def test_mackey_glass_rk4_matches_closed_form_without_feedback():
    """With beta = 0 the equation is dx/dt = -gamma x, so x(t) = x0 exp(-gamma t);
    RK4 at dt = 0.1 reproduces it to ~1e-10 relative over 60 time units."""
    x = w.mackey_glass(61, seed=4, beta=0.0, washout=0)
    t = np.arange(61, dtype=np.float64)
    exact = x[0] * np.exp(-0.1 * t)
    assert 1.1 <= x[0] <= 1.3
    assert np.max(np.abs(x - exact) / exact) < 1e-9


def test_mackey_glass_series_properties():
    x = w.mackey_glass(3000, seed=0)
    assert x.shape == (3000,)
    assert np.array_equal(x, w.mackey_glass(3000, seed=0))             # seeded
    assert not np.array_equal(x, w.mackey_glass(3000, seed=1))
    # tau = 17 is chaotic: the attractor stays within (0.2, 1.5) and is not periodic-flat
    assert 0.2 < x.min() and x.max() < 1.5 and x.std() > 0.15
    # halving dt changes the first 100 samples by little (second-order delay interpolation)
    a = w.mackey_glass(100, seed=0, washout=0)
    b = w.mackey_glass(100, seed=0, washout=0, dt=0.05)
    assert np.max(np.abs(a - b)) < 1e-3


def test_class_sequences_shapes_ar1_and_offsets():
    X, lens, labels = w.class_sequences(count=400, seed=9)
    assert X.shape == (int(lens.sum()), 13)
    assert lens.min() >= 30 and lens.max() <= 100 and labels.min() >= 0 and labels.max() <= 9
    # per-sequence demeaning by the true class mean is unknown, so estimate phi from
    # within-sequence differences of consecutive frames after subtracting the sequence mean
    num = den = 0.0
    off = 0
    means = np.zeros((10, 13))
    counts = np.zeros(10)
    for L, c in zip(lens, labels):
        s = X[off:off + L]
        off += L
        means[c] += s.sum(axis=0)
        counts[c] += L
        d = s - s.mean(axis=0)
        num += float((d[1:] * d[:-1]).sum())
        den += float((d[:-1] * d[:-1]).sum())
    phi = num / den
    assert 0.8 < phi < 0.95      # AR(1) phi = 0.9 (the short-sequence mean removal biases it down)
    # class offsets: std 0.5 across classes and channels
    m = means / counts[:, None]
    assert 0.25 < m.std() < 0.9


def test_delayed_recall_targets():
    u = np.arange(1.0, 11.0).reshape(10, 1)
    y = w.delayed_recall_targets(u, [1, 3])
    assert np.array_equal(y[:, 0], np.r_[0.0, u[:9, 0]])
    assert np.array_equal(y[:, 1], np.r_[0.0, 0.0, 0.0, u[:7, 0]])
