"""Pins of the oracle functions that the reference-subset build (oracle/_ref)
cannot check, each against something fixed independently of the oracle's
own formula: brute force, closed forms, textbook densities, library solvers
and a separate restatement of the reference's rand() draw order.

- CSR row dot (sparse.c:68-76): dense brute force in column order, and the
  skip-silent-columns sum, over random spikes and real vectors.
- PSC-homogeneous trace (psc_homogeneous.c:71-84): 1 - d^n under constant
  spikes, the convolution sum_k (1-d) d^(n-k) s_k for a random train.
- Multi-exponential synapse (fractional_multiexp.c:18-157): alpha = 1 is one
  component equal to the PSC trace; otherwise log-spaced time constants,
  weights summing to 1 in the ratios of the Mittag-Leffler spectral density,
  and the trace sum_c c_c (1 - d_c^n) under constant spikes.
- Fractional synapse (fractional.c:103-142): alpha = 1 is forward-Euler
  relaxation 1 - (1 - dt/tau)^n; alpha < 1 equals the lower-triangular
  Grunwald-Letnikov system with closed-form binomial coefficients, solved by
  scipy.
- RLS (reservoir.c:681-767): with forgetting factor 1 and P0 = I/delta it is
  the ridge solution (X'X + delta I)^-1 X'Y; with lambda < 1 the
  exponentially weighted one.
- Delta rule (reservoir.c:244-280): on a fixed state the output error
  contracts by (1 - lr |V|^2) per update.
- Power iteration (sparse.c:93-127): constant row sums give exactly that sum,
  a scaled cycle its scale, a symmetric non-negative matrix its largest
  eigenvalue (numpy eigvalsh).
- Construction (reservoir.c:307-545): W and W_in for every topology equal,
  bit for bit, those of an independent Python restatement driving glibc
  rand() (tests/topology_restate.py).
"""
import math

import mpmath
import numpy as np
import pytest
from scipy import linalg, special

from oracle import pyoracle as po
from tests import topology_restate as tr


def _random_csr(n, p, rng):
    W = np.where(rng.random((n, n)) < p, rng.standard_normal((n, n)), 0.0)
    np.fill_diagonal(W, 0.0)
    rp = np.zeros(n + 1, np.int64)
    rp[1:] = np.cumsum((W != 0).sum(axis=1))
    rows, cols = np.nonzero(W)
    return W, rp, cols.astype(np.int32), W[rows, cols]


def _seq_sum(v):
    """Sequential left-to-right float sum starting at +0.0 (cumsum accumulates in order)."""
    return float(np.cumsum(np.concatenate([[0.0], v]))[-1])


def test_csr_row_dot_brute_force():
    rng = np.random.default_rng(1)
    n = 150
    W, rp, ci, val = _random_csr(n, 0.15, rng)
    for x in (rng.integers(0, 2, n).astype(float), rng.standard_normal(n)):
        got = po.probe_csr_rows(rp, ci, val, x)
        dense = np.array([_seq_sum(W[i] * x) for i in range(n)])       # every column, in column order
        assert np.array_equal(got.view(np.uint64), dense.view(np.uint64))
        if set(np.unique(x)) <= {0.0, 1.0}:                              # spikes: silent columns skipped
            skip = np.array([_seq_sum(W[i, (W[i] != 0) & (x == 1.0)]) for i in range(n)])
            assert np.array_equal(got.view(np.uint64), skip.view(np.uint64))
    # a transposed operand would fail: W is not symmetric
    assert not np.allclose(po.probe_csr_rows(rp, ci, val, np.arange(n, dtype=float)), W.T @ np.arange(n))


def _synapse_res(n, synapse, sp, dt=1.0):
    rp = np.arange(n + 1, dtype=np.int64)                    # any weights: the probe reads the synapse vector
    ci = ((np.arange(n) + 1) % n).astype(np.int32)
    cfg = po.make_config(n, dt=dt, neuron="lif_discrete", params=[0.0, 1.0, 0.1, 0.0], synapse=synapse,
                         syn_params=sp)
    return po.OracleReservoir(cfg, csr=(rp, ci, np.ones(n)), W_in=np.zeros((n, 1)))


@pytest.mark.parametrize("tau,dt", [(3.0, 1.0), (7.5, 0.5)])
def test_psc_homogeneous_closed_forms(tau, dt):
    n, T = 3, 300
    rng = np.random.default_rng(2)
    s = np.zeros((T, n))
    s[:, 0] = 1.0                                           # constant firing
    s[:, 2] = rng.integers(0, 2, T)                         # a random train
    eff = po.probe_synapse(_synapse_res(n, "psc_homogeneous", [tau], dt), s)
    d = math.exp(-dt / tau)
    t = np.arange(1, T + 1)
    np.testing.assert_allclose(eff[:, 0], 1.0 - d ** t, rtol=0, atol=1e-13)
    assert np.all(eff[:, 1] == 0.0)
    conv = np.array([sum((1 - d) * d ** (k - j) * s[j, 2] for j in range(k + 1)) for k in range(T)])
    np.testing.assert_allclose(eff[:, 2], conv, rtol=0, atol=1e-13)


def _ml_density(r, a):
    """Spectral density of the Mittag-Leffler relaxation E_a(-t^a) (Gorenflo and
    Mainardi): K_a(r) = r^(a-1) sin(a pi) / (pi (r^(2a) + 2 r^a cos(a pi) + 1))."""
    r, a = mpmath.mpf(r), mpmath.mpf(a)
    return r ** (a - 1) * mpmath.sin(a * mpmath.pi) / (mpmath.pi * (r ** (2 * a) + 2 * r ** a * mpmath.cos(a * mpmath.pi) + 1))


def test_multiexp_single_component_is_the_psc_trace():
    n, T = 4, 120
    s = np.random.default_rng(3).integers(0, 2, (T, n)).astype(float)
    me = _synapse_res(n, "fractional_multiexp", [1.0, 4.0, 1.0, 30.0, 6])
    tau, c = po.probe_multiexp(me)
    assert tau.tolist() == [4.0] and c.tolist() == [1.0]
    a = po.probe_synapse(me, s)
    b = po.probe_synapse(_synapse_res(n, "psc_homogeneous", [4.0]), s)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("alpha,tsyn,tmin,tmax,N", [(0.6, 2.0, 0.5, 30.0, 6), (0.35, 5.0, 1.0, 200.0, 9)])
def test_multiexp_weights_and_trace(alpha, tsyn, tmin, tmax, N):
    n, T = 2, 250
    me = _synapse_res(n, "fractional_multiexp", [alpha, tsyn, tmin, tmax, N])
    tau, c = po.probe_multiexp(me)
    assert len(tau) == N
    assert tau[0] == pytest.approx(tmin, rel=1e-14) and tau[-1] == pytest.approx(tmax, rel=1e-14)
    ratio = tau[1:] / tau[:-1]
    np.testing.assert_allclose(ratio, (tmax / tmin) ** (1.0 / (N - 1)), rtol=1e-13)
    assert np.all(c >= 0) and c.sum() == pytest.approx(1.0, abs=1e-14)
    dens = np.array([float(_ml_density(tsyn / t, alpha)) for t in tau])
    np.testing.assert_allclose(c / c[0], dens / dens[0], rtol=1e-12)
    s = np.zeros((T, n))
    s[:, 0] = 1.0
    eff = po.probe_synapse(me, s)
    t = np.arange(1, T + 1)[:, None]
    expect = ((1.0 - np.exp(-1.0 / tau)[None, :] ** t) * c[None, :]).sum(axis=1)
    np.testing.assert_allclose(eff[:, 0], expect, rtol=0, atol=1e-13)
    assert np.all(eff[:, 1] == 0.0)


@pytest.mark.parametrize("tau,dt", [(4.0, 1.0), (10.0, 0.25)])
def test_fractional_alpha_one_is_forward_euler(tau, dt):
    n, T = 2, 200
    s = np.zeros((T, n))
    s[:, 0] = 1.0
    eff = po.probe_synapse(_synapse_res(n, "fractional", [1.0, tau, 60.0], dt), s)
    t = np.arange(1, T + 1)
    np.testing.assert_allclose(eff[:, 0], 1.0 - (1.0 - dt / tau) ** t, rtol=0, atol=1e-13)
    assert np.all(eff[:, 1] == 0.0)


@pytest.mark.parametrize("alpha,tau,tmem,dt", [(0.6, 3.0, 40.0, 1.0), (0.85, 2.0, 12.0, 0.5)])
def test_fractional_gl_lower_triangular_system(alpha, tau, tmem, dt):
    n, T = 3, 160
    rng = np.random.default_rng(4)
    s = rng.integers(0, 2, (T, n)).astype(float)
    eff = po.probe_synapse(_synapse_res(n, "fractional", [alpha, tau, tmem], dt), s)
    L = int(tmem / dt)
    k = np.arange(L)
    c = (-1.0) ** k * special.binom(alpha, k)               # GL weights, closed form
    da = dt ** alpha
    A = np.eye(T)
    for m in range(T):
        for q in range(1, min(m, L - 1) + 1):
            A[m, m - q] = c[q] + (da / tau if q == 1 else 0.0)
    for j in range(n):
        expect = linalg.solve_triangular(A, da * s[:, j] / tau, lower=True)
        np.testing.assert_allclose(eff[:, j], expect, rtol=1e-10, atol=1e-12)


def _lif_reservoir(n, seed):
    rng = np.random.default_rng(seed)
    _, rp, ci, val = _random_csr(n, 0.2, rng)
    val *= 0.9 / po.csr_spectral_radius(rp, ci, val)
    cfg = po.make_config(n, din=1, dout=2, gain=1.0, neuron="lif_discrete", params=[0.0, 1.0, 0.2, 0.05])
    return po.OracleReservoir(cfg, csr=(rp, ci, val), W_in=rng.uniform(-1, 1, (n, 1)))


@pytest.mark.parametrize("delta,lam", [(0.5, 1.0), (2.0, 0.98)])
def test_rls_is_weighted_ridge(delta, lam):
    n, T = 24, 180
    rng = np.random.default_rng(5)
    u = rng.uniform(0, 1, (T, 1))
    Y = np.column_stack([np.sin(np.arange(T) / 7.0), rng.standard_normal(T)])
    o = _lif_reservoir(n, 6)
    o.train_rls(u, Y, delta, lam)
    W = o.W_out()
    o.reset()
    X, _, _ = o.run_record(u, T, raster=False)
    wts = lam ** np.arange(T - 1, -1, -1)                  # lambda^(T-1-t)
    A = (X * wts[:, None]).T @ X + (lam ** T) * delta * np.eye(n)
    B = (X * wts[:, None]).T @ Y
    Wn = np.linalg.solve(A, B).T
    kappa = np.linalg.cond(A)
    assert np.linalg.norm(W - Wn) <= max(1e-10, 1e-13 * kappa) * np.linalg.norm(Wn)


def test_delta_rule_contracts_the_error():
    n = 20
    o = _lif_reservoir(n, 7)
    o.run_record(np.random.default_rng(8).uniform(0, 1, (15, 1)), 15)
    V = o.state()
    o.set_W_out(np.zeros((2, n)))
    target = np.array([0.7, -1.3])
    lr = 0.5 / float(V @ V)
    e0 = target - o.output()
    for k in range(1, 9):
        o.train_online(target, lr)
        e = target - o.output()
        np.testing.assert_allclose(e, e0 * 0.5 ** k, rtol=1e-11, atol=1e-14)


def _csr_of(W):
    rp = np.zeros(W.shape[0] + 1, np.int64)
    rp[1:] = np.cumsum((W != 0).sum(axis=1))
    r, c = np.nonzero(W)
    return rp, c.astype(np.int32), W[r, c]


def test_power_iteration_known_spectra():
    rng = np.random.default_rng(9)
    n = 64
    W = np.zeros((n, n))
    for i in range(n):                                       # four entries of 0.1875 per row: sums exactly 0.75
        W[i, rng.choice(n, 4, replace=False)] = 0.1875
    assert po.csr_spectral_radius(*_csr_of(W)) == 0.75
    P = np.roll(np.eye(n), 1, axis=1) * 1.3                  # scaled cycle: every eigenvalue has modulus 1.3
    assert po.csr_spectral_radius(*_csr_of(P)) == 1.3
    S = np.where(rng.random((n, n)) < 0.2, rng.random((n, n)), 0.0)
    S = S + S.T                                              # symmetric, non-negative: Perron root = max eigenvalue
    lam = po.csr_spectral_radius(*_csr_of(S))
    assert lam == pytest.approx(np.linalg.eigvalsh(S)[-1], rel=1e-5)


@pytest.mark.parametrize("conn_type,conn,din,ei", [("random", 0.1, 2, 0.8), ("small_world", 10 / 59, 1, 0.7),
                                                   ("scale_free", 4 / 59, 3, 0.6), ("small_world", 0.9, 1, 0.5)])
def test_construction_draw_order(conn_type, conn, din, ei):
    n = 60
    tr.srand(123)
    W_ref, Wi_ref = tr.init_weights(n, din, conn_type, conn, ei)
    po.srand(123)
    W, Wi = po.init_weights_dense(n, din, conn_type, conn, ei)
    assert np.array_equal(Wi.view(np.uint64), Wi_ref.view(np.uint64))
    assert np.array_equal(W.view(np.uint64), W_ref.view(np.uint64))
    assert np.count_nonzero(W) > 0


def test_lu_and_ridge_solves_against_numpy():
    """or_lu_solve (math_utils.c:187-241, partial pivoting) and or_ridge_solve
    (reservoir.c:573-679: Gram + lambda I, one LU per output) against numpy."""
    rng = np.random.default_rng(10)
    for n in (1, 7, 60):
        A = rng.standard_normal((n, n)) + n * np.eye(n) * (rng.random() < 0.5)
        b = rng.standard_normal(n)
        x, rc = po.lu_solve(A, b)
        assert rc == 0
        np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-10 * np.linalg.cond(A), atol=1e-13)
    x, rc = po.lu_solve(np.zeros((3, 3)), np.ones(3))
    assert rc != 0                                           # singular pivot reported
    X = rng.standard_normal((90, 20))
    Y = rng.standard_normal((90, 3))
    W, rc = po.ridge_solve(X, Y, 0.3)
    assert rc == 0
    Wn = np.linalg.solve(X.T @ X + 0.3 * np.eye(20), X.T @ Y).T
    np.testing.assert_allclose(W, Wn, rtol=1e-11, atol=1e-13)
