"""Ridge fits on the batched FP64 Gram (matrix cores), blocked Cholesky and
triangular-solve kernels (orders that are multiples of 64, exact mode):
several N=2048 systems solved together (spires_ext_fit_ridge_records_batch)
and single fits (spires_ext_fit_ridge_records) against numpy's solve of the
same normal equations; more than 8 outputs (rocSOLVER solves on the kernels'
factor); an indefinite system in a batch takes the LU fallback; an order that
is not a multiple of 64 keeps the rocBLAS + rocSOLVER path."""
import numpy as np
import pytest

import spires_amd as sa

pytestmark = pytest.mark.gpu


def _runs(n, count, T, seed, dout=1, T_mem=64):
    params = [1.0, 0.0, 0.0, 20.0, 0.7, float(T_mem), 0.05, 2.0]
    u = np.random.default_rng(seed).uniform(0, 0.5, (T, 1))
    rs = []
    for k in range(count):
        sa.srand(seed + k)
        r = sa.Reservoir(sa.make_config(n, din=1, dout=dout, rho=0.9, gain=2.0, conn=0.02, neuron="flif_gl",
                                        params=params))
        r.upload_inputs(u)
        r.reserve_records(T, sa.REC_STATE)
        r.run_resident(0, T, sa.REC_STATE)
        rs.append(r)
    y = np.stack([np.roll(u[:, 0], 5 + 7 * d) for d in range(dout)], axis=1)
    return rs, y


def _numpy_check(r, r0, r1, y, lam, W):
    """W against numpy's solve of (X'X + lam I) W' = X'Y: two backward-stable
    solvers agree to about kappa * unit roundoff."""
    X, _, _ = r.read_records(r0, r1 - r0, states=True)
    A = X.T @ X + lam * np.eye(X.shape[1])
    Wn = np.linalg.solve(A, X.T @ y[r0:r1]).T
    ev = np.linalg.eigvalsh(A)
    kappa = ev[-1] / ev[0]
    err = np.abs(W - Wn).max() / np.abs(Wn).max()
    print(f"lambda {lam:g}: kappa {kappa:.3g}, |W - numpy| {err:.3g}", flush=True)
    tol = max(1e-9, 1e-14 * kappa)
    assert err <= tol
    return tol


@pytest.mark.parametrize("n,count,dout", [(2048, 4, 1), (1024, 3, 3), (256, 2, 10)])
def test_batched_fit_matches_numpy(n, count, dout):
    T, r0, r1 = 3000, 300, 2250
    print(f"n={n} count={count} dout={dout}", flush=True)
    rs, y = _runs(n, count, T, seed=40, dout=dout)
    lams = [1e-2, 1e-4, 1e-6, 1e-8][:count]
    info = sa.fit_ridge_records_batch(rs, r0, r1, [y[r0:r1]] * count, lams)
    print(info, flush=True)
    assert np.all(info[:, 0] == 0) and np.all(info[:, 1] < 1e-10)
    for r, lam in zip(rs, lams):
        Wb = r.W_out().copy()
        tol = _numpy_check(r, r0, r1, y, lam, Wb)
        # a single fit of the same system (one system per launch)
        _, resid = r.fit_ridge_records(r0, r1, y[r0:r1], lam)
        assert resid < 1e-10
        assert np.abs(r.W_out() - Wb).max() <= tol * np.abs(Wb).max()
        r.close()


def test_indefinite_system_takes_lu_in_a_batch():
    n, T, r0, r1 = 256, 900, 100, 700
    rs, y = _runs(n, 3, T, seed=41, dout=1)
    lams = [1e-3, -1.0, 1e-5]          # X'X - I is indefinite: no Cholesky factor
    info = sa.fit_ridge_records_batch(rs, r0, r1, [y[r0:r1]] * 3, lams)
    assert list(info[:, 0]) == [0.0, 1.0, 0.0]
    for r, lam in zip(rs, lams):
        X, _, _ = r.read_records(0, T, states=True)
        Xs = X[r0:r1]
        A = Xs.T @ Xs + lam * np.eye(n)
        W = np.linalg.solve(A, Xs.T @ y[r0:r1]).T
        kappa = np.linalg.cond(A)
        assert np.abs(r.W_out() - W).max() <= max(1e-9, 1e-14 * kappa) * np.abs(W).max()
        r.close()


def test_order_not_multiple_of_64_keeps_library_path():
    n, T, r0, r1 = 200, 700, 80, 520
    rs, y = _runs(n, 2, T, seed=42, dout=2)
    lams = [1e-4, 1e-6]
    info = sa.fit_ridge_records_batch(rs, r0, r1, [y[r0:r1]] * 2, lams)
    assert np.all(info[:, 0] == 0) and np.all(info[:, 1] < 1e-10)
    for r, lam in zip(rs, lams):
        X, _, _ = r.read_records(0, T, states=True)
        Xs = X[r0:r1]
        A = Xs.T @ Xs + lam * np.eye(n)
        W = np.linalg.solve(A, Xs.T @ y[r0:r1]).T
        kappa = np.linalg.cond(A)
        assert np.abs(r.W_out() - W).max() <= max(1e-9, 1e-14 * kappa) * np.abs(W).max()
        r.close()


@pytest.mark.parametrize("n,count", [(64, 2), (128, 18)])
def test_batched_fit_edge_orders_and_counts(n, count):
    """One 64x64 block (no panel or trailing update), and more systems than one
    Gram launch takes (16); 553 rows leave a partial group of 4 for the Gram."""
    T, r0, r1 = 700, 50, 603
    rs, y = _runs(n, count, T, seed=43, dout=1)
    lams = [10.0 ** -(2 + k % 5) for k in range(count)]
    info = sa.fit_ridge_records_batch(rs, r0, r1, [y[r0:r1]] * count, lams)
    assert np.all(info[:, 0] == 0) and np.all(info[:, 1] < 1e-10)
    for r, lam in zip(rs, lams):
        _numpy_check(r, r0, r1, y, lam, r.W_out())
        r.close()
