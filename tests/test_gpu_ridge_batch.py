"""Batched ridge fits (spires_ext_fit_ridge_records_batch) at the search's size:
N=2048 systems solved together by the batched FP64 Gram (matrix cores),
blocked Cholesky and triangular-solve kernels, against one-at-a-time fits
(rocBLAS dsyrk + rocSOLVER potrf/potrs) and numpy; an indefinite system in the
batch takes the LU fallback; an order that is not a multiple of 64 keeps the
rocSOLVER path."""
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


@pytest.mark.parametrize("n,count,dout", [(2048, 4, 1), (1024, 3, 3)])
def test_batched_fit_matches_single_fits(n, count, dout):
    T, r0, r1 = 3000, 300, 2250
    print(f"n={n} count={count} dout={dout}", flush=True)
    rs, y = _runs(n, count, T, seed=40, dout=dout)
    lams = [1e-2, 1e-4, 1e-6, 1e-8][:count]
    info = sa.fit_ridge_records_batch(rs, r0, r1, [y[r0:r1]] * count, lams)
    print(info, flush=True)
    assert np.all(info[:, 0] == 0) and np.all(info[:, 1] < 1e-10)
    for r, lam in zip(rs, lams):
        Wb = r.W_out().copy()
        pb = r.predict_records(r1, T - r1)
        r.fit_ridge_records(r0, r1, y[r0:r1], lam)
        ps = r.predict_records(r1, T - r1)
        Ws = r.W_out()
        X, _, _ = r.read_records(r0, r1 - r0, states=True)
        ev = np.linalg.eigvalsh(X.T @ X + lam * np.eye(n))
        kappa = ev[-1] / ev[0]
        # two backward-stable solvers agree to about kappa * unit roundoff
        tol = max(1e-9, 1e-14 * kappa)
        print(f"lambda {lam:g}: kappa {kappa:.3g}, |dW| {np.abs(Wb - Ws).max() / np.abs(Ws).max():.3g}, "
              f"|dpred| {np.abs(pb - ps).max() / np.abs(ps).max():.3g}", flush=True)
        assert np.abs(Wb - Ws).max() <= tol * np.abs(Ws).max()
        assert np.abs(pb - ps).max() <= tol * np.abs(ps).max()
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
