"""Readout fits over a recorded row range (spires_ext_fit_ridge_records) and
the hyperparameter-search harness built on it (spires_amd.agile_search)."""
import numpy as np
import pytest

import spires_amd as sa
from spires_amd import agile_search as ag

pytestmark = pytest.mark.gpu


def _flif(n, T_mem, seed, dout=2):
    params = [1.0, 0.0, 0.0, 20.0, 0.7, float(T_mem), 0.05, 2.0]
    sa.srand(seed)
    return sa.Reservoir(sa.make_config(n, din=1, dout=dout, rho=0.9, gain=2.0, conn=0.1, neuron="flif_gl",
                                       params=params))


def test_fit_ridge_records_matches_numpy_on_row_range():
    n, T, r0, r1, lam = 160, 700, 120, 530, 1e-4
    r = _flif(n, 96, 11)
    u = np.random.default_rng(3).uniform(0, 0.5, (T, 1))
    y = np.stack([np.roll(u[:, 0], 2), np.roll(u[:, 0], 7)], axis=1)
    r.upload_inputs(u)
    r.reserve_records(T, sa.REC_STATE)
    r.run_resident(0, T, sa.REC_STATE)
    X, _, _ = r.read_records(0, T, states=True)
    _, resid = r.fit_ridge_records(r0, r1, y[r0:r1], lam)
    Xs = X[r0:r1]
    A = Xs.T @ Xs + lam * np.eye(n)
    Wn = np.linalg.solve(A, Xs.T @ y[r0:r1]).T
    kappa = np.linalg.cond(A)
    assert resid < 1e-10
    # two backward-stable solvers agree to about kappa * unit roundoff
    assert np.abs(r.W_out() - Wn).max() <= max(1e-9, 1e-14 * kappa) * np.abs(Wn).max()
    # rows past the recorded range are rejected
    with pytest.raises(sa.SpiresError):
        r.fit_ridge_records(r0, T + 1, np.zeros((T + 1 - r0, 2)), lam)
    r.close()


def test_batched_ridge_and_device_predictions_match_single_fits():
    n, T, r0, r1 = 192, 500, 60, 380
    rs = [_flif(n, 48 + 16 * k, 30 + k, dout=2) for k in range(3)]
    u = np.random.default_rng(4).uniform(0, 0.5, (T, 1))
    y = np.stack([np.roll(u[:, 0], 3), np.roll(u[:, 0], 9)], axis=1)
    for r in rs:
        r.upload_inputs(u)
        r.reserve_records(T, sa.REC_STATE)
        r.run_resident(0, T, sa.REC_STATE)
    lams = [1e-3, 1e-5, 1e-7]
    info = sa.fit_ridge_records_batch(rs, r0, r1, [y[r0:r1]] * 3, lams)
    assert info.shape == (3, 2) and np.all(info[:, 0] == 0) and np.all(info[:, 1] < 1e-10)
    Wb = [r.W_out() for r in rs]
    for r, lam, W in zip(rs, lams, Wb):
        X, _, _ = r.read_records(0, T, states=True)
        r.fit_ridge_records(r0, r1, y[r0:r1], lam)
        Ws = r.W_out()
        A = X[r0:r1].T @ X[r0:r1] + lam * np.eye(n)
        kappa = np.linalg.cond(A)
        tol = max(1e-9, 1e-14 * kappa) * np.abs(Ws).max()
        assert np.abs(W - Ws).max() <= tol
        # outputs of recorded rows under the fitted readout
        np.testing.assert_allclose(r.predict_records(r1, T - r1), X[r1:] @ Ws.T, rtol=1e-12, atol=1e-12)
        r.close()


def test_appending_runs_continue_the_record():
    """Rows [0, t1) from one run and [t1, T) from its continuation equal one
    uninterrupted run's rows, bit for bit."""
    n, T, t1 = 128, 300, 110
    a, b = _flif(n, 64, 5), _flif(n, 64, 5)
    u = np.random.default_rng(8).uniform(0, 0.5, (T, 1))
    for r in (a, b):
        r.upload_inputs(u)
    a.reserve_records(T, sa.REC_STATE)
    a.run_resident(0, T, sa.REC_STATE)
    b.reserve_records(t1, sa.REC_STATE)   # grows, keeping rows, on the second run
    b.run_resident(0, t1, sa.REC_STATE | sa.REC_APPEND)
    b.run_resident(t1, T - t1, sa.REC_STATE | sa.REC_APPEND)
    Xa, _, _ = a.read_records(0, T, states=True)
    Xb, _, _ = b.read_records(0, T, states=True)
    assert np.array_equal(Xa.view(np.uint64), Xb.view(np.uint64))
    a.close()
    b.close()


def test_evaluator_scores_do_not_depend_on_batching_or_rungs():
    """A trial's score at a fidelity is the same whether it runs alone or
    with others, and whether it reaches the fidelity directly or through a
    lower rung: the runs are bit-reproducible, and the readout differs only by
    the rounding of batched versus single factorisations (tolerance 1e-9)."""
    ev = ag.Evaluator(n=256, f_max=600, seed=2, workers=2)
    th = ag.sample_trials(5, seed=7)
    seeds = np.arange(5) + 100
    together = [ag.Trial(t, s) for t, s in zip(th, seeds)]
    ev.advance(together, 300)
    s_together = ev.advance(together, 600)
    alone = []
    for t, s in zip(th, seeds):
        q = ag.Trial(t, s)
        ev.advance([q], 600)
        alone.append(q.scores[600])
        ev.release([q])
    ev.release(together)
    ev.close()
    np.testing.assert_allclose(s_together, np.array(alone), rtol=1e-9)
    assert np.all(np.isfinite(s_together)) and np.all(s_together > 0)
