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


def test_evaluator_scores_are_batch_invariant():
    """A trial's validation score does not depend on which other trials share
    its launch (exact mode is bit-reproducible)."""
    ev = ag.Evaluator(n=256, f_max=600, seed=2)
    th = ag.sample_trials(5, seed=7)
    seeds = np.arange(5) + 100
    together = ev.evaluate(th, seeds, 600)
    alone = np.array([ev.evaluate(th[k:k + 1], seeds[k:k + 1], 600)[0] for k in range(5)])
    assert np.array_equal(together, alone)
    assert np.all(np.isfinite(together)) and np.all(together > 0)
