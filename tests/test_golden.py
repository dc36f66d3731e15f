"""The oracle against committed golden vectors of the reference's own neuron
models and AGILE (tests/golden/make_golden.py wrote them from the reference
sources); runs without /root/reference."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import pyoracle as po

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "neurons_ref.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


@pytest.mark.parametrize("name", ["lif_discrete", "lif_bio", "flif_gl", "flif_gl_dt05", "flif_caputo",
                                  "flif_diffusive"])
def test_oracle_neurons_match_golden(oracle, gold, name):
    neuron = name.replace("_dt05", "")
    p = list(gold[f"{name}/params"])
    dt = float(gold[f"{name}/dt"][0])
    u = gold[f"{name}/u"]
    T, n = u.shape
    cfg = po.make_config(n, din=n, dout=1, gain=1.0, dt=dt, neuron=neuron, params=p)
    res = po.OracleReservoir(cfg, csr=(np.zeros(n + 1, np.int64), np.zeros(0, np.int32), np.zeros(0)),
                             W_in=np.eye(n))
    X, R, _ = res.run_record(u, T)
    assert np.array_equal(X.view(np.uint64), gold[f"{name}/V"].view(np.uint64))
    assert np.array_equal(R, gold[f"{name}/S"])
    assert R.sum() > 0


def _loss(theta_p, dim, ctx):
    x = np.ctypeslib.as_array(theta_p, shape=(dim,))
    return float(np.sum((x - 0.3) ** 2 * np.arange(1, dim + 1)))


def test_oracle_agile_matches_golden(oracle, gold):
    dim = 4
    th0 = np.full(dim, 0.8); lo = np.zeros(dim); hi = np.ones(dim)
    o = po.AgileOpts(1.3, 1e-3, 0.3, 15.0, 0.9, 1e-2, 40, 10, 1e-3, 1, 0xA91E, 0)
    best = np.empty(dim); rep = po.AgileReport()
    oracle.lib().or_agile_optimize(po.dptr(th0), po.dptr(lo), po.dptr(hi), dim, po.LOSS_FN(_loss), None,
                                   C.byref(o), po.dptr(best), C.byref(rep))
    assert np.array_equal(best, gold["agile/best"])
    r = gold["agile/report"]
    assert (rep.best_loss, rep.best_iter, rep.total_evals, rep.total_iters, rep.exited_refinement) == \
        (r[0], int(r[1]), int(r[2]), int(r[3]), int(r[4]))
