"""spires_optimize on the GPU path (spires_api.c:298-498, spires_opt_agile.c:137-213).

- One evaluation of the optimizer's objective (spires_ext_optimize_evaluate:
  create -> ridge on the training slice -> reset -> MSE on the rest, per
  seed) against the same steps run on the oracle from the same rand() state.
- spires_optimize end to end on a short Mackey-Glass series: it returns 0, a
  configuration inside the search bounds, a score consistent with its own
  mean / std / cost, and the same result bit for bit when run again from the
  same srand() seed.  (The AGILE trajectory itself is pinned against the
  reference's agile.c on the CPU, tests/test_abi.py.)
"""
import ctypes as C
import time

import numpy as np
import pytest

import spires_amd as sa
from oracle import pyoracle as po
from spires_amd import workloads

pytestmark = pytest.mark.gpu

PARAMS = [1.0, 0.0, 0.0, 20.0, 0.8, 16.0, 0.05, 2.0]     # FLIF-GL, T_mem 16
SNAP_N = (200, 400, 800, 1200)


def _series(T, seed=0):
    s = workloads.mackey_glass(T + 1, seed=seed)
    return np.ascontiguousarray(s[:T].reshape(T, 1)), np.ascontiguousarray(s[1:T + 1].reshape(T, 1))


def _oracle_evaluate(theta, x, y, frac, seeds, lam_var, lam_cost):
    """eval_adapter + eval_mse (spires_api.c:298-438) on the oracle."""
    N = min(SNAP_N, key=lambda v: abs(v - theta[6]))
    topo = int(min(2, max(0, round(theta[7]))))
    T = len(x)
    n_train = int(T * frac)
    params = list(PARAMS)
    params[4] = theta[0]
    metrics = []
    for _ in range(seeds):
        cfg = po.make_config(N, din=1, dout=1, rho=theta[1], ei=theta[4], gain=theta[2], conn=theta[3],
                             conn_type=topo, neuron="flif_gl", params=params)
        o = po.OracleReservoir(cfg)
        o.train_ridge(x[:n_train], y[:n_train], 10.0 ** theta[5])
        o.reset()
        _, _, yh = o.run_record(x[n_train:], T - n_train, states=False, raster=False, outputs=True)
        mse = float(np.sum((y[n_train:] - yh) ** 2)) / (T - n_train)
        metrics.append(1.0 / (1.0 + mse))
    m = np.array(metrics)
    mean = float(m.sum() / len(m))
    std = float(np.sqrt(max(0.0, float((m * m).sum() / len(m)) - mean * mean)))
    cost = N * (1.0 / max(theta[3], 1e-8)) * 1e-6
    return mean, std, cost, mean - lam_var * std - lam_cost * cost


@pytest.mark.parametrize("theta", [
    [0.55, 0.90, 1.00, 0.20, 0.80, -2.0, 200.0, 0.0],
    [0.70, 1.00, 0.50, 0.10, 0.90, -3.0, 430.0, 1.2],
    [0.90, 0.80, 1.20, 0.30, 0.75, -1.5, 180.0, 2.0],
], ids=["random-200", "small_world-400", "scale_free-200"])
def test_optimize_evaluate_matches_oracle(theta):
    print(f"case: spires_optimize objective at {theta}", flush=True)
    x, y = _series(80)
    frac, seeds, lv, lc = 0.6, 2, 0.1, 1e-4
    cfg = sa.make_config(200, din=1, dout=1, neuron="flif_gl", params=PARAMS)
    budget = sa.OptBudget(frac, seeds, 0.0, 0, 0.0)
    score = sa.OptScore(lv, lc, 0, 0)
    out = np.zeros(5)
    th = np.array(theta, dtype=np.float64)
    sa.srand(31)
    rc = sa.lib().spires_ext_optimize_evaluate(C.byref(cfg[0]), C.byref(budget), C.byref(score), sa._dp(th),
                                               sa._dp(x), sa._dp(y), len(x), sa._dp(out))
    assert rc == 0
    assert cfg[1][4] == theta[0]                             # alpha written into the caller's params
    po.srand(31)
    mean, std, cost, sc = _oracle_evaluate(theta, x, y, frac, seeds, lv, lc)
    print(f"GPU {out.tolist()} oracle {[mean, std, cost, sc]}", flush=True)
    assert out[3] == 1.0
    assert out[0] == pytest.approx(mean, rel=1e-9)
    assert out[1] == pytest.approx(std, abs=1e-8)
    assert out[2] == cost
    assert out[4] == pytest.approx(sc, rel=1e-8, abs=1e-9)


def _optimize(x, y, seed):
    cfg = sa.make_config(200, din=1, dout=1, neuron="flif_gl", params=PARAMS)
    budget = sa.OptBudget(0.5, 1, 0.0, 0, 0.0)
    score = sa.OptScore(0.05, 1e-4, 0, 0)
    res = sa.OptResult()
    sa.srand(seed)
    t0 = time.perf_counter()
    rc = sa.lib().spires_optimize(C.byref(cfg[0]), C.byref(budget), 1, C.byref(score), C.byref(res), sa._dp(x),
                                  sa._dp(y), len(x))
    return rc, res, time.perf_counter() - t0


def test_spires_optimize_end_to_end_on_gpu():
    print("case: spires_optimize end to end, 60-step Mackey-Glass, one budget", flush=True)
    x, y = _series(60, seed=2)
    rc, res, dt = _optimize(x, y, 5)
    b = res.best_config
    print(f"rc {rc} in {dt:.1f}s: N={b.num_neurons} topo={b.connectivity_type} rho={b.spectral_radius:.4f} "
          f"gain={b.input_strength:.4f} conn={b.connectivity:.4f} ei={b.ei_ratio:.4f} "
          f"log10 lambda={res.best_log10_ridge:.4f} score={res.best_score:.6f} mean={res.metric_mean:.6f}",
          flush=True)
    assert rc == 0
    assert b.num_neurons in SNAP_N and b.connectivity_type in (0, 1, 2)
    assert 0.75 <= b.spectral_radius <= 1.05 and 0.2 <= b.input_strength <= 1.5
    assert 0.08 <= b.connectivity <= 0.35 and 0.7 <= b.ei_ratio <= 1.1 and -5.0 <= res.best_log10_ridge <= -1.0
    assert 0.0 < res.metric_mean <= 1.0 and res.metric_std == 0.0      # one seed
    cost = b.num_neurons / b.connectivity * 1e-6
    assert res.best_score == pytest.approx(res.metric_mean - 0.05 * res.metric_std - 1e-4 * cost, abs=1e-15)
    rc2, res2, _ = _optimize(x, y, 5)
    assert rc2 == 0
    def key(r):
        c = r.best_config
        return (c.num_neurons, c.connectivity_type, c.spectral_radius, c.input_strength, c.connectivity, c.ei_ratio,
                r.best_log10_ridge, r.best_score, r.metric_mean, r.metric_std)
    assert key(res2) == key(res)                                # same srand seed, same search, same result
