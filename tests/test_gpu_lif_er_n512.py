"""The reference's CPU-runnable case at full size: discrete LIF N=512
[V_0=0, V_th=1, leak 0.1, bias 0], Erdos-Renyi p=0.1, U[-1,1] weights
rescaled to rho=0.9 by the reference's power iteration, gain 1.0, a
Mackey-Glass input of 5000 steps.  Rasters and membrane traces are bit-exact
against the oracle over all 5000 steps; the ridge readout (lambda=1e-6,
one-step-ahead targets) agrees with the oracle's per-output LU to 1e-9
relative (scaled by the condition number where that is larger)."""
import numpy as np
import pytest

import spires_amd as sa
from spires_amd import workloads
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

N, T = 512, 5000
KW = dict(din=1, dout=1, rho=0.9, gain=1.0, conn=0.1, conn_type="random", neuron="lif_discrete",
          params=[0.0, 1.0, 0.1, 0.0])


def test_lif_er_n512_full_size_bit_exact_and_ridge():
    series = workloads.mackey_glass(T + 1, seed=0)
    u = series[:T].reshape(T, 1)
    y = series[1:T + 1].reshape(T, 1)
    sa.srand(7)
    r = sa.Reservoir(sa.make_config(N, **KW))
    po.srand(7)
    o = po.OracleReservoir(po.make_config(N, **KW))
    X, R, _ = r.run_record(u)
    Xo, Ro, _ = o.run_record(u, T)
    assert Ro.sum() > 0
    assert np.array_equal(R, Ro)
    assert np.array_equal(X.view(np.uint64), Xo.view(np.uint64))
    # reference ridge semantics: reset, run, fit on every state
    r.train_ridge(u, y, 1e-6)
    o.train_ridge(u, y, 1e-6)
    W, Wo = r.W_out(), o.W_out()
    A = Xo.T @ Xo + 1e-6 * np.eye(N)
    kappa = np.linalg.cond(A)
    print("kappa", kappa, "max |dW| / max |W|", np.abs(W - Wo).max() / np.abs(Wo).max())
    assert np.abs(W - Wo).max() <= max(1e-9, 1e-14 * kappa) * np.abs(Wo).max()
    # the config's task: washout 500, one-step-ahead NRMSE on the training series
    info = r.train_ridge_washout(u, y[500:], 500, 1e-6)
    Xw = Xo[500:]
    pred = Xw @ r.W_out()[0]
    nrmse = np.sqrt(np.mean((pred - y[500:, 0]) ** 2)) / np.std(y[500:, 0])
    print("washout ridge residual", info[1], "one-step NRMSE", nrmse)
    assert nrmse < 0.5
