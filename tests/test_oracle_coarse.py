"""The oracle's coarse-graining (or_coarse_grain, restating reference
reservoir.c:782-967) pinned without the reference build (its reservoir.c
needs cblas/lapacke): a hand-worked four-neuron case fixes the merge order, the
two combination rules (average when both weights are nonzero, sum otherwise)
and the V / W_in / W_out bookkeeping; on random reservoirs the result must
satisfy the loop's stopping criterion (no surviving pair exceeds the
threshold), keep each readout row's total, and be the identity when the
threshold exceeds every weight."""
import numpy as np
import pytest

from oracle import pyoracle as po

LIF = [0.0, 1.0, 0.1, 0.0]


def _csr(W):
    n = W.shape[0]
    rp, ci, val = [0], [], []
    for i in range(n):
        for j in range(n):
            if W[i, j] != 0.0:
                ci.append(j)
                val.append(W[i, j])
        rp.append(len(ci))
    return np.array(rp, np.int64), np.array(ci, np.int32), np.array(val)


def _oracle(W, W_in, W_out, steps=3, seed=0):
    n, din = W_in.shape
    cfg = po.make_config(n, din=din, dout=W_out.shape[0], conn_type="random", neuron="lif_discrete",
                         params=LIF, gain=1.0)
    o = po.OracleReservoir(cfg, csr=_csr(W), W_in=W_in, W_out=W_out)
    u = np.random.default_rng(seed).uniform(0.2, 0.9, (steps, din))
    o.run_record(u, steps)   # distinct membrane values to average
    return o


def test_hand_worked_four_neurons():
    W = np.array([[0.0, 0.9, 0.2, 0.0],
                  [0.1, 0.0, 0.3, 0.6],
                  [0.4, 0.0, 0.0, 0.05],
                  [0.0, 0.2, 0.1, 0.0]])
    W_in = np.array([[0.5, 0.0], [0.25, -0.5], [1.0, 0.75], [0.125, 0.0]])
    W_out = np.array([[0.3, -0.2, 0.7, 0.05]])
    o = _oracle(W, W_in, W_out)
    V = o.state()
    q = o.coarse_grain(0.5)
    # pass 1, row 0: (0,1) merges (0.9 > 0.5): W[0][2] = W[2][0] = (0.2 + 0.3) / 2,
    # W[0][3] = W[3][0] = 0 + 0.6 (one weight zero: sum); (0,2) stays (0.25 both ways);
    # (0,3) merges (0.6 > 0.5): W[0][2] = W[2][0] = (0.25 + W[3][2] = 0.1) / 2.
    # Row 2 meets only dead neurons; pass 2 merges nothing.  Survivors: 0 and 2.
    w02 = ((0.2 + 0.3) / 2.0 + 0.1) / 2.0
    assert q.N == 2
    np.testing.assert_array_equal(q.dense(), np.array([[0.0, w02], [w02, 0.0]]))

    def comb(a, b):
        return a + b if a * b == 0.0 else (a + b) / 2.0

    win0 = [comb(comb(W_in[0, k], W_in[1, k]), W_in[3, k]) for k in range(2)]
    np.testing.assert_array_equal(q.W_in(), np.array([win0, W_in[2]]))
    np.testing.assert_array_equal(q.W_out(), np.array([[(0.3 + -0.2) + 0.05, 0.7]]))
    np.testing.assert_array_equal(q.state(), np.array([((V[0] + V[1]) / 2.0 + V[3]) / 2.0, V[2]]))


@pytest.mark.parametrize("n,p,thr,seed", [(120, 0.1, 0.3, 1), (200, 0.05, 0.6, 2), (150, 0.2, 0.1, 3)])
def test_stopping_criterion_and_readout_totals(n, p, thr, seed):
    rng = np.random.default_rng(seed)
    W = np.where(rng.random((n, n)) < p, rng.uniform(-1.0, 1.0, (n, n)), 0.0)
    np.fill_diagonal(W, 0.0)
    W_in = rng.uniform(-1, 1, (n, 2))
    W_out = rng.uniform(-1, 1, (3, n))
    o = _oracle(W, W_in, W_out, seed=seed)
    q = o.coarse_grain(thr)
    assert 1 <= q.N < n
    D = q.dense()
    off = ~np.eye(q.N, dtype=bool)
    assert np.all(D[off] <= thr)                         # no surviving pair would merge again
    np.testing.assert_allclose(q.W_out().sum(axis=1), W_out.sum(axis=1), rtol=1e-12, atol=1e-12)


def test_threshold_above_every_weight_is_identity():
    rng = np.random.default_rng(4)
    n = 90
    W = np.where(rng.random((n, n)) < 0.1, rng.uniform(-1.0, 1.0, (n, n)), 0.0)
    W_in = rng.uniform(-1, 1, (n, 1))
    W_out = rng.uniform(-1, 1, (1, n))
    o = _oracle(W, W_in, W_out)
    q = o.coarse_grain(1.0)
    assert q.N == n
    np.testing.assert_array_equal(q.dense(), o.dense())
    np.testing.assert_array_equal(q.W_in(), W_in)
    np.testing.assert_array_equal(q.W_out(), W_out)
    np.testing.assert_array_equal(q.state(), o.state())
