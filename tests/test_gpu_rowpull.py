"""Row-pull variants of the product-wave kernel against the oracle: the edge
list in LDS (sparse rows), column bitmaps with the weights in LDS (dense
rows), bitmaps with the weights read from HBM (dense rows next to a long
history), and the edge list read from HBM in batches (N=4452, where neither
fits next to the history).  Each case checks the plan picked the variant it names, then
bit-exact rasters and membrane traces."""
import math

import numpy as np
import pytest

import spires_amd as sa
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

CASES = [  # neurons, connectivity, history length, expected (row_pull, weights_lds, neurons per workgroup)
    (2048, 0.005, 128, (0, 1, 16)),
    (2048, 0.1, 128, (1, 1, 16)),
    (2048, 0.2, 900, (1, 0, 16)),
    (5000, 0.004, 128, (0, 1, 32)),   # 16 per workgroup would need more workgroups than CUs
    (5000, 0.04, 128, (1, 1, 32)),
    (4452, 0.05, 512, (0, 0, 32)),    # neither the edge list nor the bitmaps fit next to the history
]


def _er(n, conn, seed):
    rng = np.random.default_rng(seed)
    mask = rng.random((n, n)) < conn
    np.fill_diagonal(mask, False)
    rows, cols = np.nonzero(mask)
    rp = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=n), out=rp[1:])
    val = rng.standard_normal(len(cols)) / math.sqrt(n * conn)
    return rp, cols.astype(np.int32), val, rng.uniform(-1, 1, (n, 1))


@pytest.mark.parametrize("n,conn,L,expect", CASES)
def test_row_pull_variants_bit_exact(n, conn, L, expect):
    params = [1.0, 0.0, 0.0, 20.0, 0.7, float(L), 0.05, 2.0]
    rp, ci, val, Wi = _er(n, conn, 21)
    kw = dict(din=1, dout=1, gain=3.0, neuron="flif_gl", params=params)
    r = sa.Reservoir(sa.make_config(n, **kw), csr=(rp, ci, val), W_in=Wi)
    plan = r.kernel_plan()
    print(conn, L, plan, flush=True)
    assert plan["kernel"] == 2
    assert (plan["row_pull"], plan["weights_lds"], plan["nb"]) == expect
    o = po.OracleReservoir(po.make_config(n, **kw), csr=(rp, ci, val), W_in=Wi)
    T = 40
    u = np.random.default_rng(22).uniform(0.0, 1.0, (T, 1))
    X, R, _ = r.run_record(u)
    Xo, Ro, _ = o.run_record(u, T)
    assert Ro.sum() > 0
    assert np.array_equal(R, Ro)
    assert np.array_equal(X.view(np.uint64), Xo.view(np.uint64))
    r.close()
