"""The product-wave kernel with a window of product chunks in LDS (kernel 3:
chosen when whole product rows for the workgroup's neurons do not fit), against
the oracle: rasters and membrane traces bit for bit while the history fills
(the number of chunks per step grows past the window) and after it is full,
across launches that continue a run, at 32 and 64 neurons per workgroup."""
import math

import numpy as np
import pytest

import spires_amd as sa
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _er(n, conn, seed):
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for i in range(n):
        c = np.nonzero(rng.random(n) < conn)[0]
        c = c[c != i]
        rows.append(np.full(len(c), i))
        cols.append(c)
    rows, cols = np.concatenate(rows), np.concatenate(cols)
    rp = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=n), out=rp[1:])
    val = rng.standard_normal(len(cols)) / math.sqrt(n * conn)
    return rp, cols.astype(np.int32), val, rng.uniform(-1, 1, (n, 1))


def _pair(n, conn, L, seed=31):
    params = [1.0, 0.0, 0.0, 20.0, 0.7, float(L), 0.05, 2.0]
    rp, ci, val, Wi = _er(n, conn, seed)
    kw = dict(din=1, dout=1, gain=3.0, neuron="flif_gl", params=params)
    r = sa.Reservoir(sa.make_config(n, **kw), csr=(rp, ci, val), W_in=Wi)
    o = po.OracleReservoir(po.make_config(n, **kw), csr=(rp, ci, val), W_in=Wi)
    return r, o


@pytest.mark.parametrize("n,conn,L,T,nb", [
    (5000, 0.004, 600, 700, 32),     # 19 chunks per step once full, window of 16
    (5000, 0.004, 1024, 1100, 32),   # 32 chunks, window of 16
    (9000, 0.002, 300, 350, 64),     # 10 chunks, window of 8
])
def test_windowed_products_bit_exact(n, conn, L, T, nb):
    print(f"n={n} L={L} T={T}", flush=True)
    r, o = _pair(n, conn, L)
    plan = r.kernel_plan()
    print(plan, flush=True)
    assert plan["kernel"] == 3 and plan["nb"] == nb
    u = np.random.default_rng(32).uniform(0.0, 1.0, (T, 1))
    X, R, _ = r.run_record(u)
    Xo, Ro, _ = o.run_record(u, T)
    assert Ro.sum() > 0
    bad = np.argwhere(R != Ro)
    assert np.array_equal(R, Ro), f"first raster mismatch at (step, neuron) {bad[:3]}"
    assert np.array_equal(X.view(np.uint64), Xo.view(np.uint64))
    r.close()


def test_windowed_products_continue_across_launches():
    """Three launches (300 + 400 + 300 steps) equal one oracle run: each launch
    refills the window for its first step from the history ring."""
    n, conn, L = 5000, 0.004, 600
    r, o = _pair(n, conn, L, seed=33)
    assert r.kernel_plan()["kernel"] == 3
    T = 1000
    u = np.random.default_rng(34).uniform(0.0, 1.0, (T, 1))
    r.upload_inputs(u)
    r.reserve_records(T, sa.REC_STATE | sa.REC_RASTER)
    for t0, k in ((0, 300), (300, 400), (700, 300)):
        r.run_resident(t0, k, sa.REC_STATE | sa.REC_RASTER | sa.REC_APPEND)
    X, R, _ = r.read_records(0, T, states=True, raster=True)
    Xo, Ro, _ = o.run_record(u, T)
    assert np.array_equal(R, Ro)
    assert np.array_equal(X.view(np.uint64), Xo.view(np.uint64))
    r.close()
