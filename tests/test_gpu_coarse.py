"""Coarse-graining on the device (spires_coarse_grain -> spires_dev_coarse_grain)
against the oracle's restatement of coarse_grain_reservoir (reservoir.c:782-967):
the merged matrix, W_in, W_out and membrane state bit for bit, over thresholds
that merge nearly everything, some pairs, or nothing, on the random,
small-world and scale-free topologies with several inputs and outputs; the
coarse reservoir then runs bit-exact.  Negative and NaN thresholds (on which the
reference's merge loop never ends) are rejected."""
import time

import numpy as np
import pytest

import spires_amd as sa
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def _pair(n, topo, conn, din, dout, seed):
    kw = dict(din=din, dout=dout, rho=0.9, ei=0.6, gain=2.0, conn=conn, dt=1.0, conn_type=topo,
              neuron="lif_discrete", params=[0.0, 1.0, 0.1, 0.0])
    sa.srand(seed)
    r = sa.Reservoir(sa.make_config(n, **kw))
    po.srand(seed)
    o = po.OracleReservoir(po.make_config(n, **kw))
    return r, o


@pytest.mark.parametrize("n,topo,conn,din,dout,thr", [
    (300, "random", 0.1, 3, 2, 0.0),          # -> 5 neurons
    (300, "random", 0.1, 3, 2, 0.245),        # -> 167
    (400, "small_world", 0.05, 2, 1, 0.3),    # -> 94
    (400, "scale_free", 0.02, 1, 3, 0.3),     # -> 180
    (250, "random", 0.2, 1, 1, 5.0),          # nothing merges
])
def test_coarse_grain_bit_exact(n, topo, conn, din, dout, thr):
    print(f"n={n} {topo} conn={conn} din={din} dout={dout} thr={thr}", flush=True)
    r, o = _pair(n, topo, conn, din, dout, seed=3)
    u = np.random.default_rng(5).uniform(0, 0.5, (40, din))
    r.run_record(u)
    o.run_record(u, 40)
    rq, oq = r.coarse_grain(thr), o.coarse_grain(thr)
    print(f"  {n} -> {rq.N} neurons", flush=True)
    assert rq.N == oq.N
    if thr >= 5.0:
        assert rq.N == n   # no weight exceeds the threshold: nothing merges
    else:
        assert rq.N < n
    assert np.array_equal(_bits(rq.weights()), _bits(oq.dense()))
    assert np.array_equal(_bits(rq.W_in()), _bits(oq.W_in()))
    assert np.array_equal(_bits(rq.W_out()), _bits(oq.W_out()))
    assert np.array_equal(_bits(rq.state()), _bits(oq.state()))
    X, R, _ = rq.run_record(u)
    Xo, Ro, _ = oq.run_record(u, 40)
    assert np.array_equal(R, Ro) and np.array_equal(_bits(X), _bits(Xo))


def test_coarse_grain_rejects_thresholds_that_never_end():
    r, _ = _pair(64, "random", 0.1, 1, 1, seed=1)
    for thr in (-0.1, float("nan")):
        with pytest.raises(sa.SpiresError):
            r.coarse_grain(thr)


@pytest.mark.parametrize("thr", [0.02, 0.083, 0.2])
def test_coarse_grain_time_n2048(thr):
    """N=2048 random p=0.05 (largest weight 0.086): thr=0.02 merges down to
    8 neurons, 0.083 to 475, 0.2 nothing (one scan pass).  Device passes
    against the oracle's host loop (timed; both bit-identical)."""
    r, o = _pair(2048, "random", 0.05, 1, 1, seed=7)
    u = np.random.default_rng(1).uniform(0, 0.5, (20, 1))
    r.run_record(u)
    o.run_record(u, 20)
    t0 = time.perf_counter()
    rq = r.coarse_grain(thr)
    t_dev = time.perf_counter() - t0
    t0 = time.perf_counter()
    oq = o.coarse_grain(thr)
    t_host = time.perf_counter() - t0
    print(f"N=2048 thr={thr} -> {rq.N}: device {t_dev:.3f} s, oracle host loop {t_host:.3f} s", flush=True)
    assert rq.N == oq.N
    assert np.array_equal(_bits(rq.weights()), _bits(oq.dense()))
