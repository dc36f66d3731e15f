"""The caller's glibc rand() sequence is left where the caller put it by every
GPU call: the reference's construction draws the weights from it
(reservoir.c:307-571), so after srand(s) the n-th reservoir created must get
the reference's weights whatever GPU work ran in between.  The HIP runtime's
start-up and rocBLAS's first handle draw from rand(); the device layer's
entry points hold them off the caller's sequence.  Run in a fresh process
(first GPU use in this file's process is not guaranteed, so the check also
covers the readout's first use on a new handle size)."""
import ctypes as C

import numpy as np
import pytest

import spires_amd as sa
from spires_amd import workloads

pytestmark = pytest.mark.gpu

libc = C.CDLL("libc.so.6")
PARAMS = [1.0, 0.0, 0.0, 20.0, 0.55, 16.0, 0.05, 2.0]


def _unchanged(fn):
    libc.srand(7)
    a = [libc.rand() for _ in range(4)]
    libc.srand(7)
    fn()
    b = [libc.rand() for _ in range(4)]
    return a == b


def test_gpu_calls_leave_the_rand_sequence_alone():
    print("case: rand() sequence across device calls", flush=True)
    s = workloads.mackey_glass(81, seed=0)
    x, y = np.ascontiguousarray(s[:80].reshape(80, 1)), np.ascontiguousarray(s[1:81].reshape(80, 1))
    n = 200
    csr = (np.arange(n + 1, dtype=np.int64), ((np.arange(n) + 1) % n).astype(np.int32), np.full(n, 0.5))
    held = {}
    assert _unchanged(lambda: held.setdefault("r", sa.Reservoir(sa.make_config(n, neuron="flif_gl", params=PARAMS),
                                                                 csr=csr, W_in=np.ones((n, 1)))))
    r = held["r"]
    assert _unchanged(lambda: r.run_record(x))
    assert _unchanged(lambda: r.train_ridge(x[:48], y[:48], 1e-2))            # rocBLAS path (n % 64 != 0)
    assert _unchanged(lambda: r.train_rls(x[:30], y[:30], 1.0, 1.0))
    assert _unchanged(lambda: r.reset())
    assert _unchanged(lambda: r.run_record(x[48:], outputs=True))
    assert _unchanged(lambda: sa.Surrogate(7, 16, seed=1).close())
    assert _unchanged(lambda: r.close())


def test_second_reservoir_gets_the_reference_weights_after_gpu_work():
    """srand(s); create A; train A; create B  ==  srand(s); draw A's weights; draw B's weights."""
    print("case: second reservoir's weights after GPU work", flush=True)
    kw = dict(din=1, dout=1, rho=0.9, ei=0.8, gain=1.0, conn=0.2, conn_type="random", neuron="flif_gl", params=PARAMS)
    x = np.random.default_rng(1).uniform(0, 1, (60, 1))
    sa.srand(31)
    a = sa.Reservoir(sa.make_config(150, **kw))
    a.train_ridge(x, x, 1e-2)
    a.run_record(x, outputs=True)
    b = sa.Reservoir(sa.make_config(170, **kw))
    sa.srand(31)
    sa.build_weights(sa.make_config(150, **kw))
    rp, ci, val, Wi, Wo = sa.build_weights(sa.make_config(170, **kw))
    rp2, ci2, val2 = b.csr()
    assert np.array_equal(rp, rp2) and np.array_equal(ci, ci2) and np.array_equal(val, val2)
    assert np.array_equal(Wi, b.W_in()) and np.array_equal(Wo, b.W_out())
    a.close()
    b.close()
