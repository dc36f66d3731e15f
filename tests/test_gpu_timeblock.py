"""Time blocks replayed from a HIP graph (spires_ext_timeblock_*): every
replay continues each reservoir from its own device-side step clock, and the
records and states equal those of eager batched runs of the same steps, bit
for bit -- across several replays, with eager work on some reservoirs in
between, and against the oracle."""
import numpy as np
import pytest

import spires_amd as sa
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

SPECS = [  # n, T_mem, alpha, topology, connectivity, gain, seed
    (200, 48, 0.8, "random", 0.1, 2.0, 3),
    (333, 100, 0.6, "scale_free", 0.02, 1.5, 4),
    (96, 200, 0.9, "small_world", 0.08, 2.5, 5),
]
T_ALL = 160
FLAGS = sa.REC_STATE | sa.REC_RASTER


def _kw(spec):
    n, tm, alpha, topo, conn, gain, seed = spec
    return dict(din=1, dout=2, rho=0.9, ei=0.6, gain=gain, conn=conn, dt=1.0, conn_type=topo, neuron="flif_gl",
                params=[1.0, 0.0, 0.0, 20.0, alpha, float(tm), 0.05, 2.0])


def _make(spec, u):
    sa.srand(spec[-1])
    r = sa.Reservoir(sa.make_config(spec[0], **_kw(spec)))
    r.upload_inputs(u)
    r.reserve_records(T_ALL, FLAGS)
    return r


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def _same(a, b, rows):
    Xa, Ra, _ = a.read_records(0, rows, states=True, raster=True)
    Xb, Rb, _ = b.read_records(0, rows, states=True, raster=True)
    return np.array_equal(Ra, Rb) and np.array_equal(_bits(Xa), _bits(Xb))


@pytest.mark.parametrize("one", [False, True], ids=["packed", "one_launch"])
def test_timeblock_replays_match_eager_runs_and_oracle(one):
    """Packed product-wave launches, or every reservoir in one launch of
    64-neuron one-wave slices (TB_ONE_LAUNCH): both equal the eager batched runs
    (the product-wave kernel) and the oracle bit for bit."""
    print(f"case: time block, {'one launch' if one else 'packed launches'}", flush=True)
    us = [np.random.default_rng(40 + k).uniform(0, 0.5, (T_ALL, 1)) for k in range(len(SPECS))]
    A = [_make(s, u) for s, u in zip(SPECS, us)]     # graph
    B = [_make(s, u) for s, u in zip(SPECS, us)]     # eager
    tb = sa.TimeBlock(A, 20, FLAGS | (sa.TB_ONE_LAUNCH if one else 0))
    assert tb.launches == 1
    tb.run(2)
    sa.run_resident_batch(B, 0, 40, FLAGS | sa.REC_APPEND)
    for a, b in zip(A, B):
        assert _same(a, b, 40)
    # eager work on one reservoir between replays: it runs ahead by 10 steps
    A[0].run_resident(40, 10, FLAGS | sa.REC_APPEND)
    B[0].run_resident(40, 10, FLAGS | sa.REC_APPEND)
    tb.run(1)
    B[0].run_resident(50, 20, FLAGS | sa.REC_APPEND)
    sa.run_resident_batch(B[1:], 40, 20, FLAGS | sa.REC_APPEND)
    assert _same(A[0], B[0], 70)
    for a, b in zip(A[1:], B[1:]):
        assert _same(a, b, 60)
    tb.run(3)                                          # odd and even replay counts both swap spike buffers
    B[0].run_resident(70, 60, FLAGS | sa.REC_APPEND)
    sa.run_resident_batch(B[1:], 60, 60, FLAGS | sa.REC_APPEND)
    assert _same(A[0], B[0], 130)
    for a, b in zip(A[1:], B[1:]):
        assert _same(a, b, 120)
        assert np.array_equal(_bits(a.state()), _bits(b.state()))
    # the oracle over the same 120 steps
    spec, u = SPECS[1], us[1]
    po.srand(spec[-1])
    o = po.OracleReservoir(po.make_config(spec[0], **_kw(spec)))
    Xo, Ro, _ = o.run_record(u[:120], 120)
    X, R, _ = A[1].read_records(0, 120, states=True, raster=True)
    assert Ro.sum() > 0 and np.array_equal(R, Ro) and np.array_equal(_bits(X), _bits(Xo))
    # a run past the reserved records is refused; so is one after the records move
    with pytest.raises(sa.SpiresError):
        tb.run(5)
    A[2].reserve_records(2 * T_ALL, FLAGS)
    with pytest.raises(sa.SpiresError):
        tb.run(1)
    tb.close()
    for r in A + B:
        r.close()


def test_one_launch_time_block_long_histories_and_psc():
    """The one-launch kernel on reservoirs the product-wave batch does not take
    together (T_mem 1024 beside 64, a PSC synapse) against single-reservoir
    runs of the product-wave kernels and the oracle, bit for bit."""
    print("case: one-launch time block, long histories and a PSC synapse", flush=True)
    specs = [(700, 1024, 0.7, "random", 0.05, 2.0, 7), (130, 64, 0.9, "small_world", 0.1, 1.5, 8)]
    T = 1500
    us = [np.random.default_rng(50 + k).uniform(0, 0.5, (T, 1)) for k in range(2)]

    def make(spec, u, synapse="simple"):
        sa.srand(spec[-1])
        kw = _kw(spec)
        if synapse != "simple":
            kw.update(synapse=synapse, syn_params=[0.5, 8.0])
        r = sa.Reservoir(sa.make_config(spec[0], **kw))
        r.upload_inputs(u)
        r.reserve_records(T, FLAGS)
        return r

    A = [make(specs[0], us[0]), make(specs[1], us[1], "psc_heterogeneous")]
    B = [make(specs[0], us[0]), make(specs[1], us[1], "psc_heterogeneous")]
    tb = sa.TimeBlock(A, 400, FLAGS | sa.TB_ONE_LAUNCH)
    assert tb.launches == 1
    tb.run(1)
    # a single-reservoir run between replays takes more exchange words than the
    # block's 64-neuron slices (the block must follow the reallocated words)
    A[1].run_resident(400, 300, FLAGS | sa.REC_APPEND)
    tb.run(2)
    B[0].run_resident(0, 1200, FLAGS | sa.REC_APPEND)
    B[1].run_resident(0, 1500, FLAGS | sa.REC_APPEND)
    assert _same(A[0], B[0], 1200) and _same(A[1], B[1], 1500)
    for a, b in zip(A, B):
        assert np.array_equal(_bits(a.state()), _bits(b.state()))
    spec, u = specs[0], us[0]
    po.srand(spec[-1])
    o = po.OracleReservoir(po.make_config(spec[0], **_kw(spec)))
    Xo, Ro, _ = o.run_record(u[:600], 600)
    X, R, _ = A[0].read_records(0, 600, states=True, raster=True)
    assert Ro.sum() > 0 and np.array_equal(R, Ro) and np.array_equal(_bits(X), _bits(Xo))
    tb.close()
    for r in A + B:
        r.close()


def test_autotuned_block_runner_matches_eager_runs():
    """spires_amd.autotune_blocks times one block per (tile shape, K) candidate,
    each advancing every reservoir, and BlockRunner then continues in the shape
    it picked: the records of all these blocks equal eager batched runs of the
    same steps bit for bit."""
    print("case: autotuned block runner, 2 shapes x K in (8, 16), then 3 blocks", flush=True)
    us = [np.random.default_rng(60 + k).uniform(0, 0.5, (T_ALL, 1)) for k in range(len(SPECS))]
    A = [_make(s, u) for s, u in zip(SPECS, us)]     # block runner
    B = [_make(s, u) for s, u in zip(SPECS, us)]     # eager
    shape, K, table, t = sa.autotune_blocks(A, 0, Ks=(8, 16), flags=FLAGS)
    assert set(table) == {(s, k) for s in sa.BlockRunner.SHAPES for k in (8, 16)}
    assert t == 48 and shape in sa.BlockRunner.SHAPES and K in (8, 16)
    br = sa.BlockRunner(A, K, shape, FLAGS)
    br.run(t, 3)
    br.close()
    sa.synchronize(A)
    rows = t + 3 * K
    sa.run_resident_batch(B, 0, rows, FLAGS | sa.REC_APPEND)
    for a, b in zip(A, B):
        assert _same(a, b, rows)
    for r in A + B:
        r.close()
