"""Batched runs (spires_ext_run_resident_batch): reservoirs of different size,
history length, alpha and topology advance together in one device launch, and
each one's rasters and membrane traces stay bit-identical to the oracle."""
import numpy as np
import pytest

import spires_amd as sa
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _gl(T_mem, alpha):
    return [1.0, 0.0, 0.0, 20.0, alpha, float(T_mem), 0.05, 2.0]


SPECS = [  # n, T_mem, alpha, topology, connectivity, gain, seed
    (200, 48, 0.8, "random", 0.1, 2.0, 3),
    (333, 100, 0.6, "scale_free", 0.02, 1.5, 4),
    (96, 200, 0.9, "small_world", 0.08, 2.5, 5),
]


def _pair(spec):
    n, tm, alpha, topo, conn, gain, seed = spec
    kw = dict(din=1, dout=2, rho=0.9, ei=0.6, gain=gain, conn=conn, dt=1.0, conn_type=topo,
              neuron="flif_gl", params=_gl(tm, alpha))
    sa.srand(seed)
    r = sa.Reservoir(sa.make_config(n, **kw))
    po.srand(seed)
    o = po.OracleReservoir(po.make_config(n, **kw))
    return r, o


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def test_batched_reservoirs_match_oracle_across_runs():
    pairs = [_pair(s) for s in SPECS]
    rs = [p[0] for p in pairs]
    assert sa.batch_workgroups(rs) == sum((s[0] + 31) // 32 for s in SPECS)
    T1, T2 = 70, 45
    us = [np.random.default_rng(20 + i).uniform(0, 0.5, (T1 + T2, 1)) for i in range(len(rs))]
    for r, u in zip(rs, us):
        r.upload_inputs(u)
        r.reserve_records(T1, sa.REC_STATE | sa.REC_RASTER)
    sa.run_resident_batch(rs, 0, T1, sa.REC_STATE | sa.REC_RASTER)
    for (r, o), u in zip(pairs, us):
        X, R, _ = r.read_records(0, T1, states=True, raster=True)
        Xo, Ro, _ = o.run_record(u[:T1], T1)
        assert Ro.sum() > 0
        assert np.array_equal(R, Ro) and np.array_equal(_bits(X), _bits(Xo))
    sa.run_resident_batch(rs, T1, T2, sa.REC_STATE | sa.REC_RASTER)   # continues the same state
    for (r, o), u in zip(pairs, us):
        X, R, _ = r.read_records(0, T2, states=True, raster=True)
        Xo, Ro, _ = o.run_record(u[T1:], T2)
        assert np.array_equal(R, Ro) and np.array_equal(_bits(X), _bits(Xo))


def test_batched_equals_separate_runs():
    a = [_pair(s)[0] for s in SPECS]
    b = [_pair(s)[0] for s in SPECS]
    T = 60
    for i, (ra, rb) in enumerate(zip(a, b)):
        u = np.random.default_rng(40 + i).uniform(0, 0.5, (T, 1))
        for r in (ra, rb):
            r.upload_inputs(u)
            r.reserve_records(T, sa.REC_STATE)
    sa.run_resident_batch(a, 0, T, sa.REC_STATE)
    for ra, rb in zip(a, b):
        rb.run_resident(0, T, sa.REC_STATE)
        Xa, _, _ = ra.read_records(0, T, states=True)
        Xb, _, _ = rb.read_records(0, T, states=True)
        assert np.array_equal(_bits(Xa), _bits(Xb))


def test_batch_rejects_mixed_history_classes():
    r1 = _pair(SPECS[0])[0]
    r2 = _pair((64, 900, 0.8, "random", 0.1, 2.0, 9))[0]      # L=900: 16 neurons per workgroup
    assert sa.batch_workgroups([r1, r2]) == 0
    for r in (r1, r2):
        r.upload_inputs(np.zeros((5, 1)))
        r.reserve_records(5, sa.REC_STATE)
    with pytest.raises(RuntimeError):
        sa.run_resident_batch([r1, r2], 0, 5)
