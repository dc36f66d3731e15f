"""C-ABI boundary tests that need no GPU: the library loads, exports every
declared symbol, validates arguments like the reference, builds the same
weights as the oracle from the same glibc rand() stream, and runs the same
AGILE trajectory as the reference's own agile.c."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import spires_amd as sa
from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return set(re.findall(r"\b(spires_[a-z0-9_]+)\s*\(", txt))


def test_exports_every_declared_symbol():
    L = sa.lib()
    declared = _declared("spires.h") | _declared("spires_ext.h") | _declared("spires_dev.h")
    assert len(declared) >= 21 + 15
    missing = [s for s in sorted(declared) if not hasattr(L, s)]
    assert not missing, missing
    # the 21 functions of the reference header (include/spires.h:131-267)
    ref21 = {"spires_reservoir_create", "spires_reservoir_destroy", "spires_reservoir_reset", "spires_step",
             "spires_run", "spires_train_online", "spires_train_ridge", "spires_train_rls",
             "spires_coarse_grain", "spires_copy_reservoir_state", "spires_read_reservoir_state",
             "spires_read_spike_state", "spires_copy_weights", "spires_read_weights",
             "spires_compute_output", "spires_copy_readout", "spires_read_readout", "spires_num_neurons",
             "spires_num_inputs", "spires_num_outputs", "spires_optimize"}
    assert ref21 <= _declared("spires.h")


def test_struct_layout_matches_header():
    assert C.sizeof(sa.Config) == 96
    assert sa.Config.neuron_params.offset == 72 and sa.Config.synapse_params.offset == 88
    assert C.sizeof(sa.OptBudget) == 40 and C.sizeof(sa.OptScore) == 24
    assert C.sizeof(sa.OptResult) == 96 + 32


def test_argument_validation_without_gpu():
    L = sa.lib()
    h = C.c_void_p()
    assert L.spires_reservoir_create(None, C.byref(h)) == 1
    for dt in (0.0, -1.0, 1.5, 0.3):          # spires_api.c:29-38
        cfg = sa.make_config(8, dt=dt, params=[1, 0, 0, 20, 0.8, 10, 0, 2])
        assert L.spires_reservoir_create(C.byref(cfg[0]), C.byref(h)) == 1
    cfg = sa.make_config(8, dt=0.25, neuron=7, params=[1, 0, 0, 20])
    assert L.spires_reservoir_create(C.byref(cfg[0]), C.byref(h)) == 1
    # NULL handles
    assert L.spires_step(None, None) == 1
    assert L.spires_reservoir_reset(None) == 1
    assert L.spires_read_reservoir_state(None, None) == 1
    assert not L.spires_run(None, None, 3)
    assert not L.spires_copy_reservoir_state(None)
    assert L.spires_num_neurons(None) == 0
    L.spires_reservoir_destroy(None)          # no-op
    assert L.spires_optimize(None, None, 0, None, None, None, None, 0) == -1


TOPOS = [("random", 0.1), ("small_world", 20 / 199), ("scale_free", 3 / 199)]
SYNS = [("simple", None, "sparse"), ("psc_heterogeneous", [0.5, 8.0], "sparse"),
        ("psc_heterogeneous", [0.5, 8.0], "dense"), ("psc_homogeneous", [4.0], "dense")]


@pytest.mark.parametrize("topo,conn", TOPOS)
@pytest.mark.parametrize("syn,sp,backend", SYNS)
def test_build_weights_match_oracle(oracle, topo, conn, syn, sp, backend):
    n, din, dout = 200, 2, 3
    params = [1.0, 0.0, 0.0, 20.0, 0.8, 32.0, 0.05, 2.0]
    kw = dict(din=din, dout=dout, rho=0.9, ei=0.8, conn=conn, conn_type=topo, params=params,
              synapse=syn, syn_params=sp)
    sa.srand(77)
    rp, ci, val, Wi, Wo = sa.build_weights(sa.make_config(n, backend=backend, **kw))
    po.srand(77)
    ores = po.OracleReservoir(po.make_config(n, backend=sa.BACKEND[backend], **kw))
    orp, oci, oval = (ores.csr() if backend == "sparse" else (None, None, None))
    if backend == "sparse":
        assert np.array_equal(rp, orp) and np.array_equal(ci, oci)
        assert np.array_equal(val.view(np.uint64), oval.view(np.uint64))
    else:
        W = np.zeros((n, n))
        for i in range(n):
            W[i, ci[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
        assert np.array_equal(W.view(np.uint64), ores.dense().view(np.uint64))
    assert np.array_equal(Wi.view(np.uint64), ores.W_in().view(np.uint64))
    assert np.array_equal(Wo.view(np.uint64), ores.W_out().view(np.uint64))


def test_topology_invariants():
    n = 300
    # Watts-Strogatz: K/2 forward edges per node survive rewiring (reservoir.c:399-443)
    sa.srand(3)
    rp, ci, val, _, _ = sa.build_weights(sa.make_config(n, conn=20 / (n - 1), conn_type="small_world",
                                                        params=[1, 0, 0, 20, 0.8, 8, 0, 2]))
    deg = np.diff(rp)
    assert len(val) <= n * 10 and len(val) >= n * 10 - 5          # zero-weight draws may drop an edge
    assert np.all(deg <= 10 + 10)
    # Barabasi-Albert: m0(m0-1)/2 + m(n-m0) undirected edges, each oriented once (reservoir.c:445-536)
    m = 5
    sa.srand(4)
    rp, ci, val, _, _ = sa.build_weights(sa.make_config(n, conn=m / (n - 1), conn_type="scale_free",
                                                        params=[1, 0, 0, 20, 0.8, 8, 0, 2]))
    assert abs(len(val) - (m * (m - 1) // 2 + m * (n - m))) <= 3
    rows = np.repeat(np.arange(n), np.diff(rp))
    assert not np.any(rows == ci)                                    # no self loops
    pairs = set(zip(rows.tolist(), ci.tolist()))
    assert not any((j, i) in pairs for i, j in pairs)                # one orientation per edge
    # Erdos-Renyi density ~ p
    sa.srand(5)
    rp, ci, val, _, _ = sa.build_weights(sa.make_config(n, conn=0.1, conn_type="random",
                                                        params=[1, 0, 0, 20, 0.8, 8, 0, 2]))
    assert abs(len(val) / (n * (n - 1)) - 0.1) < 0.01


def test_spectral_rescale(oracle):
    """After creation the power-iteration estimate equals the requested rho."""
    sa.srand(9)
    rp, ci, val, _, _ = sa.build_weights(sa.make_config(150, rho=0.95, conn=0.1,
                                                        params=[1, 0, 0, 20, 0.8, 8, 0, 2]))
    assert oracle.csr_spectral_radius(rp, ci, val) == pytest.approx(0.95, rel=1e-5)


def _loss(theta_p, dim, ctx):
    x = np.ctypeslib.as_array(theta_p, shape=(dim,))
    return float(np.sum((x - 0.3) ** 2 * np.arange(1, dim + 1)) + 0.1 * np.sin(7 * x[0]))


@pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built")
def test_agile_matches_reference():
    dim = 6
    th0 = np.full(dim, 0.8); lo = np.zeros(dim); hi = np.ones(dim)
    f = sa.LOSS_FN(_loss)
    ours = np.empty(dim)
    opts = np.array([1.3, 1e-3, 0.3, 20.0, 0.9, 1e-2, 60, 15, 1e-3, 1, 0xA91E], dtype=np.float64)
    assert sa.lib().spires_ext_agile_optimize(sa._dp(th0), sa._dp(lo), sa._dp(hi), dim, f, None, sa._dp(opts),
                                              sa._dp(ours)) == 0
    ref = np.empty(dim); rep = po.AgileReport()
    o = po.AgileOpts(1.3, 1e-3, 0.3, 20.0, 0.9, 1e-2, 60, 15, 1e-3, 1, 0xA91E, 0)
    po.ref().ref_agile(po.dptr(th0), po.dptr(lo), po.dptr(hi), dim, po.LOSS_FN(_loss), None, C.byref(o),
                       po.dptr(ref), C.byref(rep))
    assert np.array_equal(ours, ref)
