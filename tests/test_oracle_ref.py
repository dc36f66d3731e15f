"""Pin the CPU oracle against the reference's own sources (oracle/_ref) and
against closed forms.  CPU only."""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as po

needs_ref = pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built (no /root/reference)")


def test_gl_coeffs_closed_form(oracle):
    import mpmath as mp
    mp.mp.dps = 40
    for alpha in (0.3, 0.5, 0.8, 1.0):
        c = oracle.gl_coeffs(alpha, 600)
        assert c[0] == 1.0 and c[1] == pytest.approx(-alpha, rel=1e-15)
        for k in (2, 3, 10, 100, 511, 599):
            exact = float((-1) ** k * mp.binomial(alpha, k))
            assert c[k] == pytest.approx(exact, rel=1e-11, abs=1e-300)
        if alpha == 1.0:
            assert np.all(c[2:] == 0.0)


@needs_ref
def test_gl_coeffs_match_reference(oracle):
    for alpha, tmem in ((0.8, 512.0), (0.37, 1000.0)):
        params = np.array([1.0, 0.0, 0.0, 20.0, alpha, tmem, 0.05, 2.0])
        n = int(tmem)
        cref = np.empty(n)
        inp = np.zeros((1, 1))
        V = np.empty((1, 1)); S = np.empty((1, 1))
        po.ref().ref_neurons_run(3, po.dptr(params), 1.0, 1, 1, po.dptr(inp), po.dptr(V), po.dptr(S),
                                 None, po.dptr(cref), n)
        assert np.array_equal(oracle.gl_coeffs(alpha, n).view(np.uint64), cref.view(np.uint64))


NEURON_CASES = [
    ("lif_discrete", [0.0, 1.0, 0.1, 0.02], 1.0),
    ("lif_bio", [0.0, 1.0, 5.0, 0.9], 1.0),
    ("lif_bio", [0.0, 1.0, 5.0, 0.9], 0.25),
    ("flif_gl", [1.0, 0.0, 0.0, 20.0, 0.8, 64.0, 0.05, 2.0], 1.0),
    ("flif_gl", [1.0, -0.2, 0.1, 10.0, 0.6, 12.0, 0.0, 0.0], 0.5),
    ("flif_gl", [0.9, 0.0, 0.0, 15.0, 1.0, 40.0, 0.1, 3.0], 1.0),
    ("flif_caputo", [1.0, 0.05, 0.0, 1.0, 0.0, 2.0, 0.7, 2.0, 30.0], 1.0),
    ("flif_caputo", [1.0, 0.05, 0.0, 1.0, 0.0, 2.0, 0.5, 0.0, 20.0], 0.5),
    ("flif_diffusive", [1.0, 0.0, 0.0, 10.0, 0.6, 30.0, 0.95], 1.0),
    ("flif_diffusive", [1.0, 0.0, 0.0, 10.0, 1.0, 3.0, 0.95], 1.0),
    ("flif_diffusive", [1.0, 0.0, 0.0, 10.0, 0.0, 0.0, 0.95], 0.5),
]


def _oracle_isolated(neuron, params, dt, n, T, u):
    """Oracle reservoir with no recurrent edges and W_in = I (Din = n), so
    neuron i receives exactly u[t, i]."""
    cfg = po.make_config(n, din=n, dout=1, gain=1.0, dt=dt, neuron=neuron, params=params)
    rp = np.zeros(n + 1, np.int64)
    res = po.OracleReservoir(cfg, csr=(rp, np.zeros(0, np.int32), np.zeros(0)), W_in=np.eye(n))
    X, R, _ = res.run_record(u, T)
    return X, R


@needs_ref
@pytest.mark.parametrize("neuron,params,dt", NEURON_CASES)
def test_neuron_models_bit_exact_vs_reference(oracle, neuron, params, dt):
    rng = np.random.default_rng(11)
    n, T = 24, 300
    u = rng.uniform(-0.2, 0.6, size=(T, n))
    X, R = _oracle_isolated(neuron, params, dt, n, T, u)
    micro = int(round(1.0 / dt))
    inp = np.repeat(u, micro, axis=0).copy()            # ext is constant over micro-steps
    p = np.array(params, dtype=np.float64)
    V = np.empty((T * micro, n)); S = np.empty((T * micro, n))
    seeds = None
    if neuron == "flif_diffusive":
        seeds = (C.c_uint64 * n)(*[po.df_seed(i, p) for i in range(n)])
    po.ref().ref_neurons_run(po.NEURON[neuron], po.dptr(p), dt, n, T * micro, po.dptr(inp),
                             po.dptr(V), po.dptr(S), seeds, None, 0)
    Vm = V[micro - 1::micro]
    Sm = S[micro - 1::micro]
    assert np.array_equal(X.view(np.uint64), Vm.view(np.uint64))
    assert np.array_equal(R, (Sm != 0).astype(np.uint8))
    assert R.sum() > 0, "case should spike"


@needs_ref
@pytest.mark.parametrize("sparse", [1, 0])
def test_psc_heterogeneous_vs_reference(oracle, sparse):
    """Oracle LIF-discrete reservoir with PSC-het synapse; replay its
    membrane recursion from the reference synapse's row dots."""
    n, T = 40, 120
    rng = np.random.default_rng(5)
    W = np.where(rng.random((n, n)) < 0.2, rng.normal(0, 0.6, (n, n)), 0.0)
    np.fill_diagonal(W, 0.0)
    rp = np.zeros(n + 1, np.int64); ci = []; val = []
    for i in range(n):
        nz = np.nonzero(W[i])[0]
        ci += list(nz); val += list(W[i, nz]); rp[i + 1] = len(ci)
    Win = rng.uniform(-1, 1, (n, 1))
    params = [0.0, 0.5, 0.2, 0.05]
    syn = [0.5, 8.0]
    cfg = po.make_config(n, din=1, gain=1.0, dt=1.0, neuron="lif_discrete", params=params,
                         synapse="psc_heterogeneous", backend=0 if sparse else 1, syn_params=syn)
    u = rng.uniform(0, 0.5, T)
    po.srand(1234)
    res = po.OracleReservoir(cfg, csr=(rp, np.array(ci, np.int32), np.array(val)), W_in=Win)
    X, R, _ = res.run_record(u, T)
    last = np.vstack([np.zeros((1, n)), R[:-1].astype(np.float64)])
    rec = np.empty((T, n))
    po.srand(1234)
    sp = np.array(syn)
    po.ref().ref_psc_het_run(po.dptr(np.ascontiguousarray(W)), n, po.dptr(sp), sparse, 1.0, 1.0, T,
                             po.dptr(np.ascontiguousarray(last)), po.dptr(rec), None)
    V = np.zeros(n)
    for t in range(T):
        for i in range(n):
            ext = 1.0 * (Win[i, 0] * u[t])
            v = (1.0 - params[2]) * V[i] + (ext + rec[t, i]) + params[3]
            V[i] = params[0] if v > params[1] else v
        assert np.array_equal(V, X[t]), t
    assert R.sum() > 0


def _agile_loss(theta_p, dim, ctx):
    x = np.ctypeslib.as_array(theta_p, shape=(dim,))
    return float(np.sum((x - 0.3) ** 2 * np.arange(1, dim + 1)) + 0.1 * np.sin(7 * x[0]))


@needs_ref
def test_agile_bit_exact_vs_reference(oracle):
    dim = 5
    th0 = np.full(dim, 0.8); lo = np.zeros(dim); hi = np.ones(dim)
    opts = po.AgileOpts(1.3, 1e-3, 0.3, 20.0, 0.9, 1e-2, 60, 15, 1e-3, 1, 0xA91E, 0)
    f = po.LOSS_FN(_agile_loss)
    b1 = np.empty(dim); r1 = po.AgileReport()
    b2 = np.empty(dim); r2 = po.AgileReport()
    oracle.lib().or_agile_optimize(po.dptr(th0), po.dptr(lo), po.dptr(hi), dim, f, None,
                                   C.byref(opts), po.dptr(b1), C.byref(r1))
    po.ref().ref_agile(po.dptr(th0), po.dptr(lo), po.dptr(hi), dim, f, None, C.byref(opts),
                       po.dptr(b2), C.byref(r2))
    assert np.array_equal(b1, b2)
    assert (r1.best_loss, r1.best_iter, r1.total_evals, r1.total_iters, r1.exited_refinement) == \
           (r2.best_loss, r2.best_iter, r2.total_evals, r2.total_iters, r2.exited_refinement)
    assert r1.total_evals > 20
