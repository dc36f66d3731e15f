"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same seeded inputs.  Spike rasters and membrane traces are compared bit for
bit (FP64 exact mode); readouts built from BLAS-order-dependent reductions
are compared at the tolerance written in each test."""
import numpy as np
import pytest

import spires_amd as sa
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

GL = [1.0, 0.0, 0.0, 20.0, 0.8, 48.0, 0.05, 2.0]
CASES = [
    # neuron, params, dt, synapse, syn_params, backend, topology, conn, gain
    ("flif_gl", GL, 1.0, "simple", None, "sparse", "small_world", 0.05, 2.0),
    ("flif_gl", GL, 1.0, "simple", None, "sparse", "random", 0.1, 1.0),
    ("flif_gl", GL, 1.0, "simple", None, "sparse", "scale_free", 0.02, 1.5),
    ("flif_gl", [1.0, -0.1, 0.05, 10.0, 0.6, 12.0, 0.02, 0.0], 0.5, "simple", None, "sparse", "random", 0.1, 1.3),
    ("flif_gl", GL, 1.0, "simple", None, "dense", "random", 0.1, 2.0),
    ("flif_gl", GL, 1.0, "psc_homogeneous", [3.0], "sparse", "random", 0.1, 2.0),
    ("flif_gl", GL, 1.0, "psc_homogeneous", [3.0], "dense", "random", 0.1, 2.0),
    ("flif_gl", GL, 1.0, "psc_heterogeneous", [0.5, 8.0], "sparse", "random", 0.1, 2.0),
    ("flif_gl", GL, 1.0, "psc_heterogeneous", [0.5, 8.0], "dense", "random", 0.1, 2.0),
    ("flif_gl", GL, 1.0, "fractional", [0.7, 3.0, 20.0], "sparse", "random", 0.1, 2.0),
    ("flif_gl", GL, 1.0, "fractional_multiexp", [0.6, 2.0, 0.5, 30.0, 6], "sparse", "random", 0.1, 2.0),
    ("lif_discrete", [0.0, 1.0, 0.1, 0.0], 1.0, "simple", None, "sparse", "random", 0.1, 1.0),
    ("lif_bio", [0.0, 1.0, 5.0, 0.9], 0.25, "simple", None, "sparse", "random", 0.1, 1.5),
    ("flif_caputo", [1.0, 0.05, 0.0, 1.0, 0.0, 2.0, 0.7, 2.0, 30.0], 1.0, "simple", None, "sparse", "random",
     0.1, 2.0),
    ("flif_diffusive", [1.0, 0.0, 0.0, 10.0, 1.0, 3.0, 0.3], 1.0, "simple", None, "sparse", "random", 0.1, 3.0),
]


def _pair(case, n=160, din=1, dout=2, seed=1):
    neuron, params, dt, syn, sp, backend, topo, conn, gain = case
    kw = dict(din=din, dout=dout, rho=0.9, ei=0.6, gain=gain, conn=conn, dt=dt, conn_type=topo,
              neuron=neuron, params=params, synapse=syn, syn_params=sp)
    cfg = sa.make_config(n, backend=backend, **kw)
    cfgo = po.make_config(n, backend=sa.BACKEND[backend], **kw)
    sa.srand(seed)
    r = sa.Reservoir(cfg)
    po.srand(seed)
    o = po.OracleReservoir(cfgo)
    return r, o


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-{c[3]}-{c[5]}-{c[6]}-dt{c[2]}")
def test_run_bit_exact(case):
    print("case", case, flush=True)
    r, o = _pair(case)
    T = 250
    u = np.random.default_rng(7).uniform(0.0, 0.5, (T, 1))
    X, R, Y = r.run_record(u, states=True, raster=True, outputs=True)
    Xo, Ro, Yo = o.run_record(u, T, states=True, raster=True, outputs=True)
    assert Ro.sum() > 0, "case should spike"
    assert np.array_equal(R, Ro), f"raster mismatch at {np.argwhere(R != Ro)[:5]}"
    assert np.array_equal(_bits(X), _bits(Xo))
    # readout y = W_out V: tree reduction on the GPU vs sequential on the CPU
    np.testing.assert_allclose(Y, Yo, rtol=1e-12, atol=1e-12 * np.abs(Yo).max())


def test_diffusive_stochastic_refractory():
    """alpha < 1 draws refractory lengths with pow() on the device; draws are
    seeded identically, so rasters agree except where a device pow differs
    from glibc's in the last bit at a ceil() boundary (statistical parity)."""
    case = ("flif_diffusive", [1.0, 0.0, 0.0, 10.0, 0.6, 30.0, 0.3], 1.0, "simple", None, "sparse", "random",
            0.1, 3.0)
    r, o = _pair(case, n=256)
    T = 400
    u = np.random.default_rng(3).uniform(0.0, 0.5, (T, 1))
    _, R, _ = r.run_record(u, states=False, raster=True)
    _, Ro, _ = o.run_record(u, T, states=False, raster=True)
    assert Ro.sum() > 100
    assert (R != Ro).mean() <= 0.01
    assert abs(R.mean() - Ro.mean()) <= 0.05 * Ro.mean()


def test_step_api_and_spikes_match_oracle():
    r, o = _pair(CASES[0], din=3)
    u = np.random.default_rng(1).uniform(-0.5, 0.5, (60, 3))
    for t in range(60):
        r.step(u[t] if t % 7 else None)        # NULL input == zeros
        o.step(u[t] if t % 7 else None)
        assert np.array_equal(r.spikes(), o.spikes())
    assert np.array_equal(_bits(r.state()), _bits(o.state()))
    np.testing.assert_allclose(r.output(), o.output(), rtol=1e-12)


def test_reset_keeps_synapse_state_and_gl_table():
    case = CASES[5]   # PSC-homogeneous: its trace survives reset (reservoir.c:770-780)
    r, o = _pair(case)
    u = np.random.default_rng(2).uniform(0, 0.5, (80, 1))
    r.run_record(u); o.run_record(u, 80)
    r.reset(); o.reset()
    X, R, _ = r.run_record(u)
    Xo, Ro, _ = o.run_record(u, 80)
    assert np.array_equal(R, Ro) and np.array_equal(_bits(X), _bits(Xo))
    # reset re-reads the retained params pointer (bias change)
    cfg_arr = r._keep[1]
    cfg_arr[6] = 0.2
    o._keep[1][6] = 0.2
    r.reset(); o.reset()
    X, R, _ = r.run_record(u)
    Xo, Ro, _ = o.run_record(u, 80)
    assert np.array_equal(R, Ro) and np.array_equal(_bits(X), _bits(Xo))


def test_explicit_csr_gaussian_weights_bit_exact():
    """Scaled-down shape of the N=4096 small-world workload: FLIF-GL on a Watts-Strogatz ring with
    N(0,1) weights rescaled to rho = 0.95, gain 2.0, u ~ U[0, 0.5]."""
    n, L = 512, 64
    params = [1.0, 0.0, 0.0, 20.0, 0.8, float(L), 0.05, 2.0]
    sa.srand(11)
    rp, ci, _, Wi, _ = sa.build_weights(sa.make_config(n, conn=20 / (n - 1), conn_type="small_world",
                                                       params=params))
    val = np.random.default_rng(11).normal(0, 1, len(ci))
    val *= 0.95 / po.csr_spectral_radius(rp, ci, val)
    Wi = np.random.default_rng(12).uniform(-1, 1, (n, 1))
    kw = dict(din=1, dout=4, gain=2.0, neuron="flif_gl", params=params)
    r = sa.Reservoir(sa.make_config(n, **kw), csr=(rp, ci, val), W_in=Wi)
    o = po.OracleReservoir(po.make_config(n, **kw), csr=(rp, ci, val), W_in=Wi)
    T = 600
    u = np.random.default_rng(13).uniform(0, 0.5, (T, 1))
    X, R, _ = r.run_record(u)
    Xo, Ro, _ = o.run_record(u, T)
    assert Ro.sum() > 0
    assert np.array_equal(R, Ro)
    assert np.array_equal(_bits(X), _bits(Xo))


def test_spires_run_outputs():
    r, o = _pair(CASES[1], dout=3)
    u = np.random.default_rng(4).uniform(0, 0.5, (100, 1))
    Y = r.run(u)
    _, _, Yo = o.run_record(u, 100, states=False, raster=False, outputs=True)
    np.testing.assert_allclose(Y, Yo, rtol=1e-12, atol=1e-13)


def test_ridge_matches_oracle_lu():
    """W_out from GPU Cholesky + refinement vs the oracle's per-output LU
    (reservoir.c:573-679); tolerance 1e-9 relative on a well-conditioned system."""
    case = ("lif_discrete", [0.0, 1.0, 0.1, 0.0], 1.0, "simple", None, "sparse", "random", 0.1, 1.0)
    r, o = _pair(case, n=96, dout=3)
    T = 600
    rng = np.random.default_rng(5)
    u = rng.uniform(0, 1.0, (T, 1))
    y = np.stack([np.roll(u[:, 0], d) for d in (1, 2, 3)], axis=1)
    r.train_ridge(u, y, 1e-2)
    o.train_ridge(u, y, 1e-2)
    W, Wo = r.W_out(), o.W_out()
    assert np.abs(W - Wo).max() <= 1e-9 * np.abs(Wo).max()
    # the state after training is the end of the training run in both
    assert np.array_equal(_bits(r.state()), _bits(o.state()))


def test_ridge_washout_against_numpy():
    case = CASES[0]
    r, _ = _pair(case, n=128, dout=2)
    T, wo = 500, 50
    rng = np.random.default_rng(6)
    u = rng.uniform(0, 0.5, (T, 1))
    y = np.stack([np.roll(u[:, 0], 1), np.roll(u[:, 0], 5)], axis=1)[wo:]
    X, _, _ = r.run_record(u)   # states of the same run that training repeats after its reset
    info = r.train_ridge_washout(u, y, wo, 1e-4)
    A = X[wo:].T @ X[wo:] + 1e-4 * np.eye(128)
    Wn = np.linalg.solve(A, X[wo:].T @ y).T
    kappa = np.linalg.cond(A)
    print("kappa", kappa, "residual", info[1])
    assert info[1] < 1e-10                              # normal-equation residual after refinement
    # two backward-stable solvers agree to about kappa * u
    assert np.abs(r.W_out() - Wn).max() <= max(1e-9, 1e-14 * kappa) * np.abs(Wn).max()


def test_rls_and_online_match_oracle():
    case = ("lif_discrete", [0.0, 1.0, 0.1, 0.0], 1.0, "simple", None, "sparse", "random", 0.1, 1.0)
    r, o = _pair(case, n=64, dout=2)
    T = 200
    u = np.random.default_rng(8).uniform(0, 1, (T, 1))
    y = np.stack([np.roll(u[:, 0], 1), u[:, 0] ** 2], axis=1)
    r.train_rls(u, y, 1.0, 0.999)
    o.train_rls(u, y, 1.0, 0.999)
    np.testing.assert_allclose(r.W_out(), o.W_out(), rtol=1e-8, atol=1e-10)
    tgt = np.array([0.3, -0.2])
    r.train_online(tgt, 0.01)
    o.train_online(tgt, 0.01)
    np.testing.assert_allclose(r.W_out(), o.W_out(), rtol=1e-8, atol=1e-10)


def test_coarse_grain_matches_oracle():
    case = ("flif_gl", GL, 1.0, "simple", None, "sparse", "random", 0.1, 2.0)
    r, o = _pair(case, n=60)
    u = np.random.default_rng(9).uniform(0, 0.5, (30, 1))
    r.run_record(u); o.run_record(u, 30)
    thr = 0.25
    rq, oq = r.coarse_grain(thr), o.coarse_grain(thr)
    assert rq.N == oq.N and rq.N < 60
    assert np.array_equal(_bits(rq.weights()), _bits(oq.dense()))
    assert np.array_equal(_bits(rq.state()), _bits(oq.state()))
    X, R, _ = rq.run_record(u)
    Xo, Ro, _ = oq.run_record(u, 30)
    assert np.array_equal(R, Ro) and np.array_equal(_bits(X), _bits(Xo))


def test_invalid_inputs_raise_cleanly():
    r, _ = _pair(CASES[0])
    with pytest.raises(sa.SpiresError):
        r.read_records(0, 10 ** 6, states=True)
    L = sa.lib()
    assert L.spires_read_reservoir_state(r.h, None) == 1
    assert L.spires_train_ridge(r.h, None, None, 5, 1.0) == 1


@pytest.mark.parametrize("path", [sa.PATH_PER_STEP, sa.PATH_PERSISTENT, sa.PATH_PERSISTENT_HBM,
                                  sa.PATH_PERSISTENT_LDS])
@pytest.mark.parametrize("case", [CASES[0], CASES[2], CASES[5], CASES[7], CASES[11]],
                         ids=["gl-ws", "gl-ba", "gl-psch", "gl-psche", "lifd"])
def test_kernel_paths_agree_with_oracle(case, path):
    """Every exact-mode kernel path (one launch per micro-step; one persistent
    launch per run with product waves feeding the history sum, or a single-wave
    sum over the history in HBM or LDS) reproduces the oracle, including across several runs that
    continue the same state."""
    print("case", case, "path", path, flush=True)
    r, o = _pair(case, n=300)
    r.set_kernel_path(path)
    u = np.random.default_rng(21).uniform(0, 0.5, (90, 1))
    for part in (u[:30], u[30:31], u[31:90]):
        X, R, _ = r.run_record(part)
        Xo, Ro, _ = o.run_record(part, part.shape[0])
        assert np.array_equal(R, Ro) and np.array_equal(_bits(X), _bits(Xo))
    r.step(u[0]); o.step(u[0])
    assert np.array_equal(_bits(r.state()), _bits(o.state()))


def test_persistent_hbm_ring_path():
    """N*L too large for LDS: the persistent kernel streams the history from HBM."""
    n, L = 12288, 256
    params = [1.0, 0.0, 0.0, 20.0, 0.8, float(L), 0.05, 2.0]
    sa.srand(5)
    rp, ci, _, _, _ = sa.build_weights(sa.make_config(n, conn=10 / (n - 1), conn_type="small_world",
                                                      params=params))
    val = np.random.default_rng(5).normal(0, 1, len(ci))
    val *= 0.95 / sa.spectral_radius(rp, ci, val)
    Wi = np.random.default_rng(6).uniform(-1, 1, (n, 1))
    kw = dict(din=1, dout=1, gain=2.0, neuron="flif_gl", params=params)
    r = sa.Reservoir(sa.make_config(n, **kw), csr=(rp, ci, val), W_in=Wi)
    r.set_kernel_path(sa.PATH_PERSISTENT)
    o = po.OracleReservoir(po.make_config(n, **kw), csr=(rp, ci, val), W_in=Wi)
    T = 40
    u = np.random.default_rng(7).uniform(0, 0.5, (T, 1))
    X, R, _ = r.run_record(u)
    Xo, Ro, _ = o.run_record(u, T)
    assert np.array_equal(R, Ro) and np.array_equal(_bits(X), _bits(Xo))


def _gl_gauss(n, L, seed, topo="small_world", k=20, dout=2):
    params = [1.0, 0.0, 0.0, 20.0, 0.8, float(L), 0.05, 2.0]
    sa.srand(seed)
    rp, ci, _, _, _ = sa.build_weights(sa.make_config(n, conn=k / (n - 1), conn_type=topo, params=params))
    val = np.random.default_rng(seed).normal(0, 1, len(ci))
    val *= 0.95 / sa.spectral_radius(rp, ci, val)
    Wi = np.random.default_rng(seed + 1).uniform(-1, 1, (n, 1))
    kw = dict(din=1, dout=dout, gain=2.0, neuron="flif_gl", params=params)
    return kw, (rp, ci, val), Wi


@pytest.mark.parametrize("K,topo,k", [(16, "small_world", 20), (64, "small_world", 20), (32, "scale_free", 8)])
def test_fast_mode_tracks_fp64(K, topo, k):
    """FP32 time-blocked history (MFMA contraction + in-block triangle) against
    the FP64 oracle: identical rasters until the first threshold flip, and the
    membrane error before it stays at FP32 rounding level (tolerance 2e-4 on
    |V| <= ~1).  Rows longer than one 16-edge pull segment (k = 20, and the
    scale-free hubs with hundreds of in-edges) are summed in segments."""
    n, L, T = 640, 160, 500
    kw, csr, Wi = _gl_gauss(n, L, 17, topo=topo, k=k)
    print("longest row", np.diff(csr[0]).max())
    r = sa.Reservoir(sa.make_config(n, **kw), csr=csr, W_in=Wi)
    r.set_mode(sa.MODE_FAST, K)
    o = po.OracleReservoir(po.make_config(n, **kw), csr=csr, W_in=Wi)
    u = np.random.default_rng(18).uniform(0, 0.5, (T, 1))
    X, R, _ = r.run_record(u)
    Xo, Ro, _ = o.run_record(u, T)
    diff_rows = np.nonzero((R != Ro).any(axis=1))[0]
    t_div = diff_rows[0] if len(diff_rows) else T
    print("fast K", K, "first raster divergence at step", t_div, "agreement", (R == Ro).mean())
    assert t_div >= 40
    dv = np.abs(X[:t_div] - Xo[:t_div]).max()
    print("max |dV| before divergence", dv)
    assert dv < 2e-4
    assert (R == Ro).mean() > 0.9


def test_mode_switching_keeps_state():
    """exact -> fast -> exact on one reservoir; each exact segment continues
    from the state the fast segment left (checked against an oracle that is
    re-synchronised from the GPU state)."""
    n, L = 384, 96
    kw, csr, Wi = _gl_gauss(n, L, 23)
    r = sa.Reservoir(sa.make_config(n, **kw), csr=csr, W_in=Wi)
    o = po.OracleReservoir(po.make_config(n, **kw), csr=csr, W_in=Wi)
    u = np.random.default_rng(24).uniform(0, 0.5, (300, 1))
    X, R, _ = r.run_record(u[:100])
    Xo, Ro, _ = o.run_record(u[:100], 100)
    assert np.array_equal(_bits(X), _bits(Xo))
    r.set_mode(sa.MODE_FAST, 32)
    Xf, Rf, _ = r.run_record(u[100:200])
    Xfo, Rfo, _ = o.run_record(u[100:200], 100)
    assert np.abs(Xf[:10] - Xfo[:10]).max() < 2e-4
    r.set_mode(sa.MODE_EXACT, 0)
    X2, R2, _ = r.run_record(u[200:300])
    assert np.isfinite(X2).all() and R2.sum() > 0


def _nrmse(y, yh):
    return float(np.sqrt(np.mean((y - yh) ** 2)) / np.std(y))


def test_fast_readout_split_bf16_gram_against_fp64_solve():
    """Fast-mode ridge on the fast-mode states against numpy's FP64 ridge on the
    same states.  N = 384 is a multiple of 128, so the Gram comes from the
    split-BF16 matrix-core kernel (flag bit 2) and its factor preconditions
    conjugate gradients on the FP64 normal equations: with lambda large the
    solution is the FP64 one to 1e-8; with a tiny lambda the factor needs a
    diagonal shift (flag bit 3) and the solution agrees to the accuracy the
    conditioning allows (kappa-scaled), with a small FP64 normal-equation
    residual."""
    n, L, T, w = 384, 96, 2400, 200
    kw, csr, Wi = _gl_gauss(n, L, 41, dout=1)
    r = sa.Reservoir(sa.make_config(n, **kw), csr=csr, W_in=Wi)
    r.set_mode(sa.MODE_FAST, 32)
    u = np.random.default_rng(42).uniform(0, 0.5, (T, 1))
    y = np.roll(u[:, 0], 3)
    r.upload_inputs(u)
    r.reserve_records(T, sa.REC_STATE)
    r.run_resident(0, T, sa.REC_STATE)
    X, _, _ = r.read_records(0, T, states=True)
    Xs = X[w:]
    for lam in (0.05 * T, 1e-8):
        flags, resid = r.fit_ridge_records(w, T, y[w:], lam)
        A = Xs.T @ Xs + lam * np.eye(n)
        Wn = np.linalg.solve(A, Xs.T @ y[w:])
        Wg = r.W_out()[0]
        err = np.abs(Wg - Wn).max() / np.abs(Wn).max()
        kappa = np.linalg.cond(A)
        print("lambda", lam, "flags", flags, "residual", resid, "rel err", err, "kappa", kappa, flush=True)
        assert flags & 4                                  # split-BF16 Gram kernel + CG
        if lam > 1.0:
            assert err < 1e-8 and resid < 1e-10
        else:
            assert err <= max(1e-8, 1e-13 * kappa)
    r.close()


def test_fast_ridge_split_bf16_matches_fp64_readout():
    """Fast mode (FP32 history, split-BF16 Gram, FP64 Cholesky + refinement)
    against the FP64 exact path on a delayed-recall task: validation NRMSE of
    the two readouts within 0.02 of each other."""
    n, L, T, wo, Tv = 512, 128, 4000, 200, 1000
    kw, csr, Wi = _gl_gauss(n, L, 31, dout=4)
    rng = np.random.default_rng(32)
    u = rng.uniform(0, 0.5, (T + Tv, 1))
    delays = [1, 2, 4, 8]
    y = np.zeros((T + Tv, 4))
    for j, d in enumerate(delays):
        y[d:, j] = u[:-d, 0]
    scores = {}
    for mode in (sa.MODE_EXACT, sa.MODE_FAST):
        r = sa.Reservoir(sa.make_config(n, **kw), csr=csr, W_in=Wi)
        r.set_mode(mode, 32)
        r.train_ridge_washout(u[:T], y[wo:T], wo, 1e-5)
        _, _, Yh = r.run_record(u[T:], states=False, raster=False, outputs=True)
        scores[mode] = _nrmse(y[T:], Yh)
    print("NRMSE exact", scores[sa.MODE_EXACT], "fast", scores[sa.MODE_FAST])
    assert scores[sa.MODE_EXACT] < 0.9
    assert abs(scores[sa.MODE_FAST] - scores[sa.MODE_EXACT]) < 0.02
