"""GPU parity at the benchmark workloads' own sizes.

- FLIF-GL N=4096, L=512, Watts-Strogatz k=20, T=20000 (the bench workload):
  every raster and membrane row bit for bit against the oracle; the ridge
  readout of the 100 delayed-recall targets against numpy's FP64 solve of the
  same normal equations (relative error <= max(1e-9, 1e-14 * kappa)).
- 64 uneven FLIF-GL reservoirs under HIP-graph time blocks: every reservoir's
  records bit for bit against the same reservoir run alone, two of them
  against the oracle.
- Fast mode at N=16384, L=1024, Barabasi-Albert m=8, T=50000: raster
  agreement, max |dV| before the first raster divergence and readout NRMSE
  against FP64 exact mode (the fast readout's Gram from the split-BF16
  matrix-core kernel); exact mode against the oracle over the first 1000
  steps.
- The AGILE search's evaluator at N=2048: trial runs bit for bit against the
  oracle, trial scores against a numpy ridge on the oracle's states.
"""
import numpy as np
import pytest

import spires_amd as sa
from oracle import pyoracle as po
from spires_amd import agile_search as ag
from spires_amd import workloads

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def _first_diff(a, b):
    bad = np.argwhere(a != b)
    return bad[0].tolist() if len(bad) else None


def test_flifgl_n4096_l512_small_world_full_run_and_readout():
    print("case: FLIF-GL N=4096 L=512 WS k=20 T=20000, exact mode vs oracle", flush=True)
    T, n, D, lam = 20000, 4096, 100, 1e-5
    cfg, csr, W_in, u = workloads.flifgl_small_world(T=T)
    params = [cfg[1][i] for i in range(8)]
    r = sa.Reservoir(cfg, csr=csr, W_in=W_in)
    assert r.kernel_plan()["kernel"] == 2          # the product-wave kernel the bench times
    r.upload_inputs(u)
    flags = sa.REC_STATE | sa.REC_RASTER | sa.REC_APPEND
    r.reserve_records(T, flags)
    for t0 in range(0, T, 5000):                   # four launches: continuation across launches too
        r.run_resident(t0, 5000, flags)
    o = po.OracleReservoir(po.make_config(n, din=1, dout=D, gain=2.0, neuron="flif_gl", params=params),
                           csr=csr, W_in=W_in)
    X = np.empty((T, n))
    spikes = 0
    chunk = 2000
    for a in range(0, T, chunk):
        Xo, Ro, _ = o.run_record(u[a:a + chunk], chunk)
        Xg, Rg, _ = r.read_records(a, chunk, states=True, raster=True)
        assert np.array_equal(Rg, Ro), f"raster differs at row {a} + {_first_diff(Rg, Ro)}"
        assert np.array_equal(_bits(Xg), _bits(Xo)), f"V differs at row {a} + {_first_diff(Xg, Xo)}"
        X[a:a + chunk] = Xg
        spikes += int(Ro.sum())
    print(f"bit-exact over {T} steps, {spikes} spikes", flush=True)
    assert spikes > 0
    # ridge readout over all recorded states (train_output_ridge_regression records every step)
    Y = workloads.delayed_recall_targets(u, list(range(1, D + 1)))
    flags_fit, resid = r.fit_ridge_records(0, T, Y, lam)
    W = r.W_out()
    A = X.T @ X
    A[np.diag_indices(n)] += lam
    B = X.T @ Y
    Wn = np.linalg.solve(A, B).T
    ev = np.linalg.eigvalsh(A)
    kappa = float(ev[-1] / ev[0])
    rel = float(np.linalg.norm(W - Wn) / np.linalg.norm(Wn))
    tol = max(1e-9, 1e-14 * kappa)
    # backward error of the GPU solution in the normal equations (kappa-free)
    berr = float(np.linalg.norm(A @ W.T - B) / (np.linalg.norm(A) * np.linalg.norm(W) + np.linalg.norm(B)))
    print(f"ridge: solver {'LU' if flags_fit & 1 else 'Cholesky'} kappa {kappa:.3e} rel {rel:.3e} "
          f"tol {tol:.3e} backward {berr:.3e} device residual {resid:.3e}", flush=True)
    assert rel <= tol
    assert berr <= 1e-14
    r.close()


def _uneven_batch():
    u, _, _ = workloads.class_sequences()
    specs = workloads.uneven_batch_specs()
    built = []
    for sp in specs:
        params = workloads.flif_gl_params(alpha=sp["alpha"], T_mem=sp["T_mem"])
        cfg, csr, W_in = workloads.gaussian_reservoir(sp["n"], sp["topology"], sp["conn"], sp["rho"], params,
                                                      sp["seed"], din=13, dout=10, gain=1.0)
        built.append((sp, params, cfg, csr, W_in))
    return u, built


def test_uneven_batch_of_64_time_blocks_bit_exact():
    print("case: 64 uneven FLIF-GL reservoirs, 13-channel AR(1) input, graph time blocks K=256 x 3", flush=True)
    u, built = _uneven_batch()
    K, blocks = 256, 3
    T = K * blocks
    flags = sa.REC_STATE | sa.REC_RASTER
    res = []
    for sp, params, cfg, csr, W_in in built:
        r = sa.Reservoir(cfg, csr=csr, W_in=W_in)
        r.upload_inputs(u[:T])
        r.reserve_records(T, flags)
        res.append(r)
    plans = [r.kernel_plan() for r in res]
    tp = [i for i in range(len(res)) if plans[i]["batch_nb"] > 0]
    singles = [i for i in range(len(res)) if plans[i]["batch_nb"] == 0]
    assert len(tp) >= 48
    tb = sa.TimeBlock([res[i] for i in tp], K, flags)
    print(f"{len(tp)} reservoirs in {tb.launches} graph launches, {len(singles)} alone", flush=True)
    for b in range(blocks):
        tb.run(1)
        for i in singles:
            res[i].run_resident(b * K, K, flags | sa.REC_APPEND)
    tb.close()
    # every reservoir against a twin run alone over the same T steps
    for i, (sp, params, cfg, csr, W_in) in enumerate(built):
        twin = sa.Reservoir(cfg, csr=csr, W_in=W_in)
        twin.upload_inputs(u[:T])
        X2, R2, _ = twin.run_record(u[:T])
        X1, R1, _ = res[i].read_records(0, T, states=True, raster=True)
        assert np.array_equal(R1, R2), f"reservoir {i} (N={sp['n']}) raster differs at {_first_diff(R1, R2)}"
        assert np.array_equal(_bits(X1), _bits(X2)), f"reservoir {i} V differs at {_first_diff(X1, X2)}"
        twin.close()
    # the smallest Erdos-Renyi and Barabasi-Albert reservoirs against the oracle
    for topo in ("random", "scale_free"):
        i = min((k for k in range(len(built)) if built[k][0]["topology"] == topo), key=lambda k: built[k][0]["n"])
        sp, params, cfg, csr, W_in = built[i]
        o = po.OracleReservoir(po.make_config(sp["n"], din=13, dout=10, gain=1.0, neuron="flif_gl", params=params),
                               csr=csr, W_in=W_in)
        Xo, Ro, _ = o.run_record(u[:T], T)
        X1, R1, _ = res[i].read_records(0, T, states=True, raster=True)
        assert Ro.sum() > 0
        assert np.array_equal(R1, Ro) and np.array_equal(_bits(X1), _bits(Xo)), f"{topo} N={sp['n']} vs oracle"
    for r in res:
        r.close()


def test_fast_mode_n16384_l1024_full_run_against_exact():
    print("case: fast vs exact, FLIF-GL N=16384 L=1024 BA m=8 T=50000", flush=True)
    n, L, T, H = 16384, 1024, 50000, 84
    params = workloads.flif_gl_params(alpha=0.7, T_mem=L)
    cfg, csr, W_in = workloads.gaussian_reservoir(n, "scale_free", 8.0 / (n - 1), 0.9, params, 3, dout=1, gain=1.0)
    series = workloads.mackey_glass(T + H + 1, seed=0)
    u = np.ascontiguousarray(series[:T].reshape(T, 1))
    y = series[H:T + H]
    flags = sa.REC_STATE | sa.REC_RASTER
    runs = {}
    for mode in ("exact", "fast"):
        r = sa.Reservoir(cfg, csr=csr, W_in=W_in)
        if mode == "fast":
            r.set_mode(sa.MODE_FAST, 64)
        r.upload_inputs(u)
        r.reserve_records(T, flags)
        r.run_resident(0, T, flags)
        r.sync()
        runs[mode] = r
    # exact mode is the reference's arithmetic: first 1000 steps against the oracle
    o = po.OracleReservoir(po.make_config(n, din=1, dout=1, gain=1.0, neuron="flif_gl", params=params),
                           csr=csr, W_in=W_in)
    Xo, Ro, _ = o.run_record(u[:1000], 1000)
    Xe, Re, _ = runs["exact"].read_records(0, 1000, states=True, raster=True)
    assert np.array_equal(Re, Ro) and np.array_equal(_bits(Xe), _bits(Xo))
    del Xo, Ro, Xe, Re, o
    agree, t_div, chunk = 0, T, 2500
    for a in range(0, T, chunk):
        _, Re, _ = runs["exact"].read_records(a, chunk, raster=True)
        _, Rf, _ = runs["fast"].read_records(a, chunk, raster=True)
        eq = Re == Rf
        agree += int(eq.sum())
        if t_div == T:
            bad = np.nonzero(~eq.all(axis=1))[0]
            if len(bad):
                t_div = a + int(bad[0])
    dv = 0.0
    for a in range(0, t_div, chunk):
        m = min(chunk, t_div - a)
        Xe, _, _ = runs["exact"].read_records(a, m, states=True)
        Xf, _, _ = runs["fast"].read_records(a, m, states=True)
        dv = max(dv, float(np.abs(Xe - Xf).max()))
    ratio = agree / float(T * n)
    w, sp = T // 10, (3 * T) // 4
    nrmse = {}
    fit_flags = {}
    lams = (1e-2, 1.0, 100.0)
    for mode, r in runs.items():
        for lam in lams:
            fit_flags[(mode, lam)] = r.fit_ridge_records(w, sp, y[w:sp].reshape(-1, 1), lam)
            pred = r.predict_records(sp, T - sp)[:, 0]
            nrmse[(mode, lam)] = float(np.sqrt(np.mean((pred - y[sp:T]) ** 2)) / np.std(y[sp:T]))
    print(f"raster agreement {ratio:.4f}, first divergence at step {t_div}, max|dV| before it {dv:.3e}, "
          f"NRMSE {nrmse}, fits (flags, residual) {fit_flags}", flush=True)
    for lam in lams:
        assert fit_flags[("fast", lam)][0] & 4            # the split-BF16 Gram kernel was used
    assert ratio >= 0.85
    assert t_div >= 100
    assert dv <= 1e-3
    # lambda 1 and 100: the readouts agree within 0.02.  lambda 0.01: the normal
    # equations are so ill-conditioned that two FP64 solvers (the exact mode's
    # Cholesky of the formed Gram, the fast mode's CG applied through the record,
    # both at their floating-point floor) differ in directions the data does not
    # fix; the fast readout must then be no worse than 0.02 above the exact one.
    for lam in lams:
        d = nrmse[("fast", lam)] - nrmse[("exact", lam)]
        assert d <= 0.02 and (lam < 0.1 or abs(d) <= 0.02)
        assert nrmse[("exact", lam)] < 1.0
    for r in runs.values():
        r.close()


def test_agile_evaluator_n2048_against_oracle():
    print("case: AGILE evaluator N=2048, 2000-step rung, 3 trials vs oracle", flush=True)
    F = 2000
    ev = ag.Evaluator(n=2048, f_max=F, seed=0)
    th = ag.sample_trials(3, seed=77)
    th[:, 4] = [96.0, 400.0, 1000.0]                  # short, middle and long histories
    trials = [ag.Trial(t, 500 + i) for i, t in enumerate(th)]
    scores = ev.advance(trials, F)
    w, sp = ev.split(F)
    for t, s in zip(trials, scores):
        rp, ci, val, W_in = ev._weights(t.theta, t.seed)
        params = workloads.flif_gl_params(alpha=float(t.theta[1]), T_mem=float(t.theta[4]))
        o = po.OracleReservoir(po.make_config(2048, din=1, dout=1, gain=float(t.theta[2]), neuron="flif_gl",
                                              params=params), csr=(rp, ci, val), W_in=W_in)
        Xo, Ro, _ = o.run_record(ev.u, F)
        Xg, _, _ = t.res.read_records(0, F, states=True)
        assert np.array_equal(_bits(Xg), _bits(Xo)), f"trial {t.theta} differs at {_first_diff(Xg, Xo)}"
        # the trial's W_out and score from a numpy ridge on the oracle's states, at a
        # tolerance scaled by the condition number of its normal equations
        Xt = Xo[w:sp]
        A = Xt.T @ Xt
        A[np.diag_indices(2048)] += t.theta[5]
        Wn = np.linalg.solve(A, Xt.T @ ev.y[w:sp])
        ev_a = np.linalg.eigvalsh(A)
        kappa = float(ev_a[-1] / ev_a[0])
        tol = max(1e-9, 1e-14 * kappa)
        rel = float(np.linalg.norm(t.res.W_out()[0] - Wn) / np.linalg.norm(Wn))
        pred = Xo[sp:F] @ Wn
        yv = ev.y[sp:F]
        e = float(np.sqrt(np.mean((pred - yv) ** 2)) / np.std(yv))
        print(f"theta {np.round(t.theta, 4).tolist()}: GPU score {s:.8f}, numpy {e:.8f}, kappa {kappa:.2e}, "
              f"W_out rel {rel:.2e} (tol {tol:.2e})", flush=True)
        assert rel <= tol
        assert abs(s - e) <= max(1e-6, 10 * tol) * max(1.0, e)
    ev.release(trials)
    ev.close()
