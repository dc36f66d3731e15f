#!/usr/bin/env python3
"""Benchmark of the reservoir step on one MI355X.

Workload (FP64 exact mode, single reservoir): FLIF-GL N=4096, T_mem=512
(L=512), dt=1, Watts-Strogatz ring k=20 p=0.1 with N(0,1) weights rescaled to
rho=0.95, input gain 2.0, one input channel u ~ U[0, 0.5], T=20000 time steps.
A reservoir time step is the recurrent CSR pull over the spike vector, the
FLIF-GL update with the (up to) 511-lag GL history sum, and the state row
written to HBM.  A bench "step" is a block of T/(steps+warmup) reservoir time
steps (800 at --steps 20 --warmup 5), run as one persistent launch; the
warm-up blocks fill the history, so the timed blocks sum all 511 lags.
Inputs, weights and the state record are resident in HBM before the timed
region.  value = N * timed time steps / HIP-event time of the timed region
(neuron-steps/s).  The roofline divides the algorithmic bytes of the lags
actually summed by the mean launch time; a 1 GiB device-to-device copy, timed
first, gives the measured copy bandwidth beside the 8 TB/s peak (and, under
rocprofv3 --pmc, the FETCH_SIZE / WRITE_SIZE calibration).

The CPU baseline is the oracle's C restatement of the reference step
(oracle/liboracle.so, OpenMP over neurons) on a bounded sample of the same
workload, timed on this host; its first steady-state chunk is compared bit for
bit with the GPU's records of the same steps.

The default line also carries the metric's second half, the AGILE search
(N=2048, 512 trials): the plain batched FP64 search over 512 trials at 18000
steps sets the target NRMSE, and the surrogate-guided search is timed per seed
until it comes within 2 % of it (--no-agile leaves it out).  --agile runs the
full comparison (with successive halving) as its own line.

With --fast it runs FLIF-GL N=16384, L=1024, BA m=8 in FP64 exact and
FP32 fast mode on the same inputs and reports both rates, raster agreement, the first
raster divergence, max |dV| before it and each readout's NRMSE (the block size K is
picked by timing the candidates on a twin reservoir).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-agile]
                       [--agile [--agile-seeds S]] [--fast [--fast-steps T]] [--batched] [--readout]
                       [--ridge-batch]
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "neuron-steps/s FLIF-GL N=4096 L=512; AGILE GPU-s to matched val NRMSE"
HBM_PEAK_GBS = 8000.0


class HipEvents:
    """HIP events recorded on the reservoir's own stream."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [C.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.hip.hipEventDestroy.argtypes = [C.c_void_p]

    def create(self):
        e = C.c_void_p()
        assert self.hip.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.hip.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.hip.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return ms.value


def measured_traffic_per_step():
    """HBM bytes per step of the step kernel from the committed rocprofv3 --pmc
    summary (FETCH_SIZE / WRITE_SIZE, copy-calibrated) for this workload."""
    f = os.path.join(ROOT, "profiles", "k_persist_tp_flifgl_n4096_L512.json")
    if not os.path.exists(f):
        return None, None
    with open(f) as fh:
        d = json.load(fh)
    return d["hbm_read_bytes_per_step"] + d["hbm_write_bytes_per_step"], d["source"]


def cpu_baseline(cfg, csr, W_in, L, budget_s=12.0, check=None):
    """The oracle's C restatement (OpenMP over neurons) stepping the same
    workload on this host's cores.  It draws its own input series from the
    workload's seed (long enough for the history ramp plus the timed sample,
    whatever the GPU run's length), runs steps 0..L-1 untimed (the history
    ramp), then times chunks of steady-state steps (full L-1-lag history) until
    `budget_s` is spent.  `check` = (first GPU state rows, their first step):
    the oracle's states of the first timed chunk are compared bit for bit with
    the GPU's records of the same steps."""
    import numpy as np
    from oracle import pyoracle as po
    from spires_amd import workloads
    po.build()
    n = cfg[0].num_neurons
    params = [cfg[1][i] for i in range(8)]
    o = po.OracleReservoir(po.make_config(n, din=1, dout=cfg[0].num_outputs, gain=cfg[0].input_strength,
                                          neuron="flif_gl", params=params), csr=csr, W_in=W_in)
    chunk, cap = 25, 200000
    u = workloads.flifgl_small_world_inputs(L + cap)
    o.run_record(u[:L], L, states=False, raster=False)
    done, parity = 0, None
    t0 = time.perf_counter()
    while done + chunk <= cap:
        X, _, _ = o.run_record(u[L + done:L + done + chunk], chunk, states=True, raster=False)
        if done == 0 and check is not None:
            G, row0 = check
            m = min(chunk, G.shape[0])
            parity = {"steps": [L, L + m - 1], "bit_exact": bool(row0 == L and np.array_equal(
                G[:m].view(np.uint64), X[:m].view(np.uint64)))}
        done += chunk
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    out = {"value": n * done / dt, "unit": "neuron-steps/s", "cores": cores, "kind": "port",
           "sample": f"steps {L}..{L + done - 1} of the same N={n} L={L} workload after an untimed "
                     f"{L}-step history ramp (oracle C restatement, OpenMP over neurons, {dt:.1f} s)"}
    if parity is not None:
        out["gpu_parity"] = parity
    return out


def agile_cpu_baseline(budget_s=10.0):
    """The oracle's C restatement (OpenMP over neurons) stepping one search
    trial (N=2048 FLIF-GL, T_mem=544 -- the mean of the T_mem range -- ER
    p=0.045, the median connection probability) for a bounded sample of steps;
    the plain search's 512 x 18000-step runs are priced at that rate."""
    import numpy as np
    from oracle import pyoracle as po
    from spires_amd import agile_search as ag
    po.build()
    th = np.array([1.0, 0.65, 1.0, 0.045, 544.0, 1e-5])
    ev = ag.Evaluator.__new__(ag.Evaluator)
    ev.n = 2048
    rp, ci, val, W_in = ag.Evaluator._weights(ev, th, 1)
    from spires_amd import workloads
    params = workloads.flif_gl_params(alpha=th[1], T_mem=th[4])
    o = po.OracleReservoir(po.make_config(2048, din=1, dout=1, gain=th[2], neuron="flif_gl", params=params),
                           csr=(rp, ci, val), W_in=W_in)
    u = workloads.mackey_glass(4000, seed=0).reshape(-1, 1)
    L = int(th[4])
    o.run_record(u[:L], L, states=False, raster=False)       # history ramp, untimed
    done, chunk = 0, 25
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        a = L + done % (len(u) - L - chunk)                    # the input series, cycled
        o.run_record(u[a:a + chunk], chunk, states=True, raster=False)
        done += chunk
    dt = time.perf_counter() - t0
    rate = 2048 * done / dt
    cores = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    plain_s = 512 * 18000 * 2048 / rate
    return {"value": 3600.0 / (plain_s / 512), "unit": "trials/hour (18000-step trials, simulation only)",
            "cores": cores, "kind": "port", "neuron_steps_per_s": rate,
            "sample": f"{done} steady-state steps of one N=2048 T_mem=544 trial (oracle C restatement, "
                      f"OpenMP over neurons, {dt:.1f} s); plain search priced at {plain_s:.0f} s"}


def run_agile(args):
    """AGILE hyperparameter search (N=2048, 512 trials, rungs 2k/6k/18k): the plain batched FP64 search (all 512
    trials at 18000 steps) sets the target; successive halving (no surrogate)
    and the surrogate-guided search (several search seeds) are timed until a
    top-rung validation NRMSE is within 2 % of it."""
    import numpy as np
    from spires_amd import agile_search as ag
    rungs = (2000, 6000, 18000)
    count = 512
    th = ag.sample_trials(count, seed=2024)
    seeds = np.arange(count) + 10000
    log = (lambda m: print(m, file=sys.stderr, flush=True))
    ev = ag.Evaluator(n=2048, f_max=rungs[-1], seed=0)
    plain = ag.plain_search(ev, th, seeds, rungs[-1], log=log)
    ev.close()
    plain.pop("scores")
    target = plain["best"]
    ev = ag.Evaluator(n=2048, f_max=rungs[-1], seed=0)
    halving = ag.halving_search(ev, th, seeds, rungs, target=target, log=log)
    ev.close()
    runs = []
    for sd in range(1, args.agile_seeds + 1):
        ev = ag.Evaluator(n=2048, f_max=rungs[-1], seed=0)
        r = ag.surrogate_search(ev, rungs, target=target, seed=sd, log=log)
        ev.close()
        r.pop("history")
        r["seed"] = sd
        r["matched"] = r["best"] <= 1.02 * target
        runs.append(r)
    times = sorted(r["elapsed"] if r["matched"] else math.inf for r in runs)
    med = float(np.median(times))
    halving["matched"] = halving["best"] <= 1.02 * target
    gbs = plain["alg_bytes"] / plain["run_s"] / 1e9
    line = {
        "metric": METRIC,
        "value": med,
        "unit": "GPU-s to matched val NRMSE (median over surrogate-search seeds)",
        "n_gpus": 1,
        "higher_is_better": False,
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "AGILE search: FLIF-GL N=2048, 512 trials, rungs 2000/6000/18000 steps, "
                               "84-step-ahead Mackey-Glass, NRMSE; 6 hyperparameters; surrogate 7-64-64-1 FP32 "
                               "Adam lr=1e-3", "target_nrmse": target, "target_tolerance": 0.02},
        "plain_gpu_s": plain["elapsed"],
        "speedup_median": plain["elapsed"] / med if med > 0 else None,
        "matched": f"{sum(r['matched'] for r in runs)}/{len(runs)}",
        "trials_per_hour_plain": 3600.0 * count / plain["elapsed"],
        "plain": plain,
        "halving": halving,
        "surrogate": runs,
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_mux_gl (plain search: 32 trials per launch of 64-neuron slices, 1000-step blocks; algorithmic bytes / run time)"},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = agile_cpu_baseline()
    print(json.dumps(line), flush=True)


def fast_k_autotune(make, u, Ks=(16, 32, 48, 64), steps=300):
    """Pick the fast-mode time block K by timing `steps` steps of a twin
    reservoir per candidate (HIP events on its stream); returns (K, {K: us/step})."""
    import spires_amd as sa
    ev = HipEvents()
    times = {}
    for K in Ks:
        r = make()
        r.set_mode(sa.MODE_FAST, K)
        r.upload_inputs(u[:steps + 50])
        r.reserve_records(steps, sa.REC_STATE)
        r.run_resident(0, 50, sa.REC_STATE)      # warm-up
        r.sync()
        e0, e1 = ev.create(), ev.create()
        ev.record(e0, r.stream())
        r.run_resident(50, steps, sa.REC_STATE)
        ev.record(e1, r.stream())
        r.sync()
        times[K] = ev.elapsed_ms(e0, e1) * 1e3 / steps
        r.close()
    return min(times, key=times.get), times


def run_batched(args):
    """64 FLIF-GL reservoirs of uneven N (log-U[256, 8192]), L, alpha
    and topology (ER p=0.05 / BA m=5), N(0,1) weights rescaled to rho, on one
    13-channel input of AR(1) class sequences, advanced in time blocks of K
    steps captured once in a HIP graph and replayed every block
    (spires_amd.BlockRunner).  The tile shape -- all 64 reservoirs in one
    cooperative launch of 64-neuron one-wave slices per block, or product-wave
    workgroups of 16/32 neurons packed per class into launches of at most one
    workgroup per CU (the others one launch each per block) -- and K are picked
    by spires_amd.autotune_blocks, which times one block per candidate; --packed
    keeps the product-wave shape.  The product-wave launches are also timed
    eagerly for comparison."""
    import numpy as np
    import spires_amd as sa
    from spires_amd import workloads
    u, lens, labels = workloads.class_sequences()
    T = u.shape[0]
    specs = workloads.uneven_batch_specs()
    res, bps = [], []
    for sp in specs:
        params = workloads.flif_gl_params(alpha=sp["alpha"], T_mem=sp["T_mem"])
        cfg, csr, W_in = workloads.gaussian_reservoir(sp["n"], sp["topology"], sp["conn"], sp["rho"], params, sp["seed"],
                                                      din=13, dout=10, gain=1.0)
        r = sa.Reservoir(cfg, csr=csr, W_in=W_in)
        r.upload_inputs(u)
        r.reserve_records(T, sa.REC_STATE)
        res.append(r)
        bps.append(workloads.algorithmic_bytes_per_neuron_step(sp["n"], len(csr[2]), sp["T_mem"]))
    plans = [r.kernel_plan() for r in res]
    tp = [i for i in range(len(res)) if plans[i]["batch_nb"] > 0]
    # launch-mates finish together when their step times are alike: order by
    # row-pull kind, history length and size
    tp.sort(key=lambda i: (plans[i]["weights_lds"], plans[i]["row_pull"], specs[i]["T_mem"], -specs[i]["n"]))
    singles = [i for i in range(len(res)) if plans[i]["batch_nb"] == 0]
    total_n = sum(sp["n"] for sp in specs)
    flags = sa.REC_STATE | sa.REC_APPEND

    def eager_block(t, K):
        """The product-wave launches without the graph (run_resident_batch per launch)."""
        groups = []
        for nb in (32, 16):
            cur = []
            for i in tp:
                if plans[i]["batch_nb"] != nb:
                    continue
                if cur and sa.batch_workgroups([res[j] for j in cur + [i]]) == 0:
                    groups.append(cur)
                    cur = []
                cur.append(i)
            if cur:
                groups.append(cur)
        for g in groups:
            sa.run_resident_batch([res[i] for i in g], t, K, flags)
        for i in singles:
            res[i].run_resident(t, K, flags)
        sa.synchronize(res)

    # untimed warm-up until every history is full
    t = 1024
    br = sa.BlockRunner(res, t, "slices")
    br.run(0)
    br.close()
    sa.synchronize(res)
    # autotune: the tile shape (one launch of 64-neuron one-wave slices, or packed
    # product-wave workgroups) and K, one block per candidate (each advances every reservoir)
    shapes = ("products",) if args.packed else sa.BlockRunner.SHAPES
    shape, K, table, t = sa.autotune_blocks(res, t, Ks=(64, 256, 1024), shapes=shapes)
    # eager product-wave launches for comparison: two blocks
    t0 = time.perf_counter()
    for _ in range(2):
        eager_block(t, K)
        t += K
    eager_us = (time.perf_counter() - t0) / (2 * K) * 1e6
    # timed: graph replays of whole blocks
    br = sa.BlockRunner(res, K, shape)
    launches = br.launches
    gsingles = [i for i in range(len(res)) if shape == "products" and plans[i]["batch_nb"] == 0]
    blocks = (T - t) // K
    t0 = time.perf_counter()
    br.run(t, blocks)
    sa.synchronize(res)
    wall = time.perf_counter() - t0
    t += blocks * K
    br.close()
    one_launch = shape == "slices"
    steps = blocks * K
    value = total_n * steps / wall
    alg = sum(b * sp["n"] for b, sp in zip(bps, specs)) * steps / wall / 1e9
    for r in res:
        r.close()
    line = {
        "metric": "batched neuron-steps/s, 64 uneven FLIF-GL reservoirs",
        "value": value, "unit": "neuron-steps/s", "n_gpus": 1, "steps": steps, "higher_is_better": True,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "64 FLIF-GL reservoirs, N log-U[256,8192] (total %d neurons), ER p=0.05 / BA m=5, "
                               "alpha U[0.5,1], T_mem in {128,256,512,1024}, rho U[0.7,1.2], N(0,1) weights; "
                               "13-channel AR(1) class sequences (512 sequences, %d frames)" % (total_n, T),
                   "block_k": K, "tile_shape": shape,
                   "graph": ("HIP graph per time block: one launch of 64-neuron slices (all %d reservoirs)" % len(res))
                   if one_launch else "HIP graph per time block (%d product-wave reservoirs)" % len(tp)},
        "autotune_us_per_step": {"%s/K=%d" % k: v for k, v in table.items()},
        "eager_us_per_step": eager_us,
        "launches_per_block": launches,
        "single_launches": [dict(n=specs[i]["n"], T_mem=specs[i]["T_mem"], kernel=plans[i]["kernel"]) for i in gsingles],
        "ms_per_step": wall / steps * 1e3,
        "roofline": {"bound": "hbm", "achieved": alg, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": alg / HBM_PEAK_GBS,
                     "traffic": None,
                     "kernel": "k_mux_gl (one launch per block)" if one_launch
                     else "k_persist_tp_batch (graph) + k_persist (all launches of a block)"},
    }
    print(json.dumps(line), flush=True)


LAMBDAS = (1e-2, 1.0, 1e2)   # ridge lambdas of the N=16384 readout comparison


def run_fast(args):
    """FLIF-GL N=16384, L=1024, alpha=0.7, Barabasi-Albert m=8, N(0,1)
    weights at rho=0.9, Mackey-Glass input, T steps, in FP64 exact mode (the
    reference's arithmetic, bit-exact with the oracle) and in fast mode (FP32
    MFMA history, split-BF16 Gram) on the same inputs.  Reports both rates,
    raster agreement, the first raster divergence, max |dV| before it, and the
    readout NRMSE of each (84-step-ahead forecast, fit on [T/10, 3T/4), for each
    ridge lambda in LAMBDAS)."""
    import numpy as np
    import spires_amd as sa
    from spires_amd import workloads
    n, L, T, H = 16384, 1024, args.fast_steps, 84
    params = workloads.flif_gl_params(alpha=0.7, T_mem=L)
    cfg, csr, W_in = workloads.gaussian_reservoir(n, "scale_free", 8.0 / (n - 1), 0.9, params, 3, dout=1, gain=1.0)
    series = workloads.mackey_glass(T + H + 1, seed=0)
    u = np.ascontiguousarray(series[:T].reshape(T, 1))
    y = series[H:T + H]
    make = lambda: sa.Reservoir(cfg, csr=csr, W_in=W_in)
    K, ktimes = fast_k_autotune(make, u)
    ev = HipEvents()
    out = {}
    res = {}
    for mode in ("exact", "fast"):
        r = make()
        if mode == "fast":
            r.set_mode(sa.MODE_FAST, K)
        r.upload_inputs(u)
        r.reserve_records(T, sa.REC_STATE | sa.REC_RASTER)
        e0, e1 = ev.create(), ev.create()
        ev.record(e0, r.stream())
        r.run_resident(0, T, sa.REC_STATE | sa.REC_RASTER)
        ev.record(e1, r.stream())
        r.sync()
        ms = ev.elapsed_ms(e0, e1)
        w, sp = T // 10, (3 * T) // 4
        fits = {}
        for lam in LAMBDAS:
            t_fit = time.perf_counter()
            info = r.fit_ridge_records(w, sp, y[w:sp].reshape(-1, 1), lam)
            t_fit = time.perf_counter() - t_fit
            pred = r.predict_records(sp, T - sp)[:, 0]
            if mode == "exact":
                gram, solver = "fp64 (matrix-core kernel)", "LU" if info[0] & 1 else "Cholesky"
            elif info[0] & 4:
                gram = "split-bf16 (matrix-core kernel)"
                solver = "CG on the FP64 normal equations, preconditioned by the Cholesky factor of the " \
                         "split-bf16 Gram" + (" with a diagonal shift" if info[0] & 8 else "")
            else:
                gram = "fp64 (split-bf16 Gram not positive definite)" if info[0] & 2 else "split-bf16 (rocBLAS)"
                solver = "LU" if info[0] & 1 else "Cholesky"
            fits[f"{lam:g}"] = {"val_nrmse": float(np.sqrt(np.mean((pred - y[sp:T]) ** 2)) / np.std(y[sp:T])),
                                "gram": gram, "solver": solver, "residual": info[1], "fit_seconds": t_fit}
        out[mode] = {"us_per_step": ms * 1e3 / T, "neuron_steps_per_s": n * T / (ms * 1e-3), "readout": fits}
        res[mode] = r
    # raster agreement and first divergence, in row chunks
    agree, t_div, chunk = 0, T, 2000
    for a in range(0, T, chunk):
        m = min(chunk, T - a)
        _, Re, _ = res["exact"].read_records(a, m, raster=True)
        _, Rf, _ = res["fast"].read_records(a, m, raster=True)
        eq = Re == Rf
        agree += int(eq.sum())
        if t_div == T:
            bad = np.nonzero(~eq.all(axis=1))[0]
            if len(bad):
                t_div = a + int(bad[0])
    dv = 0.0
    for a in range(0, t_div, chunk):
        m = min(chunk, t_div - a)
        Xe, _, _ = res["exact"].read_records(a, m, states=True)
        Xf, _, _ = res["fast"].read_records(a, m, states=True)
        dv = max(dv, float(np.abs(Xe - Xf).max()))
    for r in res.values():
        r.close()
    fs = out["fast"]["us_per_step"] * 1e-6
    tflops = 2.0 * (L - 1) * n / fs / 1e12
    line = {
        "metric": "fast mode vs FP64 exact, FLIF-GL N=16384 L=1024",
        "value": out["fast"]["neuron_steps_per_s"],
        "unit": "neuron-steps/s",
        "n_gpus": 1, "steps": T, "higher_is_better": True, "dtype": "f32 history / f64 update",
        "data": "synthetic",
        "config": {"workload": "FLIF-GL N=16384 L=1024 alpha=0.7, Barabasi-Albert m=8, N(0,1) weights rho=0.9, "
                               "gain 1.0, Mackey-Glass input, 84-step-ahead readout", "time_steps": T,
                   "block_k": K},
        "k_autotune_us_per_step": ktimes,
        "exact": out["exact"], "fast": out["fast"],
        "raster_agreement": agree / float(T * n),
        "first_raster_divergence_step": t_div,
        "max_abs_dV_before_divergence": dv,
        "nrmse_fast_minus_exact": {k: out["fast"]["readout"][k]["val_nrmse"] - out["exact"]["readout"][k]["val_nrmse"]
                                   for k in out["exact"]["readout"]},
        "speedup_fast_vs_exact": out["exact"]["us_per_step"] / out["fast"]["us_per_step"],
        "roofline": {"bound": "mfma", "achieved": tflops, "peak": 157.3, "unit": "TFLOP/s", "frac": tflops / 157.3,
                     "traffic": None, "kernel": "k_fast (GL history 2(L-1) flop per neuron-step on FP32 MFMA)"},
    }
    print(json.dumps(line), flush=True)


def run_readout(args):
    """Ridge readout of the FLIF-GL N=4096 L=512 small-world reservoir: 100
    delayed-recall targets u[t-d], d=1..100, fitted on all T=20000 recorded
    states with lambda=1e-5 (train_output_ridge_regression records every step,
    reservoir.c:573-679).  A = X'X + lambda I and B = X'Y are formed on the
    device from the state record in HBM (655 MB) (the Gram by the FP64
    matrix-core kernel), then one FP64 Cholesky (blocked kernels) for all 100
    outputs, batched triangular solves and two refinement steps.  Reports the
    fit time, its FP64 rate against the MFMA peak and the relative
    normal-equation residual."""
    import numpy as np
    import spires_amd as sa
    from spires_amd import workloads
    n, K, D, lam = 4096, 20000, 100, 1e-5
    cfg, csr, W_in, u = workloads.flifgl_small_world(T=K)
    Y = np.ascontiguousarray(workloads.delayed_recall_targets(u, list(range(1, D + 1))))
    r = sa.Reservoir(cfg, csr=csr, W_in=W_in)
    r.upload_inputs(u)
    r.reserve_records(K, sa.REC_STATE)
    r.run_resident(0, K, sa.REC_STATE)
    r.sync()
    r.fit_ridge_records(0, K, Y, lam)          # handle and workspace set-up
    times, flags, resid = [], None, None
    for _ in range(3):
        t0 = time.perf_counter()
        flags, resid = r.fit_ridge_records(0, K, Y, lam)
        times.append(time.perf_counter() - t0)
    t = min(times)
    # X'X (symmetric: N^2 T), X'Y, Cholesky N^3/3, three solve pairs and two residual products
    flop = n * n * K + 2.0 * n * K * D + n ** 3 / 3.0 + 3 * 2.0 * n * n * D + 2 * 2.0 * n * n * D
    W = r.W_out()
    line = {
        "metric": "ridge readout seconds, FLIF-GL N=4096 T=20000 Dout=100",
        "value": t, "unit": "s", "n_gpus": 1, "higher_is_better": False, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "ridge readout of the FLIF-GL N=4096 L=512 Watts-Strogatz k=20 reservoir: "
                               "100 delayed-recall targets, all 20000 recorded states, lambda 1e-5",
                   "state_record_bytes": n * K * 8},
        "fit_seconds_all": times,
        "solver": "LU" if flags & 1 else "Cholesky",
        "relative_residual": resid,
        "w_out_finite": bool(np.isfinite(W).all()),
        "roofline": {"bound": "mfma", "achieved": flop / t / 1e12, "peak": 78.6, "unit": "TFLOP/s",
                     "frac": flop / t / 1e12 / 78.6, "traffic": None,
                     "kernel": "k_gram_lower + k_chol_* + rocBLAS dgemm/dtrsm (whole fit)"},
    }
    print(json.dumps(line), flush=True)


def run_ridge_batch(args):
    """Batched ridge fits at the AGILE search's 6000-step rung: 16 FLIF-GL N=2048 trials
    (hyperparameters sampled as in the search, seed 5), each run 6000 steps with its states
    recorded, then one spires_ext_fit_ridge_records_batch call over rows [600, 4500) (3900
    rows per system, each with its own lambda).  Reports the time per call (best of five),
    its FP64 rate (N^2 T Gram + N^3/3 Cholesky + solves per system) and the largest relative
    normal-equation residual."""
    import numpy as np
    import spires_amd as sa
    from spires_amd import agile_search as ag
    n, F, count = 2048, 6000, 16
    ev = ag.Evaluator(n=n, f_max=F, seed=0)
    trials = [ag.Trial(th, 100 + i) for i, th in enumerate(ag.sample_trials(count, seed=5))]
    ev.advance(trials, F)                     # runs and a first (set-up) fit
    w, sp = ev.split(F)
    ytr = np.ascontiguousarray(ev.y[w:sp].reshape(-1, 1))
    res = [t.res for t in trials]
    lams = [float(t.theta[5]) for t in trials]
    times, info = [], None
    for _ in range(5):
        t0 = time.perf_counter()
        info = sa.fit_ridge_records_batch(res, w, sp, [ytr] * count, lams)
        times.append(time.perf_counter() - t0)
    t = min(times)
    T = sp - w
    flop = count * (n * n * T + 2.0 * n * T + n ** 3 / 3.0 + 10.0 * n * n)
    line = {
        "metric": "batched ridge fit seconds, 16 x FLIF-GL N=2048, 3900 rows",
        "value": t, "unit": "s", "n_gpus": 1, "higher_is_better": False, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "ridge fits of 16 AGILE-search trials at the 6000-step rung "
                               "(FLIF-GL N=2048, 84-step-ahead Mackey-Glass target, own lambda each)",
                   "systems": count, "rows": T},
        "fit_seconds_all": times,
        "solvers": [int(x) for x in info[:, 0]],
        "max_relative_residual": float(info[:, 1].max()),
        "roofline": {"bound": "mfma", "achieved": flop / t / 1e12, "peak": 78.6, "unit": "TFLOP/s",
                     "frac": flop / t / 1e12 / 78.6, "traffic": None,
                     "kernel": "k_gram_lower + k_cross_* + k_chol_* + k_resid_* (whole batched fit)"},
    }
    ev.release(trials)
    ev.close()
    print(json.dumps(line), flush=True)


T_CONFIG = 20000          # the single-reservoir workload's run length (time steps)
CLOCK_GHZ = 2.4           # MI355X peak engine clock
ADD_LATENCY_CYCLES = 8.5  # dependent v_add_f64 latency (one-wave probe)


def copy_bandwidth(hip, nbytes=1 << 30, reps=5):
    """Device-to-device copy of `nbytes` (hipMemcpyAsync on a stream of its
    own), best of `reps`: read + write bytes / time, in GB/s."""
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemsetAsync.argtypes = [C.c_void_p, C.c_int, C.c_size_t, C.c_void_p]
    hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    src, dst, st = C.c_void_p(), C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), nbytes) == 0 and hip.hipMalloc(C.byref(dst), nbytes) == 0
    assert hip.hipStreamCreate(C.byref(st)) == 0
    ev = HipEvents()
    best = 0.0
    try:
        assert hip.hipMemsetAsync(src, 1, nbytes, st) == 0
        assert hip.hipMemcpyAsync(dst, src, nbytes, 3, st) == 0          # warm-up
        for _ in range(reps):
            e0, e1 = ev.create(), ev.create()
            ev.record(e0, st)
            assert hip.hipMemcpyAsync(dst, src, nbytes, 3, st) == 0      # hipMemcpyDeviceToDevice
            ev.record(e1, st)
            best = max(best, 2.0 * nbytes / (ev.elapsed_ms(e0, e1) * 1e-3) / 1e9)
    finally:
        hip.hipStreamSynchronize(st)
        hip.hipStreamDestroy(st)
        hip.hipFree(src)
        hip.hipFree(dst)
    return best


def run_agile_summary(args):
    """The AGILE half of the metric for the default line: plain search, then the
    surrogate search over args.agile_seeds seeds (no halving leg)."""
    import numpy as np
    from spires_amd import agile_search as ag
    rungs = (2000, 6000, 18000)
    th = ag.sample_trials(512, seed=2024)
    seeds = np.arange(512) + 10000
    log = (lambda m: print(m, file=sys.stderr, flush=True))
    ev = ag.Evaluator(n=2048, f_max=rungs[-1], seed=0)
    plain = ag.plain_search(ev, th, seeds, rungs[-1], log=log)
    ev.close()
    target = plain["best"]
    per_seed = []
    for sd in range(1, args.agile_seeds + 1):
        ev = ag.Evaluator(n=2048, f_max=rungs[-1], seed=0)
        r = ag.surrogate_search(ev, rungs, target=target, seed=sd, log=log)
        ev.close()
        matched = r["best"] <= 1.02 * target
        per_seed.append({"seed": sd, "gpu_s": r["elapsed"], "best": r["best"], "matched": bool(matched),
                         "trials": r.get("trials")})
    times = sorted(x["gpu_s"] if x["matched"] else math.inf for x in per_seed)
    med = float(np.median(times))
    return {"unit": "GPU-s to matched val NRMSE (median over surrogate-search seeds; unmatched = never)",
            "value": med, "plain_gpu_s": plain["elapsed"], "target_nrmse": target,
            "speedup_median": plain["elapsed"] / med if med > 0 else None,
            "matched": f"{sum(x['matched'] for x in per_seed)}/{len(per_seed)}",
            "workload": "FLIF-GL N=2048, 512 trials, rungs 2000/6000/18000, 84-step-ahead Mackey-Glass, "
                        "6 hyperparameters, surrogate 7-64-64-1 FP32 Adam lr=1e-3; plain = 512 x 18000 steps",
            "seeds": per_seed}


def run_step(args):
    """The single-reservoir workload (FP64 exact FLIF-GL N=4096 L=512, T=20000).
    A bench step is a block of T/(steps+warmup) reservoir time steps (one
    persistent launch); the warm-up blocks fill the 511-lag history, so the timed
    blocks run in steady state.  When the warm-up blocks are too short for that,
    an untimed ramp runs first."""
    import numpy as np
    import spires_amd as sa
    from spires_amd import workloads
    K, W = args.steps, args.warmup
    L = 512
    block = max(1, T_CONFIG // max(1, K + W))
    ramp = max(0, (L - 1) - W * block)
    T = ramp + (K + W) * block
    cfg, csr, W_in, u = workloads.flifgl_small_world(T=T)
    n = cfg[0].num_neurons
    ev = HipEvents()
    copy_gbs = copy_bandwidth(ev.hip)
    r = sa.Reservoir(cfg, csr=csr, W_in=W_in)
    plan = r.kernel_plan()
    r.upload_inputs(u)
    flags = sa.REC_STATE | sa.REC_APPEND
    r.reserve_records(T, flags)
    stream = r.stream()
    t = 0
    if ramp:
        r.run_resident(0, ramp, flags)
        t = ramp
    for _ in range(W):
        r.run_resident(t, block, flags)
        t += block
    r.sync()
    t_timed0 = t
    marks = [(ev.create(), ev.create()) for _ in range(K)]
    w0 = time.perf_counter()
    for k in range(K):
        ev.record(marks[k][0], stream)
        r.run_resident(t, block, flags)
        ev.record(marks[k][1], stream)
        t += block
    r.sync()
    wall = time.perf_counter() - w0
    region_ms = ev.elapsed_ms(marks[0][0], marks[-1][1])
    launch_ms = np.array([ev.elapsed_ms(a, b) for a, b in marks])
    steps_timed = K * block
    value = n * steps_timed / (region_ms * 1e-3)
    nnz = len(csr[2])
    # algorithmic bytes of the lags actually summed: step st sums min(st, L-1) lags
    per_lag = 8 * n
    fixed = (workloads.algorithmic_bytes_per_neuron_step(n, nnz, L) - 8 * (L - 1)) * n
    lags = np.minimum(np.arange(t_timed0, t_timed0 + steps_timed), L - 1)
    alg_bytes_timed = float(per_lag * lags.sum() + fixed * steps_timed)
    alg_per_launch = alg_bytes_timed / K
    achieved = alg_per_launch / (float(launch_ms.mean()) * 1e-3) / 1e9
    us_step_med = float(np.median(launch_ms)) * 1e3 / block
    floor_us = (L - 1) * ADD_LATENCY_CYCLES / (CLOCK_GHZ * 1e3)
    xch_us = sa.exchange_floor(plan["workgroups"], 20000)
    traffic_step, traffic_src = measured_traffic_per_step()
    G, _, _ = r.read_records(L, 25, states=True) if T >= L + 25 else (None, None, None)
    X, _, _ = r.read_records(T - 1, 1, states=True)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "neuron-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": region_ms / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "FLIF-GL reservoir, FP64 exact mode: N=4096, T_mem=512 (L=512), dt=1, "
                               "Watts-Strogatz k=20 p=0.1, N(0,1) weights rho=0.95, gain 2.0, u~U[0,0.5], T=%d" % T,
                   "neurons": n, "memory_len": L, "nnz": nnz, "time_steps": T, "mode": "exact",
                   "bench_step": "%d reservoir time steps (one persistent launch)" % block,
                   "untimed_ramp_steps": ramp, "timed_time_steps": [t_timed0, t_timed0 + steps_timed - 1],
                   "record": "state V [T x N] f64"},
        "us_per_time_step": {"region": region_ms * 1e3 / steps_timed, "launch_median": us_step_med,
                             "launch_min": float(launch_ms.min()) * 1e3 / block,
                             "launch_max": float(launch_ms.max()) * 1e3 / block, "launches": K},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "frac_of_measured_copy": achieved / copy_gbs if copy_gbs else None,
                     "measured_copy_GBs": copy_gbs,
                     "traffic": traffic_step * block if traffic_step else None,
                     "traffic_source": traffic_src,
                     "algorithmic_bytes_per_launch": alg_per_launch,
                     "kernel": "k_persist_tp<%d> (one launch per bench step, %d workgroups)" % (plan["nb"],
                                                                                            plan["workgroups"])},
        "latency_floor": {"bound": "spike exchange among the launch's workgroups (measured), above the GL add chain",
                          "floor_us_per_time_step": max(floor_us, xch_us), "frac": max(floor_us, xch_us) / us_step_med,
                          "exchange_us": xch_us,
                          "exchange_basis": "publish + gather of one tagged word per workgroup per step, %d "
                                            "workgroups, no other work (spires_ext_exchange_floor)" % plan["workgroups"],
                          "chain_us": floor_us,
                          "chain_basis": "%d lags x %.1f cycles at %.1f GHz" % (L - 1, ADD_LATENCY_CYCLES, CLOCK_GHZ)},
        "wall_s": wall,
        "final_state_checksum": float(np.sum(X)),
    }
    r.close()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(cfg, csr, W_in, L, check=(G, L) if G is not None else None)
    if not args.no_agile:
        try:
            line["agile"] = run_agile_summary(args)
        except Exception as e:           # the step line is still reported
            line["agile"] = {"error": repr(e)}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-agile", action="store_true",
                    help="leave the AGILE search (the metric's second half) out of the default line")
    ap.add_argument("--agile", action="store_true",
                    help="run the AGILE hyperparameter search comparison instead of the step benchmark")
    ap.add_argument("--agile-seeds", type=int, default=12)
    ap.add_argument("--fast", action="store_true",
                    help="run the fast-mode comparison (FLIF-GL N=16384, L=1024) instead of the step benchmark")
    ap.add_argument("--fast-steps", type=int, default=50000)
    ap.add_argument("--readout", action="store_true",
                    help="time the ridge readout of the N=4096 workload (100 delayed-recall outputs)")
    ap.add_argument("--ridge-batch", action="store_true",
                    help="time one batched ridge fit of 16 N=2048 search trials at the 6000-step rung")
    ap.add_argument("--batched", action="store_true",
                    help="run the batched workload (64 uneven reservoirs) instead of the step benchmark")
    ap.add_argument("--packed", action="store_true",
                    help="with --batched: product-wave launches packed per neurons-per-workgroup class "
                         "instead of one launch of 64-neuron slices per time block")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library runs one reservoir per device and has no "
              "multi-GPU path (a reservoir would shard by neuron blocks, one spike exchange per step "
              "over the GPUs).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import spires_amd as sa
    from spires_amd import workloads

    if not os.path.exists(sa.LIB_PATH):
        sa.build()
    if args.agile:
        run_agile(args)
        return
    if args.fast:
        run_fast(args)
        return
    if args.batched:
        run_batched(args)
        return
    if args.readout:
        run_readout(args)
        return
    if args.ridge_batch:
        run_ridge_batch(args)
        return
    run_step(args)


if __name__ == "__main__":
    main()
