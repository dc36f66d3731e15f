"""Hyperparameter search over FLIF-GL reservoirs on one GPU (the AGILE
workload with fidelity rungs and a learned surrogate).

Task: 84-step-ahead forecasting of the Mackey-Glass series (tau=17, RK4
dt=0.1, sampled every 1.0).  A trial is a FLIF-GL reservoir of N neurons with
six hyperparameters:
    rho    U[0.5, 1.5]       spectral radius of the recurrent weights
    alpha  U[0.3, 1.0]       fractional order
    gain   log-U[0.1, 10]    input gain
    conn   log-U[0.01, 0.2]  Erdos-Renyi connection probability
    T_mem  U[64, 1024]       history length (integer, dt = 1)
    lam    log-U[1e-8, 1e-2] ridge regulariser
evaluated at a fidelity of F steps: the reservoir runs F steps (FP64 exact
mode, several trials per device launch), the ridge readout is fitted on the
recorded states of steps [F/10, 3F/4) and scored on [3F/4, F) by
NRMSE = RMSE / std(target).

Recurrent weights are N(0,1) on an Erdos-Renyi graph, scaled by
rho / sqrt(N * conn): by the circular law the spectral radius of such a
matrix is sqrt(N * conn) to within a few percent at these sizes, which keeps
trial construction cheap (the exact power iteration of the reference costs
1000 sparse products per trial).  Input weights are U[-1, 1].

Searches (all report only scores measured by full reservoir runs):
  plain_search      every trial at the top fidelity;
  halving_search    successive halving over the rungs without a surrogate;
  surrogate_search  the same rungs, with the MLP surrogate (spires_amd.Surrogate,
                    trained on the GPU on every (hyperparameters, fidelity,
                    log10 NRMSE) record) screening candidates before they are
                    built and ranking promotions between rungs.
The last two keep promoting candidates to the top rung, best-ranked first,
until a target score is matched; `elapsed` is the wall time of the search on
one GPU.
"""
import math
import time

import numpy as np

import spires_amd as sa
from spires_amd import workloads

NAMES = ("rho", "alpha", "gain", "conn", "T_mem", "lam")
HORIZON = 84


def sample_trials(count, seed):
    """[count x 6] hyperparameters from the search distributions."""
    rng = np.random.default_rng(seed)
    th = np.empty((count, 6))
    th[:, 0] = rng.uniform(0.5, 1.5, count)
    th[:, 1] = rng.uniform(0.3, 1.0, count)
    th[:, 2] = 10.0 ** rng.uniform(-1.0, 1.0, count)
    th[:, 3] = 10.0 ** rng.uniform(-2.0, math.log10(0.2), count)
    th[:, 4] = np.floor(rng.uniform(64, 1025, count))
    th[:, 5] = 10.0 ** rng.uniform(-8.0, -2.0, count)
    return th


def encode(th, fidelity, f_min, f_max):
    """Surrogate input: the six hyperparameters mapped to [0, 1] (log scales
    where sampled log-uniformly) and log fidelity mapped to [0, 1]."""
    th = np.atleast_2d(th)
    z = np.empty((th.shape[0], 7), np.float32)
    z[:, 0] = (th[:, 0] - 0.5) / 1.0
    z[:, 1] = (th[:, 1] - 0.3) / 0.7
    z[:, 2] = (np.log10(th[:, 2]) + 1.0) / 2.0
    z[:, 3] = (np.log10(th[:, 3]) + 2.0) / (math.log10(0.2) + 2.0)
    z[:, 4] = (th[:, 4] - 64.0) / 960.0
    z[:, 5] = (np.log10(th[:, 5]) + 8.0) / 6.0
    z[:, 6] = math.log(fidelity / f_min) / max(1e-12, math.log(f_max / f_min))
    return z


class Evaluator:
    """Builds trials and scores them at a fidelity on the GPU."""

    def __init__(self, n=2048, f_max=18000, seed=0):
        self.n = n
        self.series = workloads.mackey_glass(f_max + HORIZON + 1, seed=seed)
        self.kernel_steps = 0          # trial-steps simulated
        self.trials_built = 0
        self.evaluations = 0
        self.run_seconds = 0.0         # synchronous reservoir launches
        self.fit_seconds = 0.0         # ridge fits and validation readback
        self.build_seconds = 0.0       # host-side trial construction and upload

    def _weights(self, th, trial_seed):
        rng = np.random.default_rng(trial_seed)
        n = self.n
        mask = rng.random((n, n)) < th[3]
        np.fill_diagonal(mask, False)
        rows, cols = np.nonzero(mask)
        rp = np.zeros(n + 1, np.int64)
        np.cumsum(np.bincount(rows, minlength=n), out=rp[1:])
        val = rng.standard_normal(len(cols)) * (th[0] / math.sqrt(n * th[3]))
        W_in = rng.uniform(-1.0, 1.0, (n, 1))
        return rp, cols.astype(np.int32), val, W_in

    def build(self, th, trial_seed):
        params = workloads.flif_gl_params(alpha=float(th[1]), T_mem=float(th[4]))
        cfg = sa.make_config(self.n, din=1, dout=1, gain=float(th[2]), conn=float(th[3]), neuron="flif_gl",
                             params=params)
        rp, ci, val, W_in = self._weights(th, trial_seed)
        self.trials_built += 1
        return sa.Reservoir(cfg, csr=(rp, ci, val), W_in=W_in)

    def _groups(self, res):
        """Greedy groups of reservoirs that run together in one launch."""
        groups, cur = [], []
        for r in res:
            if cur and sa.batch_workgroups(cur + [r]) == 0:
                groups.append(cur)
                cur = []
            cur.append(r)
        if cur:
            groups.append(cur)
        return groups

    def evaluate(self, thetas, seeds, F):
        """Validation NRMSE of each trial at fidelity F (FP64 exact, batched)."""
        u = self.series[:F].reshape(F, 1)
        y = self.series[HORIZON:F + HORIZON]
        w, split = max(50, F // 10), (3 * F) // 4
        tb = time.perf_counter()
        res = [self.build(th, s) for th, s in zip(thetas, seeds)]
        for r in res:
            r.upload_inputs(u)
            r.reserve_records(F, sa.REC_STATE)
        tr = time.perf_counter()
        self.build_seconds += tr - tb
        # one launch per group of trials with the same neurons per workgroup
        nbs = [r.kernel_plan()["nb"] for r in res]
        order = sorted(range(len(res)), key=lambda k: (nbs[k], thetas[k][4]))
        for grp in self._groups([res[k] for k in order]):
            if len(grp) > 1:
                sa.run_resident_batch(grp, 0, F, sa.REC_STATE)
            else:
                grp[0].run_resident(0, F, sa.REC_STATE)
        tf = time.perf_counter()
        self.run_seconds += tf - tr
        out = np.empty(len(res))
        yv = y[split:F]
        sd = float(np.std(yv))
        for k, (r, th) in enumerate(zip(res, thetas)):
            try:
                r.fit_ridge_records(w, split, y[w:split].reshape(-1, 1), float(th[5]))
                X, _, _ = r.read_records(split, F - split, states=True)
                pred = X @ r.W_out().reshape(-1)
                e = float(np.sqrt(np.mean((pred - yv) ** 2)) / sd)
                out[k] = e if np.isfinite(e) else 1e3
            except RuntimeError:
                out[k] = 1e3
            r.close()
        self.fit_seconds += time.perf_counter() - tf
        self.kernel_steps += F * len(res)
        self.evaluations += len(res)
        return out


def plain_search(ev, thetas, seeds, F):
    """Every trial at fidelity F, in batches."""
    t0 = time.perf_counter()
    scores = np.empty(len(thetas))
    step = 32
    for a in range(0, len(thetas), step):
        scores[a:a + step] = ev.evaluate(thetas[a:a + step], seeds[a:a + step], F)
    return {"best": float(scores.min()), "best_trial": int(scores.argmin()), "scores": scores,
            "elapsed": time.perf_counter() - t0}


def _promote_until(ev, thetas, seeds, ranked, F, target, best, log, t0, chunk=4):
    """Run trials `ranked` (best first) at the top fidelity, `chunk` per launch,
    until the best measured score is within 2 % of target."""
    done = 0
    while done < len(ranked) and not (target is not None and best <= 1.02 * target):
        idx = ranked[done:done + chunk]
        sc = ev.evaluate(thetas[idx], seeds[idx], F)
        for k, s in zip(idx, sc):
            log.append((int(k), F, float(s)))
        best = min(best, float(sc.min()))
        done += len(idx)
        if target is None and done >= chunk:
            break
    return best, time.perf_counter() - t0


def halving_search(ev, thetas, seeds, rungs, target=None, eta=3):
    """Successive halving without a surrogate: every trial at rung 0, the best
    1/eta move up a rung; top-rung candidates in rung order until target."""
    t0 = time.perf_counter()
    log = []
    alive = np.arange(len(thetas))
    best = math.inf
    for ri, F in enumerate(rungs[:-1]):
        sc = np.empty(len(alive))
        for a in range(0, len(alive), 32):
            sc[a:a + 32] = ev.evaluate(thetas[alive[a:a + 32]], seeds[alive[a:a + 32]], F)
        for k, s in zip(alive, sc):
            log.append((int(k), F, float(s)))
        alive = alive[np.argsort(sc)]
        alive = alive[:max(1, len(alive) // eta)] if ri < len(rungs) - 2 else alive
    best, el = _promote_until(ev, thetas, seeds, alive, rungs[-1], target, best, log, t0)
    return {"best": best, "elapsed": el, "log": log}


def surrogate_search(ev, thetas, seeds, rungs, target=None, warm=64, keep=0.25, promote=24, seed=1,
                     epochs=400, lr=1e-3):
    """Rungs as halving_search, with the surrogate screening candidates before
    their first run and ranking promotions by its top-fidelity prediction."""
    t0 = time.perf_counter()
    f_min, f_max = rungs[0], rungs[-1]
    net = sa.Surrogate(7, 64, seed=seed)
    X, Y, log = [], [], []

    def record(idx, F, sc):
        for k, s in zip(idx, sc):
            X.append(encode(thetas[k], F, f_min, f_max)[0])
            Y.append(math.log10(max(float(s), 1e-6)))
            log.append((int(k), F, float(s)))

    def train():
        net.train(np.array(X), np.array(Y, np.float32), epochs=epochs, lr=lr)

    def predict(idx, F):
        return net.predict(encode(thetas[idx], F, f_min, f_max))

    # rung 0 on a warm-up wave, then screen the rest by predicted rung-0 score
    first = np.arange(min(warm, len(thetas)))
    sc = np.concatenate([ev.evaluate(thetas[first[a:a + 32]], seeds[first[a:a + 32]], f_min)
                         for a in range(0, len(first), 32)])
    record(first, f_min, sc)
    train()
    rest = np.arange(len(first), len(thetas))
    if len(rest):
        pr = predict(rest, f_min)
        chosen = rest[np.argsort(pr)[:max(1, int(keep * len(rest)))]]
        sc = np.concatenate([ev.evaluate(thetas[chosen[a:a + 32]], seeds[chosen[a:a + 32]], f_min)
                             for a in range(0, len(chosen), 32)])
        record(chosen, f_min, sc)
        train()
    seen = np.array(sorted({k for k, _, _ in log}))
    # middle rungs: promote by predicted top-fidelity score
    for F in rungs[1:-1]:
        pred = predict(seen, f_max)
        up = seen[np.argsort(pred)[:promote]]
        sc = np.concatenate([ev.evaluate(thetas[up[a:a + 32]], seeds[up[a:a + 32]], F)
                             for a in range(0, len(up), 32)])
        record(up, F, sc)
        train()
        seen = up
    ranked = seen[np.argsort(predict(seen, f_max))]
    best, el = _promote_until(ev, thetas, seeds, ranked, f_max, target, math.inf, log, t0)
    net.close()
    return {"best": best, "elapsed": el, "log": log}


def spearman(a, b):
    ra = np.argsort(np.argsort(a)).astype(float)
    rb = np.argsort(np.argsort(b)).astype(float)
    ra -= ra.mean()
    rb -= rb.mean()
    return float((ra * rb).sum() / math.sqrt((ra * ra).sum() * (rb * rb).sum()))
