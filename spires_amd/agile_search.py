"""Hyperparameter search over FLIF-GL reservoirs on one GPU (the AGILE
workload with fidelity rungs and a learned surrogate).

Task: 84-step-ahead forecasting of the Mackey-Glass series (tau=17, RK4
dt=0.1, sampled every 1.0).  A trial is a FLIF-GL reservoir of N neurons with
six hyperparameters
    rho    U[0.5, 1.5]       spectral radius of the recurrent weights
    alpha  U[0.3, 1.0]       fractional order
    gain   log-U[0.1, 10]    input gain
    conn   log-U[0.01, 0.2]  Erdos-Renyi connection probability
    T_mem  U[64, 1024]       history length (integer, dt = 1)
    lam    log-U[1e-8, 1e-2] ridge regulariser
and a seed for its random weights.  At a fidelity of F steps the reservoir
has run F steps (FP64 exact mode, several trials per device launch), the
ridge readout is fitted on the recorded states of steps [F/10, 3F/4) and
scored on [3F/4, F) by NRMSE = RMSE / std(target).  A trial promoted to a
higher rung continues its run (the records of the lower rung are kept), so
its scores at every rung are those of one uninterrupted run.

Recurrent weights are N(0,1) on an Erdos-Renyi graph, scaled by
rho / sqrt(N * conn): by the circular law the spectral radius of such a matrix
is sqrt(N * conn) to within a few percent at these sizes, which keeps trial
construction cheap (the reference's power iteration costs 1000 sparse
products per trial).  Input weights are U[-1, 1].

Searches (every reported score is a full reservoir run at that fidelity):
  plain_search      all trials at the top rung;
  halving_search    successive halving over the rungs, no surrogate;
  surrogate_search  the MLP surrogate (spires_amd.Surrogate, trained on the
                    GPU on every (hyperparameters, fidelity, log10 NRMSE)
                    record) picks which proposals to run and which trials to
                    promote; proposals are global samples and Levy-flight
                    moves around the best trials so far.
`elapsed` is the wall time of a search on one GPU, host work included.
"""
import math
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

import spires_amd as sa
from spires_amd import workloads

NAMES = ("rho", "alpha", "gain", "conn", "T_mem", "lam")
HORIZON = 84
LO = np.array([0.5, 0.3, -1.0, -2.0, 64.0, -8.0])           # rho, alpha, log10 gain, log10 conn, T_mem, log10 lam
HI = np.array([1.5, 1.0, 1.0, math.log10(0.2), 1024.0, -2.0])


def to_unit(th):
    """Hyperparameters -> [0, 1]^6 (log scales where sampled log-uniformly)."""
    th = np.atleast_2d(th)
    z = np.column_stack([th[:, 0], th[:, 1], np.log10(th[:, 2]), np.log10(th[:, 3]), th[:, 4], np.log10(th[:, 5])])
    return (z - LO) / (HI - LO)


def from_unit(z):
    z = np.clip(np.atleast_2d(z), 0.0, 1.0) * (HI - LO) + LO
    return np.column_stack([z[:, 0], z[:, 1], 10.0 ** z[:, 2], 10.0 ** z[:, 3], np.floor(z[:, 4]), 10.0 ** z[:, 5]])


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


def latin_hypercube(count, seed):
    """[count x 6] hyperparameters from a Latin hypercube over the search
    distributions: each axis is cut into `count` equal-probability strata and
    every stratum holds one sample, so a small first wave still covers every
    axis end to end."""
    rng = np.random.default_rng(seed)
    z = (np.argsort(rng.random((6, count)), axis=1).T + rng.random((count, 6))) / count
    return from_unit(z)


def encode(th, fidelity, f_min, f_max):
    """Surrogate input: the six hyperparameters on [0, 1] and log fidelity on [0, 1]."""
    z = to_unit(th)
    fz = math.log(fidelity / f_min) / max(1e-12, math.log(f_max / f_min))
    return np.column_stack([z, np.full(len(z), fz)]).astype(np.float32)


class Trial:
    """One (hyperparameters, weight seed) candidate and its run so far."""
    __slots__ = ("theta", "seed", "res", "future", "done", "scores", "nb", "bps")

    def __init__(self, theta, seed):
        self.theta = np.asarray(theta, dtype=np.float64)
        self.seed = int(seed)
        self.res = None
        self.future = None
        self.done = 0
        self.scores = {}
        self.nb = 0
        self.bps = 0.0          # algorithmic bytes per neuron-step


class Evaluator:
    """Builds trials (on host threads, overlapping device work) and advances
    them to a fidelity on the GPU: batched runs, batched ridge fits, readout
    predictions on the device."""

    def __init__(self, n=2048, f_max=18000, seed=0, workers=16, runner="auto", block=1000, slices_min=16):
        """runner: "batch" runs trials that start at the same step as packed
        product-wave launches (run_resident_batch, at most one workgroup per
        CU); "slices" runs them all in one launch of 64-neuron one-wave slices
        per block of at most `block` steps (spires_amd.BlockRunner; the largest
        length that divides the group's steps); "auto" takes the slices for groups of at least
        `slices_min` trials (a slices launch lasts as long as its slowest trial
        and suits large groups) and packed launches for smaller ones."""
        self.n = n
        self.f_max = f_max
        self.runner = runner
        self.block = block
        self.slices_min = slices_min
        series = workloads.mackey_glass(f_max + HORIZON + 1, seed=seed)
        self.u = np.ascontiguousarray(series[:f_max].reshape(f_max, 1))
        self.y = np.ascontiguousarray(series[HORIZON:f_max + HORIZON])
        self.pool = ThreadPoolExecutor(max_workers=workers)
        # released trials are destroyed on a thread of their own (their frees then
        # overlap the next launches); drain() waits for them
        self.closer = ThreadPoolExecutor(max_workers=1)
        self.closing = []
        self.trial_steps = 0           # reservoir steps simulated, all trials
        self.trials_built = 0
        self.fits = 0
        self.run_seconds = 0.0         # device launches (synchronous)
        self.fit_seconds = 0.0         # ridge fits and validation predictions
        self.wait_seconds = 0.0        # waiting for trial construction
        self.alg_bytes = 0.0           # algorithmic HBM bytes of the simulated steps

    def drain(self):
        """Wait until every released trial has been destroyed."""
        for f in self.closing:
            f.result()
        self.closing = []

    def close(self):
        self.drain()
        self.closer.shutdown(wait=True)
        self.pool.shutdown(wait=True)

    @staticmethod
    def split(F):
        return max(50, F // 10), (3 * F) // 4

    def _weights(self, th, trial_seed):
        """Erdos-Renyi edges by geometric gaps between successive edges of the
        row-major n x n index space (no self loops), N(0,1) values."""
        rng = np.random.default_rng(trial_seed)
        n, c = self.n, float(th[3])
        total = n * n
        inv = 1.0 / math.log1p(-c)

        def gaps(k):   # Geometric(c) on {1, 2, ...} by inversion: floor(log(1 - U) / log(1 - c)) + 1
            g = rng.random(k)
            np.subtract(1.0, g, out=g)
            np.log(g, out=g)
            g *= inv
            out = g.astype(np.int64)
            out += 1
            return out

        pos = np.cumsum(gaps(int(total * c * 1.1) + 64))
        pos -= 1
        while pos[-1] < total:
            pos = np.concatenate([pos, np.cumsum(gaps(int(total * c * 0.1) + 64)) + pos[-1]])
        pos = pos[:np.searchsorted(pos, total)]   # increasing
        if n & (n - 1) == 0:
            rows, cols = pos >> (n.bit_length() - 1), pos & (n - 1)
        else:
            rows, cols = np.divmod(pos, n)
        keep = rows != cols
        rows, cols = rows[keep], cols[keep]
        rp = np.zeros(n + 1, np.int64)
        np.cumsum(np.bincount(rows, minlength=n), out=rp[1:])
        val = rng.standard_normal(len(cols), dtype=np.float32) * (th[0] / math.sqrt(n * th[3]))
        W_in = rng.uniform(-1.0, 1.0, (n, 1))
        return rp, cols.astype(np.int32), val, W_in

    def build(self, th, trial_seed):
        params = workloads.flif_gl_params(alpha=float(th[1]), T_mem=float(th[4]))
        cfg = sa.make_config(self.n, din=1, dout=1, gain=float(th[2]), conn=float(th[3]), neuron="flif_gl",
                             params=params)
        rp, ci, val, W_in = self._weights(th, trial_seed)
        r = sa.Reservoir(cfg, csr=(rp, ci, val), W_in=W_in)
        r.upload_inputs(self.u)
        if self.runner in ("slices", "auto"):   # time blocks write into reserved records
            r.reserve_records(self.f_max, sa.REC_STATE)
        r.bps = workloads.algorithmic_bytes_per_neuron_step(self.n, len(ci), int(th[4]))
        return r

    def prefetch(self, trials):
        """Start building trials on the host threads."""
        for t in trials:
            if t.res is None and t.future is None:
                t.future = self.pool.submit(self.build, t.theta, t.seed)

    def _ensure(self, trials):
        self.prefetch(trials)
        t0 = time.perf_counter()
        for t in trials:
            if t.res is None:
                t.res = t.future.result()
                t.future = None
                t.nb = t.res.kernel_plan()["batch_nb"]
                t.bps = t.res.bps
                self.trials_built += 1
        self.wait_seconds += time.perf_counter() - t0

    def release(self, trials):
        """Free the device memory of trials that will not run further."""
        for t in trials:
            if t.future is not None:
                t.res = t.future.result()
                t.future = None
            if t.res is not None:
                self.closing.append(self.closer.submit(t.res.close))
                t.res = None

    def _launch_groups(self, trials):
        """Trials that start at the same step, grouped by neurons per
        workgroup and packed greedily into launches that fit the device."""
        by_key = {}
        for t in trials:
            by_key.setdefault((t.nb, t.done), []).append(t)
        out = []
        for (_, done), ts in sorted(by_key.items()):
            ts.sort(key=lambda q: q.theta[4])
            cur = []
            for t in ts:
                if cur and sa.batch_workgroups([q.res for q in cur + [t]]) == 0:
                    out.append((done, cur))
                    cur = []
                cur.append(t)
            if cur:
                out.append((done, cur))
        return out

    def advance(self, trials, F, fit_batch=32):
        """Run each trial on to fidelity F and score it there.  Returns the
        validation NRMSE of each trial (1e3 for a singular fit or a blow-up)."""
        trials = list(trials)
        self._ensure(trials)
        t0 = time.perf_counter()
        by_done = {}
        for t in trials:
            if t.done < F:
                by_done.setdefault(t.done, []).append(t)
        groups = []
        for done, grp in sorted(by_done.items()):
            if self.runner == "slices" or (self.runner == "auto" and len(grp) >= self.slices_min):
                groups.append((done, grp, True))
            else:
                groups += [(d, g, False) for d, g in self._launch_groups(grp)]
        for done, grp, slices in groups:
            steps = F - done
            if slices:
                K = min(self.block, steps)   # the largest block length that divides the steps
                while steps % K:
                    K -= 1
                br = sa.BlockRunner([t.res for t in grp], K, "slices")
                try:
                    br.run(done, steps // K)
                finally:
                    br.close()
            elif len(grp) > 1:
                sa.run_resident_batch([t.res for t in grp], done, steps, sa.REC_STATE | sa.REC_APPEND)
            else:
                grp[0].res.run_resident(done, steps, sa.REC_STATE | sa.REC_APPEND)
            for t in grp:
                t.done = F
                self.alg_bytes += t.bps * self.n * steps
            self.trial_steps += steps * len(grp)
        t1 = time.perf_counter()
        self.run_seconds += t1 - t0
        w, sp = self.split(F)
        yv = self.y[sp:F]
        sd = float(np.std(yv))
        ytr = np.ascontiguousarray(self.y[w:sp].reshape(-1, 1))
        todo = [t for t in trials if F not in t.scores]
        for a in range(0, len(todo), fit_batch):
            part = todo[a:a + fit_batch]
            info = sa.fit_ridge_records_batch([t.res for t in part], w, sp, [ytr] * len(part),
                                              [t.theta[5] for t in part])
            for t, inf in zip(part, info):
                e = 1e3
                if inf[0] >= 0:
                    pred = t.res.predict_records(sp, F - sp)[:, 0]
                    e = float(np.sqrt(np.mean((pred - yv) ** 2)) / sd)
                    e = e if np.isfinite(e) else 1e3
                t.scores[F] = min(e, 1e3)
            self.fits += len(part)
        self.fit_seconds += time.perf_counter() - t1
        return np.array([t.scores[F] for t in trials])


def _stats(ev, t0):
    drain = getattr(ev, "drain", None)
    if drain is not None:   # destroys still pending count towards the search's time
        drain()
    return dict(elapsed=time.perf_counter() - t0, trials_built=ev.trials_built, fits=ev.fits,
                trial_steps=ev.trial_steps, run_s=round(ev.run_seconds, 3), fit_s=round(ev.fit_seconds, 3),
                wait_s=round(ev.wait_seconds, 3), alg_bytes=ev.alg_bytes)


def plain_search(ev, thetas, seeds, F, batch=32, log=None):
    """Every trial at fidelity F, `batch` trials at a time (the next batch is
    built on the host while the current one runs)."""
    t0 = time.perf_counter()
    trials = [Trial(th, s) for th, s in zip(thetas, seeds)]
    scores = np.empty(len(trials))
    for a in range(0, len(trials), batch):
        part = trials[a:a + batch]
        ev.prefetch(trials[a + batch:a + 2 * batch])
        scores[a:a + batch] = ev.advance(part, F)
        ev.release(part)
        if log:
            log(f"plain {a + len(part)}/{len(trials)} best {scores[:a + len(part)].min():.4f} "
                f"t {time.perf_counter() - t0:.1f}s")
    out = _stats(ev, t0)
    out.update(best=float(scores.min()), best_trial=int(scores.argmin()), scores=scores)
    return out


def _best_at(trials, F):
    s = [t.scores[F] for t in trials if F in t.scores]
    return min(s) if s else math.inf


def halving_search(ev, thetas, seeds, rungs, target=None, eta=3, batch=32, log=None):
    """Successive halving without a surrogate: every trial at the first rung,
    the best 1/eta of each rung continue to the next; at the top rung trials
    run best-first, four at a time, until one is within 2 % of target."""
    t0 = time.perf_counter()
    alive = [Trial(th, s) for th, s in zip(thetas, seeds)]
    for ri, F in enumerate(rungs):
        top = ri == len(rungs) - 1
        step = 4 if (top and target is not None) else batch
        for a in range(0, len(alive), step):
            part = alive[a:a + step]
            ev.prefetch(alive[a + step:a + 2 * step])
            ev.advance(part, F)
            if top and target is not None and _best_at(alive, F) <= 1.02 * target:
                break
        if log:
            log(f"halving rung {F}: {len(alive)} trials, best {_best_at(alive, F):.4f} "
                f"t {time.perf_counter() - t0:.1f}s")
        if top:
            break
        alive.sort(key=lambda t: t.scores[F])
        keep = max(1, len(alive) // eta)
        ev.release(alive[keep:])
        alive = alive[:keep]
    best = _best_at(alive, rungs[-1])
    ev.release(alive)
    out = _stats(ev, t0)
    out.update(best=best)
    return out


def step_cost(th):
    """Time of one step of a batched launch of four trials like th, in us, as
    measured for N=2048: a fixed 4.6 us (spike exchange, neuron tail, launch
    share), the GL history at about 1 us per 200 lags, and the row pull, which
    grows with the connection probability once rows are dense enough for the
    column-bitmap pull (0.3 us at p = 0.01, 3 us at 0.05, 4.5-7 us at 0.2)."""
    th = np.atleast_2d(th)
    pull = np.maximum(0.3, 3.0 + 2.2 * np.log2(th[:, 3] / 0.05))
    return 4.6 + th[:, 4] / 200.0 + pull


def _levy_steps(rng, count, dim, mu=1.3, lo=1e-3, hi=0.3):
    """Step lengths from a power law p(s) ~ s^-mu truncated to [lo, hi]
    (inverse transform), times random unit directions."""
    a = 1.0 - mu
    u = rng.random(count)
    s = (lo ** a + u * (hi ** a - lo ** a)) ** (1.0 / a)
    d = rng.standard_normal((count, dim))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return d * s[:, None]


def surrogate_search(ev, rungs, target=None, budget=512, warm=96, wave=32, promote=6, promote_warm=4, pool=2048, seed=1,
                     epochs=3000, round_epochs=500, lr=1e-3, reseed=6, reseed_top=3, reseed_direct=True,
                     explore=8, top_weight=4, clip=2.0, novelty=0.5, levy_hi=0.15, low_rungs_warm_only=True,
                     cost_weight=0.15, log=None):
    """Surrogate-guided search with the fidelity rungs.

    Every trial first runs to the middle rung (the lowest rung is a checkpoint
    of the same run and is recorded too).  The surrogate is trained on all
    (hyperparameters, fidelity, log10 NRMSE) records, the NRMSE clipped at
    `clip` (a diverged trial's size carries no ranking information) and each
    top-rung record counted `top_weight` times; the network is trained once
    for `epochs` and then continued for `round_epochs` per round.  Of each
    wave, half of the `promote` trials that continue to the top rung
    (`promote_warm` in the first wave) are the best by measured middle-rung
    score and half the best by the surrogate's top-rung prediction; the others
    are pruned.  The defaults fill whole launches at
    N=2048, where one launch holds four trials (one workgroup per CU): 4 in the
    first wave, 6 + `reseed` = 12 in later ones; a launch costs about the same
    with two trials as with four, since its steps are latency-bound.
    The first wave holds `warm` trials from a Latin hypercube; each later wave
    holds `wave` proposals:
      - `reseed` fresh weight seeds for the `reseed_top` best hyperparameter
        sets so far (a trial's score varies strongly with its weight draw);
        with `reseed_direct` they go to the top rung without the middle-rung
        gate (whose score ranks the best low-lambda sets poorly: 3900 training
        rows for N=2048 features overfit them);
      - `explore` global samples, half unscreened and half the best of `pool`
        by predicted score minus `novelty` x distance to the nearest trial run;
      - the rest from its ranking of `pool` Levy-flight moves around the eight
        best top-rung trials.
    Global and Levy proposals are ranked by predicted log10 NRMSE plus
    `cost_weight` x log2 of their step cost (step_cost): of two candidates
    predicted alike, the cheaper one runs.
    With `low_rungs_warm_only`, rungs below the middle one are scored for the
    first wave only (they tell the surrogate how scores change with fidelity);
    later waves run straight to the middle rung.
    Every reported score comes from a full run; the search stops when a
    top-rung score is within 2 % of target or after `budget` trials."""
    t0 = time.perf_counter()
    rng = np.random.default_rng(seed)
    f_min, f_max = rungs[0], rungs[-1]
    lower = list(rungs[:-1])
    mid = lower[-1]
    net = sa.Surrogate(7, 64, seed=seed)
    X, Y = [], []
    done = []           # (theta, top-rung score)
    history = []        # (theta, seed, {fidelity: score}) of every trial
    best = math.inf
    ymax = math.log10(clip)
    train_s = 0.0

    def record(ts, fids):
        for t in ts:
            for F in fids:
                if F in t.scores:
                    for _ in range(top_weight if F == f_max else 1):
                        X.append(encode(t.theta, F, f_min, f_max)[0])
                        Y.append(min(math.log10(max(t.scores[F], 1e-6)), ymax))

    def train(ep):
        nonlocal train_s
        a = time.perf_counter()
        net.train(np.array(X), np.array(Y, np.float32), epochs=ep, lr=lr)
        train_s += time.perf_counter() - a

    def seeds(k):
        return rng.integers(1 << 31, size=k)

    def rank(v):
        return np.argsort(np.argsort(v))

    def run_wave(new, ep, forced=0, fids=None, promote=promote):
        nonlocal best
        # the first `forced` trials go to the top rung whatever their middle-rung score
        fids = lower if fids is None else fids
        for a in range(0, len(new), 32):
            part = new[a:a + 32]
            ev.prefetch(new[a + 32:a + 64])
            for F in fids:
                ev.advance(part, F)
        record(new, fids)
        train(ep)
        pred = net.predict(encode(np.array([t.theta for t in new]), f_max, f_min, f_max))
        meas = np.array([t.scores[mid] for t in new])
        meas[:forced] = -np.inf
        pred[:forced] = -np.inf
        chosen = list(np.argsort(meas, kind="stable")[:forced + promote // 2])
        chosen += [i for i in np.argsort(pred, kind="stable") if i not in chosen][:forced + promote - len(chosen)]
        up = [new[i] for i in chosen]
        ev.release([t for t in new if t not in up])
        ev.advance(up, f_max)
        record(up, [f_max])
        done.extend((t.theta, t.scores[f_max]) for t in up)
        history.extend((t.theta.tolist(), t.seed, dict(t.scores)) for t in new)
        best = min(best, _best_at(up, f_max))
        ev.release(up)

    count = warm
    run_wave([Trial(th, s) for th, s in zip(latin_hypercube(warm, int(rng.integers(1 << 31))), seeds(warm))], epochs,
             promote=promote_warm)
    rounds = 0
    if log:
        log(f"surrogate warm-up: {count} trials, best {best:.4f} t {time.perf_counter() - t0:.1f}s "
            f"(training {train_s:.2f}s)")
    while not (target is not None and best <= 1.02 * target) and count < budget:
        rounds += 1
        train(round_epochs)
        inc = sorted(done, key=lambda d: d[1])
        pick = [inc[q % min(reseed_top, len(inc))][0] for q in range(reseed)]
        glob = sample_trials(pool, int(rng.integers(1 << 31)))
        pg = net.predict(encode(glob, f_max, f_min, f_max))
        # screened global picks favour unexplored regions: predicted score minus
        # `novelty` x distance (unit cube) to the nearest trial run so far
        seen = to_unit(np.array([h[0] for h in history]))
        gz = to_unit(glob)
        dist = np.sqrt(((gz[:, None, :] - seen[None, :, :]) ** 2).sum(axis=2)).min(axis=1)
        acq = pg - novelty * dist + cost_weight * np.log2(step_cost(glob))
        pick += list(glob[:explore // 2])                                   # unscreened
        pick += [glob[i] for i in np.argsort(acq) if i >= explore // 2][:explore - explore // 2]
        base = to_unit(np.array([d[0] for d in inc[:8]]))
        loc = from_unit(base[rng.integers(len(base), size=pool)] + _levy_steps(rng, pool, 6, hi=levy_hi))
        pl = net.predict(encode(loc, f_max, f_min, f_max)) + cost_weight * np.log2(step_cost(loc))
        pick += [loc[i] for i in np.argsort(pl)[:max(0, wave - len(pick))]]
        run_wave([Trial(th, s) for th, s in zip(pick, seeds(len(pick)))], round_epochs,
                 forced=reseed if reseed_direct else 0, fids=[mid] if low_rungs_warm_only else None)
        count += len(pick)
        if log:
            log(f"surrogate round {rounds}: {count} trials, best {best:.4f} t {time.perf_counter() - t0:.1f}s "
                f"(training {train_s:.2f}s)")
    net.close()
    out = _stats(ev, t0)
    out.update(best=best, rounds=rounds, trials=count, train_s=round(train_s, 3), history=history)
    return out


def spearman(a, b):
    ra = np.argsort(np.argsort(a)).astype(float)
    rb = np.argsort(np.argsort(b)).astype(float)
    ra -= ra.mean()
    rb -= rb.mean()
    return float((ra * rb).sum() / math.sqrt((ra * ra).sum() * (rb * rb).sum()))
