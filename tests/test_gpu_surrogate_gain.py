"""The surrogate-guided search's gain over the plain search, on a synthetic
objective with the traits measured on the FLIF-GL workload (a narrow optimum
near the upper gain and alpha bounds, a weight-seed spread of the score, and
middle rungs that penalise small ridge lambdas, which overfit few rows).

The search code and its HIP surrogate (7->64->64->1, Adam) run unchanged; only
the evaluator is replaced by one that scores a trial from its hyperparameters,
weight seed and fidelity and charges the steps it would have run.  The plain
search scores the same 512 sampled trials at the top rung; the surrogate
search must reach within 2 % of that best NRMSE with several times fewer
trial-steps, in most search seeds.  (The same comparison on real reservoirs
is `python bench.py --agile`.)"""
import math

import numpy as np
import pytest

from spires_amd import agile_search as ag

pytestmark = pytest.mark.gpu

RUNGS = (2000, 6000, 18000)
CENTRE = np.array([0.35, 0.95, 0.9, 0.5, 0.55, 0.35])   # unit-cube optimum
WIDTH = np.array([1.0, 1.6, 1.6, 0.7, 0.5, 1.0])


def objective(theta, seed, F):
    """Scores between 0.55 (the optimum) and about 1.05, a 6 % weight-seed
    spread, and a middle-rung penalty for high gain with small lambda: the
    rank correlation with the top rung is 0.53 at 6000 steps and about 0 at
    2000 steps over the plain trials (0.58 and 0.09 measured on FLIF-GL)."""
    z = ag.to_unit(theta)[0]
    d2 = float(np.sum(((z - CENTRE) * WIDTH) ** 2))
    base = 0.55 + 0.5 * (1.0 - math.exp(-d2 / 0.3))
    score = base * (1.0 + 0.06 * np.random.default_rng(int(seed)).normal())
    return score + 0.5 * z[2] * (1.0 - z[5]) * (RUNGS[-1] / F - 1.0) / 2.0


class SyntheticEvaluator:
    """The Evaluator interface used by the searches, scoring with objective()."""

    def __init__(self):
        self.trials_built = self.fits = self.trial_steps = 0
        self.run_seconds = self.fit_seconds = self.wait_seconds = self.alg_bytes = 0.0

    def prefetch(self, trials):
        pass

    def release(self, trials):
        pass

    def advance(self, trials, F):
        for t in trials:
            if t.done == 0:
                self.trials_built += 1
            if t.done < F:
                self.trial_steps += F - t.done
                t.done = F
            if F not in t.scores:
                t.scores[F] = objective(t.theta, t.seed, F)
                self.fits += 1
        return np.array([t.scores[F] for t in trials])

    def close(self):
        pass


def test_surrogate_search_matches_plain_best_with_fewer_steps():
    th = ag.sample_trials(512, seed=2024)
    seeds = np.arange(512) + 10000
    ev = SyntheticEvaluator()
    plain = ag.plain_search(ev, th, seeds, RUNGS[-1])
    target = plain["best"]
    plain_steps = plain["trial_steps"]
    assert plain_steps == 512 * RUNGS[-1]
    ratios = []
    for sd in range(1, 7):
        ev = SyntheticEvaluator()
        r = ag.surrogate_search(ev, RUNGS, target=target, seed=sd)
        matched = r["best"] <= 1.02 * target
        ratios.append(plain_steps / r["trial_steps"] if matched else 0.0)
        print(f"seed {sd}: best {r['best']:.4f} (target {target:.4f}) trials {r['trials']} "
              f"trial-steps {r['trial_steps'] / 1e6:.2f}M ratio {ratios[-1]:.2f}", flush=True)
    ratios.sort()
    median = 0.5 * (ratios[2] + ratios[3])
    print(f"median trial-step ratio {median:.2f}", flush=True)
    assert sum(x > 0 for x in ratios) >= 4
    assert median >= 3.0
