"""The surrogate-guided search against the plain search on real FLIF-GL
reservoirs at a reduced scale (N=512, 128 sampled trials, rungs 500 / 1500 /
4500 steps, the Evaluator and search code of the config-4 workload, smaller
waves: 32 warm-up trials, waves of 16).  The plain search runs all 128 trials
to 4500 steps; each of eight search seeds stops when a top-rung score is within
2 % of the plain best.  Measured on MI355X: 6 of 8 seeds match, the median
needs 4.4x fewer trial-steps than the plain search (unmatched seeds count as
never); the assertion keeps a margin (5 of 8, 2.5x)."""
import math

import numpy as np
import pytest

from spires_amd import agile_search as ag

pytestmark = pytest.mark.gpu

RUNGS = (500, 1500, 4500)


def test_surrogate_search_beats_plain_on_real_reservoirs():
    print("case: N=512, 128 trials, rungs 500/1500/4500, 8 search seeds", flush=True)
    M = 128
    ev = ag.Evaluator(n=512, f_max=RUNGS[-1], seed=0)
    plain = ag.plain_search(ev, ag.sample_trials(M, seed=2024), np.arange(M) + 10000, RUNGS[-1])
    ev.close()
    target = plain["best"]
    steps = []
    for sd in range(1, 9):
        ev = ag.Evaluator(n=512, f_max=RUNGS[-1], seed=0)
        r = ag.surrogate_search(ev, RUNGS, target=target, budget=M, seed=sd, warm=32, wave=16, reseed=4, explore=4,
                                promote=4)
        ev.close()
        ok = r["best"] <= 1.02 * target
        steps.append(r["trial_steps"] if ok else math.inf)
        print(f"seed {sd}: best {r['best']:.4f} ({'matched' if ok else 'no'}), {r['trials']} trials, "
              f"{r['trial_steps']} trial-steps", flush=True)
    ratio = plain["trial_steps"] / float(np.median(steps))
    print(f"plain best {target:.4f} over {plain['trial_steps']} trial-steps; median ratio {ratio:.2f}, "
          f"matched {int(np.isfinite(steps).sum())}/8", flush=True)
    assert np.isfinite(steps).sum() >= 5
    assert ratio >= 2.5
