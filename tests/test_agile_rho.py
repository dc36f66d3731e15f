"""The AGILE trials (spires_amd.agile_search.Evaluator) scale N(0,1) weights by
rho / sqrt(N p) instead of rescaling W with the reference power iteration
(reservoir.c:548-557).  By the circular law the spectral radius of such a
matrix is rho to within a few percent; here it is checked with numpy's dense
eigenvalues on eight N=512 trials (within +-8 %; measured +2..+5.5 %, and
+1..+3 % at N=2048).  The reference power iteration would do worse on these
matrices: their leading eigenvalues are complex pairs, so its max-norm
estimate does not converge (printed for comparison)."""
import numpy as np

from oracle import pyoracle as po
from spires_amd import agile_search as ag


def test_trial_spectral_radius_is_rho():
    ev = ag.Evaluator.__new__(ag.Evaluator)   # only the weight generator is used (no GPU)
    ev.n = 512
    th = ag.sample_trials(8, seed=7)
    worst = 0.0
    for k in range(8):
        rp, ci, val, _ = ag.Evaluator._weights(ev, th[k], 100 + k)
        W = np.zeros((512, 512))
        W[np.repeat(np.arange(512), np.diff(rp)), ci] = val
        ev_ = np.linalg.eigvals(W)
        rad = float(np.abs(ev_).max())
        lead = ev_[np.argmax(np.abs(ev_))]
        pi = po.csr_spectral_radius(rp, ci, val.astype(np.float64))
        dev = (rad - th[k][0]) / th[k][0]
        print(f"rho {th[k][0]:.3f} p {th[k][3]:.4f}: radius {rad:.4f} ({dev:+.2%}), leading eigenvalue {lead:.3f}, "
              f"power-iteration estimate {pi:.4f} ({(pi - th[k][0]) / th[k][0]:+.2%})")
        worst = max(worst, abs(dev))
    assert worst < 0.08
