"""The device Erdos-Renyi construction, through its numpy restatement
(tests/er_restate.py), against the reference's construction in distribution.

The reference draws its random topology from rand() (init_weights,
reservoir.c:345-545): an edge per ordered pair i != j with probability
`connectivity`, weight magnitude U[0, 1] with sign + at probability
`ei_ratio`.  The hashed generator must give the same law: ascending columns
and no self loops, density within 6 sigma of p n (n - 1) and of the
reference's edge count, magnitudes U(0, 1] (KS against the uniform law and,
after dividing out each matrix's scale, two-sample KS against the reference's
own rescaled weights), the sign fraction, and a spectral radius where random
matrix theory puts it: at the mean row sum when the mean weight is positive
(a matrix whose row sums all equal c has eigenvalue c), inside the circular
law's sqrt(n p E[w^2]) when it is zero.  The seed fixes the matrix."""
import numpy as np
import pytest
from scipy import stats

import spires_amd as sa
from oracle import pyoracle as po
from tests import er_restate as er


@pytest.mark.parametrize("n,p,ei", [(2048, 0.05, 0.8), (1000, 0.2, 0.5), (4096, 0.01, 1.0)])
def test_er_law(n, p, ei):
    rp, ci, val = er.er_unscaled(n, p, ei, seed=11)
    rp2, ci2, val2 = er.er_unscaled(n, p, ei, seed=11)
    assert np.array_equal(ci, ci2) and np.array_equal(val, val2)
    _, ci3, _ = er.er_unscaled(n, p, ei, seed=12)
    assert not (len(ci3) == len(ci) and np.array_equal(ci3, ci))
    assert rp[0] == 0 and np.all(np.diff(rp) >= 0) and rp[-1] == len(ci) == len(val)
    rows = np.repeat(np.arange(n), np.diff(rp))
    assert np.all((ci >= 0) & (ci < n)) and not np.any(ci == rows)
    same = rows[1:] == rows[:-1]
    assert np.all(ci[1:][same] > ci[:-1][same])
    pairs = n * (n - 1)
    assert abs(len(ci) - p * pairs) < 6 * np.sqrt(pairs * p * (1 - p))
    mag = np.abs(val)
    assert mag.max() <= 1.0 and mag.min() > 0.0
    assert stats.kstest(mag, "uniform").pvalue > 1e-4
    assert abs(np.mean(val > 0) - ei) <= 6 * np.sqrt(ei * (1 - ei) / len(val)) + 1e-12
    lam = po.csr_spectral_radius(rp, ci, val)
    row_sum = np.bincount(rows, weights=val, minlength=n).mean()
    if ei > 0.5:
        assert abs(lam - row_sum) < 0.02 * row_sum
    else:
        bulk = np.sqrt(n * p / 3.0)
        assert 0.85 * bulk < lam < 1.2 * bulk


def test_er_matches_reference_construction_in_law():
    n, p, ei = 1024, 0.1, 0.8
    sa.srand(4)
    cfg = sa.make_config(n, conn=p, conn_type="random", rho=0.9, ei=ei)
    rp_r, ci_r, val_r, _, _ = sa.build_weights(cfg)
    rp_d, ci_d, val_d, lam = er.random_er(n, p, ei, seed=4, rho=0.9)
    pairs = n * (n - 1)
    assert abs(len(ci_d) - len(ci_r)) < 6 * np.sqrt(2 * pairs * p * (1 - p))
    a = np.abs(val_r) / np.abs(val_r).max()
    b = np.abs(val_d) / np.abs(val_d).max()
    assert stats.ks_2samp(a, b).pvalue > 1e-4
    sd = 6 * np.sqrt(ei * (1 - ei) * (1 / len(val_r) + 1 / len(val_d)))
    assert abs(np.mean(val_r > 0) - np.mean(val_d > 0)) < sd
    assert abs(po.csr_spectral_radius(rp_d, ci_d, val_d) - 0.9) < 1e-6   # the iteration stops at |d lambda| < 1e-6
    assert abs(po.csr_spectral_radius(rp_r, ci_r, val_r) - 0.9) < 1e-6


def test_io_weights_law():
    w_in, w_out = er.io_weights(4096, 3, 2, ei=0.8, seed=9)
    assert w_in.shape == (4096, 3) and w_out.shape == (8192,)
    for w, ei in ((w_in.ravel(), 0.8), (w_out, 0.5)):
        assert np.all(np.abs(w) <= 1.0) and np.all(w != 0.0)
        assert stats.kstest(np.abs(w), "uniform").pvalue > 1e-4
        assert abs(np.mean(w > 0) - ei) < 6 * np.sqrt(ei * (1 - ei) / len(w))
