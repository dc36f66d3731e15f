"""Host-side pieces of the hyperparameter search: trial sampling follows the
stated distributions and the surrogate encoding maps them onto [0, 1]."""
import math

import numpy as np

from spires_amd import agile_search as ag


def test_sample_trials_ranges_and_scales():
    th = ag.sample_trials(4000, seed=1)
    lo = [0.5, 0.3, 0.1, 0.01, 64, 1e-8]
    hi = [1.5, 1.0, 10.0, 0.2, 1024, 1e-2]
    for j in range(6):
        assert th[:, j].min() >= lo[j] and th[:, j].max() <= hi[j]
    assert np.all(th[:, 4] == np.floor(th[:, 4]))
    # log-uniform axes: the median sits at the geometric mean of the bounds
    for j in (2, 3, 5):
        gm = math.sqrt(lo[j] * hi[j])
        assert abs(math.log10(np.median(th[:, j]) / gm)) < 0.05
    # uniform axes: the median sits at the midpoint
    for j in (0, 1, 4):
        assert abs(np.median(th[:, j]) - 0.5 * (lo[j] + hi[j])) < 0.03 * (hi[j] - lo[j])


def test_encode_unit_range_and_fidelity_axis():
    th = ag.sample_trials(1000, seed=2)
    z = ag.encode(th, 6000, 2000, 18000)
    assert z.dtype == np.float32 and z.shape == (1000, 7)
    assert z[:, :6].min() >= 0.0 and z[:, :6].max() <= 1.0
    assert np.allclose(z[:, 6], 0.5)                       # 6000 is the geometric middle of 2000..18000
    assert np.allclose(ag.encode(th[:3], 2000, 2000, 18000)[:, 6], 0.0)
    assert np.allclose(ag.encode(th[:3], 18000, 2000, 18000)[:, 6], 1.0)


def test_spearman_known_values():
    a = np.arange(10.0)
    assert ag.spearman(a, a ** 3) == 1.0
    assert ag.spearman(a, -a) == -1.0


def test_latin_hypercube_strata_and_units():
    n = 64
    th = ag.latin_hypercube(n, seed=3)
    z = ag.to_unit(th)
    # every axis has exactly one sample per stratum [k/n, (k+1)/n) (T_mem is floored, so allow its edge)
    for j in (0, 1, 2, 3, 5):
        assert np.array_equal(np.sort(np.floor(z[:, j] * n).astype(int)), np.arange(n))
    assert np.all(th[:, 4] == np.floor(th[:, 4])) and th[:, 4].min() >= 64 and th[:, 4].max() <= 1024
    # to_unit / from_unit round trip on the continuous axes
    back = ag.to_unit(ag.from_unit(z))
    assert np.allclose(back[:, [0, 1, 2, 3, 5]], z[:, [0, 1, 2, 3, 5]])
