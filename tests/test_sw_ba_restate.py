"""The device small-world and scale-free construction, restated in
tests/sw_ba_restate.py, against the reference's own construction
(spires_reservoir_create's rand() order, restated bit-exact in
csrc/host/topology.cpp and pinned in tests/test_oracle_pins.py).  The two use
different random streams, so they are compared in law: the small world's
exact out-degree K/2 and its rewired fraction (p = 0.1), the scale-free graph's
exact edge count, its degree distribution (two-sample KS) and its tail."""
import numpy as np
import pytest

import spires_amd as sa
from tests import sw_ba_restate as sb


def reference(n, conn_type, conn, seed):
    sa.srand(seed)
    rp, ci, val, _, _ = sa.build_weights(sa.make_config(n, conn=conn, conn_type=conn_type, rho=0.9))
    return rp, ci


def ks_stat(a, b):
    a, b = np.sort(a), np.sort(b)
    grid = np.union1d(a, b)
    fa = np.searchsorted(a, grid, side="right") / len(a)
    fb = np.searchsorted(b, grid, side="right") / len(b)
    return float(np.max(np.abs(fa - fb)))


def off_ring_fraction(rp, ci, n, half):
    rows = np.repeat(np.arange(n), np.diff(rp))
    off = (ci - rows) % n
    return float(np.mean((off < 1) | (off > half)))


@pytest.mark.parametrize("n,k", [(2048, 20), (1000, 6)])
def test_small_world_law(n, k):
    conn = k / (n - 1)
    rp, ci, _ = sb.small_world_unscaled(n, conn, 0.8, 7)
    rp_r, ci_r = reference(n, "small_world", conn, 7)
    half = k // 2
    # exact: every row keeps K/2 edges (a rewired edge is replaced by a new one), no self loops
    assert np.all(np.diff(rp) == half) and np.all(np.diff(rp_r) == half)
    rows = np.repeat(np.arange(n), np.diff(rp))
    assert not np.any(ci == rows)
    f, f_r = off_ring_fraction(rp, ci, n, half), off_ring_fraction(rp_r, ci_r, n, half)
    # a rewired edge lands off the ring unless it picks a ring slot again (about half/n)
    expect = 0.1 * (1 - half / n)
    sd = np.sqrt(0.1 * 0.9 / (n * half))
    print(f"n={n} K={k}: off-ring fraction restated {f:.4f}, reference {f_r:.4f}, expected {expect:.4f} +- {sd:.4f}")
    assert abs(f - expect) < 5 * sd and abs(f_r - expect) < 5 * sd
    indeg, indeg_r = np.bincount(ci, minlength=n), np.bincount(ci_r, minlength=n)
    assert indeg.mean() == indeg_r.mean() == half
    assert ks_stat(indeg, indeg_r) < 0.08


@pytest.mark.parametrize("n,m", [(2000, 4), (1500, 8)])
def test_scale_free_law(n, m):
    conn = m / (n - 1)
    rp, ci, _ = sb.scale_free_unscaled(n, conn, 0.8, 9)
    rp_r, ci_r = reference(n, "scale_free", conn, 9)
    _, _, E = sb.scale_free_edges(n, conn)
    assert len(ci) == E and len(ci_r) == E          # exact: m partners per new node over the clique
    rows = np.repeat(np.arange(n), np.diff(rp))
    assert not np.any(ci == rows)
    pairs = np.minimum(rows, ci).astype(np.int64) * n + np.maximum(rows, ci)
    assert len(np.unique(pairs)) == E                # no pair in both directions
    deg = np.bincount(rows, minlength=n) + np.bincount(ci, minlength=n)
    rows_r = np.repeat(np.arange(n), np.diff(rp_r))
    deg_r = np.bincount(rows_r, minlength=n) + np.bincount(ci_r, minlength=n)
    D = ks_stat(deg, deg_r)
    q99, q99_r = np.quantile(deg, 0.99), np.quantile(deg_r, 0.99)
    fo, fo_r = np.mean(rows < ci), np.mean(rows_r < ci_r)   # fair-coin orientation
    print(f"n={n} m={m}: degree KS {D:.4f}; 99th pct {q99:.1f} vs {q99_r:.1f}; max {deg.max()} vs {deg_r.max()}; "
          f"lower->higher {fo:.3f} vs {fo_r:.3f}")
    assert D < 0.07
    assert 0.67 < q99 / q99_r < 1.5 and 0.4 < deg.max() / deg_r.max() < 2.5
    sd = np.sqrt(0.25 / E)
    assert abs(fo - 0.5) < 5 * sd and abs(fo_r - 0.5) < 5 * sd
    _, m0, _ = sb.scale_free_edges(n, conn)
    assert deg[m0:].min() >= m and deg_r[m0:].min() >= m   # each new node brings m partners
