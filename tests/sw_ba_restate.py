"""Python restatement of the device small-world and scale-free construction
(spires_ext_build_topology, test helper).

The device generator takes the reference's steps (init_weights,
reservoir.c:399-536) with SplitMix64 draws instead of rand():
- small world: ring of K/2 forward neighbours; each ring edge (i, i+s), s in
  order, is rewired when unit0(key(1, i, s, 0)) < 0.1: the old edge is
  dropped and the first attempt a = 0, 1, ... whose target
  c = int(unit0(key(2, i, s, a)) * n) is neither i nor an edge of row i wins;
  after 10 n attempts the draw a = 10 n is kept (added unless it is i), as the
  reference's fail-safe keeps its last draw.
- scale free: seed clique on m0 nodes (pairs in lexicographic order), then
  node v draws partners with attempt t: r = unit0(key(3, v, 0, t)) *
  (2 E_v + k + v); r < v picks node int(r), otherwise endpoint position
  q = int(r - v) of the edges among 0..v-1 (even: the edge's first node, odd:
  its second) or, past them, one of v's own k partners (a duplicate, rejected);
  attempts from 50 m on draw uniformly, int(unit0(key(4, v, 0, t - 50 m)) * v).
  Each undirected edge is oriented lower -> higher when
  unit0(pair_key(lo, hi, 3)) < 0.5.
Weights are edge_weight(i, j) of the random generator; the device resolves the
scale-free draws in rounds, but each draw is a function of (v, t) and of
earlier edges, so the sequential order here gives the same graph.  Pure
Python loops: small n only."""
import numpy as np

from oracle import pyoracle as po
from tests import er_restate as er

M64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15
INV53 = 1.0 / 9007199254740992.0


def mix64(z):
    z = (z + GOLDEN) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def draw_key(seed, tag, i, s, a):
    return mix64(seed ^ mix64((tag << 60) | (i << 45) | (s << 30) | a))


def unit0(h):
    return (h >> 11) * INV53


def pair_key(seed, n, i, j, stream):
    return mix64(seed ^ mix64((((i * n + j) << 2) | stream) & M64))


def edge_weight(seed, n, ei, i, j):
    mag = ((pair_key(seed, n, i, j, 1) >> 11) + 1.0) * INV53
    return mag if unit0(pair_key(seed, n, i, j, 2)) < ei else -mag


def _csr(n, rows, seed, ei):
    rp = np.zeros(n + 1, np.int64)
    cols, vals = [], []
    for i in range(n):
        cs = sorted(rows[i])
        cols.extend(cs)
        vals.extend(edge_weight(seed, n, ei, i, j) for j in cs)
        rp[i + 1] = len(cols)
    return rp, np.array(cols, np.int32), np.array(vals, np.float64)


def small_world_unscaled(n, conn, ei, seed):
    rows = [set() for _ in range(n)]
    if n >= 3:
        K = int(conn * (n - 1) + 0.5)
        K = max(K, 2)
        K += K % 2
        K = min(K, n - 1)
        half = K // 2
        limit = 10 * n
        for i in range(n):
            row = rows[i]
            for s in range(1, half + 1):
                j = (i + s) % n
                if j != i:
                    row.add(j)
            for s in range(1, half + 1):
                if not unit0(draw_key(seed, 1, i, s, 0)) < 0.1:
                    continue
                row.discard((i + s) % n)
                a = 0
                while True:
                    c = min(int(unit0(draw_key(seed, 2, i, s, a)) * n), n - 1)
                    if a == limit or (c != i and c not in row):
                        break
                    a += 1
                if c != i:
                    row.add(c)
    return _csr(n, rows, seed, ei)


def scale_free_edges(n, conn):
    """(m, m0, expected undirected edge count) as reservoir.c:451-463."""
    m = int(conn * (n - 1) + 0.5)
    m = min(max(m, 1), n - 1)
    m0 = m if m > 2 else 2
    m0 = min(m0, n - 1)
    return m, m0, m0 * (m0 - 1) // 2 + (n - m0) * m


def scale_free_unscaled(n, conn, ei, seed):
    rows = [set() for _ in range(n)]
    if n >= 2:
        m, m0, _ = scale_free_edges(n, conn)
        src, tgt = [], []
        for u in range(m0):
            for v in range(u + 1, m0):
                src.append(u)
                tgt.append(v)
        guard = 50 * m
        for v in range(m0, n):
            Eb = len(src)
            mine = []
            t = 0
            while len(mine) < m:
                k = len(mine)
                if t < guard:
                    total = 2.0 * Eb + k + v
                    r = unit0(draw_key(seed, 3, v, 0, t)) * total
                    if r < v:
                        u = min(int(r), v - 1)
                    else:
                        q = min(int(r - v), 2 * Eb + k - 1)
                        if q < 2 * Eb:
                            e = q >> 1
                            u = tgt[e] if q & 1 else src[e]
                        else:
                            u = mine[q - 2 * Eb]
                else:
                    u = min(int(unit0(draw_key(seed, 4, v, 0, t - guard)) * v), v - 1)
                t += 1
                if u not in mine:
                    mine.append(u)
            for u in mine:
                src.append(v)
                tgt.append(u)
        for x, y in zip(src, tgt):
            lo, hi = min(x, y), max(x, y)
            if unit0(pair_key(seed, n, lo, hi, 3)) < 0.5:
                rows[lo].add(hi)
            else:
                rows[hi].add(lo)
    return _csr(n, rows, seed, ei)


def build(topology, n, conn, ei, seed, rho):
    """(row_ptr, col_idx, values, lambda) as spires_ext_build_topology returns them."""
    if topology == "random":
        return er.random_er(n, conn, ei, seed, rho)
    f = small_world_unscaled if topology == "small_world" else scale_free_unscaled
    rp, ci, val = f(n, conn, ei, seed)
    lam = po.csr_spectral_radius(rp, ci, val)
    if lam > 1e-9:
        val = val * (rho / lam)
    return rp, ci, val, lam
