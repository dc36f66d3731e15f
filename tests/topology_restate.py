"""Python restatement of the reference's random weight construction
(init_weights, reservoir.c:345-545, with generate_weight and the rand()
helpers at reservoir.c:307-343), driving glibc rand() through ctypes.

TEST INFRASTRUCTURE: written from the reference source, independently of
oracle/spires_oracle.c, so that tests/test_oracle_pins.py can check the
oracle's rand() consumption order (W_in first, then the topology; ER one draw
per ordered pair; WS ring + rewiring; BA clique + preferential attachment +
random orientation) matrix for matrix.  Small n only (pure Python loops).
"""
import ctypes as C

import numpy as np

_libc = C.CDLL("libc.so.6")
RAND_MAX = 2147483647.0


def srand(seed):
    _libc.srand(C.c_uint(seed))


def urand01():
    return float(_libc.rand()) / RAND_MAX


def generate_weight(ei):                      # reservoir.c:308-317
    mag = float(_libc.rand()) / RAND_MAX
    return mag if float(_libc.rand()) / RAND_MAX < ei else -mag


def init_weights(n, din, conn_type, conn, ei):
    """(W [n x n] dense, W_in [n x din]) in the reference's draw order."""
    W_in = np.array([generate_weight(ei) for _ in range(n * din)]).reshape(n, din)
    W = np.zeros((n, n))
    if conn_type == "random":                 # reservoir.c:386-398
        for i in range(n):
            for j in range(n):
                if i != j and urand01() < conn:
                    W[i, j] = generate_weight(ei)
    elif conn_type == "small_world" and n >= 3:   # reservoir.c:399-443
        K = int(conn * (n - 1) + 0.5)
        K = max(K, 2)
        if K % 2:
            K += 1
        K = min(K, n - 1)
        half = K // 2
        for i in range(n):
            for s in range(1, half + 1):
                j = (i + s) % n
                if i != j:
                    W[i, j] = generate_weight(ei)
        for i in range(n):
            for s in range(1, half + 1):
                j_old = (i + s) % n
                if urand01() < 0.1:
                    W[i, j_old] = 0.0
                    attempts = 0
                    while True:
                        j_new = int(urand01() * n)
                        if j_new >= n:
                            j_new = n - 1
                        attempts += 1
                        if attempts > 10 * n:
                            break
                        if not (j_new == i or W[i, j_new] != 0.0):
                            break
                    if j_new != i:
                        W[i, j_new] = generate_weight(ei)
    elif conn_type == "scale_free" and n >= 2:    # reservoir.c:445-536
        m = int(conn * (n - 1) + 0.5)
        m = min(max(m, 1), n - 1)
        m0 = min(max(m, 2), n - 1)
        A = 1.0
        adj = np.zeros((n, n), dtype=bool)
        deg = [0] * n
        for u in range(m0):
            for v in range(u + 1, m0):
                adj[u, v] = adj[v, u] = True
                deg[u] += 1
                deg[v] += 1
        for v in range(m0, n):
            added = guard = 0
            while added < m and guard < 50 * m:
                total = 0.0
                for u in range(v):
                    total += float(deg[u]) + A
                if total <= 0.0:
                    total = float(v)
                r = urand01() * total
                acc, pick = 0.0, 0
                for u in range(v):
                    acc += float(deg[u]) + A
                    if r <= acc:
                        pick = u
                        break
                if pick != v and not adj[v, pick]:
                    adj[v, pick] = adj[pick, v] = True
                    deg[v] += 1
                    deg[pick] += 1
                    added += 1
                guard += 1
            while added < m:
                pick = min(int(urand01() * v), v - 1)
                if pick != v and not adj[v, pick]:
                    adj[v, pick] = adj[pick, v] = True
                    deg[v] += 1
                    deg[pick] += 1
                    added += 1
        for i in range(n):
            for j in range(i + 1, n):
                if adj[i, j]:
                    if urand01() < 0.5:
                        W[i, j] = generate_weight(ei)
                    else:
                        W[j, i] = generate_weight(ei)
    return W, W_in
