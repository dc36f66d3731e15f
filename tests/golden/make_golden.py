"""Writes tests/golden/neurons_ref.npz: input/output vectors of the reference's
own neuron models (src/neuron.c + src/neurons/*.c compiled from /root/reference
by oracle/Makefile.ref into oracle/_ref) and of its AGILE optimizer
(src/agile.c) on a fixed quadratic loss.  Run in a container that has
/root/reference; the committed fixture lets the oracle be checked without it."""
import ctypes as C
import os
import sys
import zlib

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import pyoracle as po  # noqa: E402

CASES = {
    "lif_discrete": ([0.0, 1.0, 0.1, 0.02], 1.0),
    "lif_bio": ([0.0, 1.0, 5.0, 0.9], 0.25),
    "flif_gl": ([1.0, 0.0, 0.0, 20.0, 0.8, 64.0, 0.05, 2.0], 1.0),
    "flif_gl_dt05": ([1.0, -0.2, 0.1, 10.0, 0.6, 12.0, 0.0, 0.0], 0.5),
    "flif_caputo": ([1.0, 0.05, 0.0, 0.5, 0.0, 2.0, 0.7, 2.0, 30.0], 1.0),
    "flif_diffusive": ([1.0, 0.0, 0.0, 10.0, 0.6, 30.0, 0.95], 1.0),
}


def agile_loss(theta_p, dim, ctx):
    x = np.ctypeslib.as_array(theta_p, shape=(dim,))
    return float(np.sum((x - 0.3) ** 2 * np.arange(1, dim + 1)))


def main():
    po.build()
    out = {}
    n, T = 6, 150
    for name, (params, dt) in CASES.items():
        neuron = name.replace("_dt05", "")
        micro = int(round(1 / dt))
        rng = np.random.default_rng(zlib.crc32(name.encode()))
        u = rng.uniform(-0.2, 0.6, size=(T, n))
        inp = np.repeat(u, micro, axis=0).copy()
        p = np.array(params, dtype=np.float64)
        V = np.empty((T * micro, n)); S = np.empty((T * micro, n))
        seeds = None
        if neuron == "flif_diffusive":
            seeds = (C.c_uint64 * n)(*[po.df_seed(i, p) for i in range(n)])
        po.ref().ref_neurons_run(po.NEURON[neuron], po.dptr(p), dt, n, T * micro, po.dptr(inp),
                                 po.dptr(V), po.dptr(S), seeds, None, 0)
        out[f"{name}/params"] = p
        out[f"{name}/dt"] = np.array([dt])
        out[f"{name}/u"] = u
        out[f"{name}/V"] = V[micro - 1::micro]
        out[f"{name}/S"] = (S[micro - 1::micro] != 0).astype(np.uint8)
    dim = 4
    th0 = np.full(dim, 0.8); lo = np.zeros(dim); hi = np.ones(dim)
    o = po.AgileOpts(1.3, 1e-3, 0.3, 15.0, 0.9, 1e-2, 40, 10, 1e-3, 1, 0xA91E, 0)
    best = np.empty(dim); rep = po.AgileReport()
    po.ref().ref_agile(po.dptr(th0), po.dptr(lo), po.dptr(hi), dim, po.LOSS_FN(agile_loss), None,
                       C.byref(o), po.dptr(best), C.byref(rep))
    out["agile/best"] = best
    out["agile/report"] = np.array([rep.best_loss, rep.best_iter, rep.total_evals, rep.total_iters,
                                    rep.exited_refinement], dtype=np.float64)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "neurons_ref.npz"), **out)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
