"""The exchange wave of the product-wave kernel keeps the inputs of 64 macro
steps in its lanes, loaded a window ahead.  A small-world FLIF-GL reservoir
whose rows all fit the register row pull (16 neurons per workgroup), with three
input channels, runs in launches that start and end off the 64-step windows;
rasters and states must equal the oracle's bit for bit (flif_gl.c:50-100,
reservoir.c:132-136 for W_in u)."""
import numpy as np
import pytest

import spires_amd as sa
from spires_amd import workloads
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def test_input_windows_across_launches_bit_exact():
    print("case: small world N=512 k=10, Din=3, launches 0+70, 70+100, 170+130", flush=True)
    n, T, din = 512, 300, 3
    params = workloads.flif_gl_params(T_mem=64)
    cfg, csr, W_in = workloads.gaussian_reservoir(n, "small_world", 10.0 / (n - 1), 0.95, params, seed=11, din=din,
                                                  dout=2, gain=1.5)
    u = np.random.default_rng(12).uniform(0.0, 0.5, (T, din))
    r = sa.Reservoir(cfg, csr=csr, W_in=W_in)
    plan = r.kernel_plan()
    # the register row pull: product-wave kernel, 16 neurons per workgroup, edge list, CSR in LDS
    assert (plan["kernel"], plan["nb"], plan["row_pull"], plan["weights_lds"]) == (2, 16, 0, 1), plan
    flags = sa.REC_STATE | sa.REC_RASTER | sa.REC_APPEND
    r.upload_inputs(u)
    r.reserve_records(T, flags)
    for t0, steps in ((0, 70), (70, 100), (170, 130)):
        r.run_resident(t0, steps, flags)
    X, R, _ = r.read_records(0, T, states=True, raster=True)
    o = po.OracleReservoir(po.make_config(n, din=din, dout=2, gain=1.5, neuron="flif_gl", params=params),
                           csr=csr, W_in=W_in)
    Xo, Ro, _ = o.run_record(u, T)
    assert R.sum() > 0
    assert np.array_equal(R, Ro)
    assert np.array_equal(np.ascontiguousarray(X).view(np.uint64), np.ascontiguousarray(Xo).view(np.uint64))
    r.close()
