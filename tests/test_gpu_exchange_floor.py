"""The measured exchange floor (spires_ext_exchange_floor): the persistent
kernels' spike exchange alone.  It must be a few microseconds, and, at the
workgroup count the N=4096 workload's persistent kernel launches, below the
time of a full step of that workload (the same exchange plus the neuron
work)."""
import numpy as np
import pytest

import spires_amd as sa
from spires_amd import workloads

pytestmark = pytest.mark.gpu


def test_exchange_floor_is_below_the_step():
    print("case: exchange floor, 256 and 64 workgroups", flush=True)
    f256 = sa.exchange_floor(256, 5000)
    f64 = sa.exchange_floor(64, 5000)
    cfg, csr, W_in, u = workloads.flifgl_small_world(T=3000)
    r = sa.Reservoir(cfg, csr=csr, W_in=W_in)
    wg = r.kernel_plan()["workgroups"]
    sa.exchange_floor(wg, 1000)   # warm
    fwg = sa.exchange_floor(wg, 5000)
    r.upload_inputs(u)
    r.run_resident(0, 1000)
    r.sync()
    import time
    t0 = time.perf_counter()
    r.run_resident(1000, 2000)
    r.sync()
    step = (time.perf_counter() - t0) / 2000 * 1e6
    print(f"exchange floor {f256:.3f} us (256 workgroups), {f64:.3f} us (64), {fwg:.3f} us ({wg}); "
          f"step {step:.3f} us", flush=True)
    assert 0.2 < f256 < 10.0 and 0.2 < f64 < 10.0
    assert fwg < step
    with pytest.raises(sa.SpiresError):
        sa.exchange_floor(100000, 10)
    r.close()
