"""Error path of the asynchronous persistent launch, forced from the host:
spires_ext_inject_exchange_timeout raises the kernel's own abort flag, the
kernel gives up at its first unfinished exchange wait (every wave leaves its
step loop; no wait is left unbounded), the enqueueing call still returns OK,
and the call that waits for the run reports SPIRES_ERR_INTERNAL.  The flag is
then cleared: after a reset the reservoir runs bit-exact against the oracle."""
import numpy as np
import pytest

import spires_amd as sa
from oracle import pyoracle as po
from spires_amd import workloads

pytestmark = pytest.mark.gpu


def test_injected_exchange_timeout_is_reported_once_and_cleared():
    print("case: injected exchange timeout, FLIF-GL N=1024 L=128 product-wave kernel", flush=True)
    n, T = 1024, 300
    params = workloads.flif_gl_params(T_mem=128)
    cfg, csr, W_in = workloads.gaussian_reservoir(n, "small_world", 20.0 / (n - 1), 0.95, params, 5, gain=2.0)
    u = np.random.default_rng(9).uniform(0.0, 0.5, (T, 1))
    r = sa.Reservoir(cfg, csr=csr, W_in=W_in)
    assert r.kernel_plan()["kernel"] == 2
    r.upload_inputs(u)
    r.reserve_records(T, sa.REC_STATE)
    r.run_resident(0, 50, sa.REC_STATE)
    r.sync()                                            # a normal run settles cleanly
    assert sa.lib().spires_ext_inject_exchange_timeout(r.h) == 0
    r.run_resident(50, 100, sa.REC_STATE)               # enqueued: no error yet
    with pytest.raises(sa.SpiresError) as e:
        r.sync()
    assert e.value.status == 3 and "timed out" in str(e.value)
    r.sync()                                            # reported once
    r.reset()
    r.run_resident(0, T, sa.REC_STATE)
    X, _, _ = r.read_records(0, T, states=True)
    o = po.OracleReservoir(po.make_config(n, din=1, dout=1, gain=2.0, neuron="flif_gl", params=params),
                           csr=csr, W_in=W_in)
    Xo, _, _ = o.run_record(u, T, raster=False)
    assert np.array_equal(X.view(np.uint64), Xo.view(np.uint64))
    r.close()
