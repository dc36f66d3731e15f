"""Reservoir construction on the GPU.

The device power iteration restates the reference's (sparse.c:93-127) with
the same operations in the same order, so its lambda must equal the oracle's
bit for bit, on the reference's own random, small-world and scale-free
matrices and on the cases where the iteration stops early (a non-negative
matrix) or at once (no edges).

The device Erdos-Renyi generator and the W_in / W_out streams of
spires_ext_create_device are pure functions of the seed; they must equal the
numpy restatement in tests/er_restate.py bit for bit (whose law is checked
against the reference's construction in tests/test_er_restate.py), including
p = 0, p = 1 and n = 1.  A reservoir built on the device then runs bit-exact
against the oracle given the same weights, and its ridge readout forecasts as
well as one built by the reference's rand() construction."""
import struct
import time

import numpy as np
import pytest

import spires_amd as sa
from spires_amd import workloads
from oracle import pyoracle as po
from tests import er_restate as er
from tests import sw_ba_restate as sb

pytestmark = pytest.mark.gpu


def bits(x):
    return struct.pack("<d", x)


def ref_weights(n, conn_type, conn, seed):
    sa.srand(seed)
    rp, ci, val, _, _ = sa.build_weights(sa.make_config(n, conn=conn, conn_type=conn_type, rho=0.9))
    return rp, ci, val


@pytest.mark.parametrize("n,conn_type,conn", [(512, "random", 0.1), (2048, "small_world", 0.01),
                                              (4096, "scale_free", 0.002)])
def test_power_iteration_bit_exact(n, conn_type, conn):
    rp, ci, val = ref_weights(n, conn_type, conn, 5)
    lam, iters = sa.spectral_radius_device(rp, ci, val)
    lam_o = po.csr_spectral_radius(rp, ci, val)
    print(f"n={n} {conn_type}: lambda={lam!r} iterations={iters}", flush=True)
    assert bits(lam) == bits(lam_o)
    assert bits(lam) == bits(sa.spectral_radius(rp, ci, val))
    assert 1 <= iters <= 1000


def test_power_iteration_early_stop_and_empty():
    rp, ci, val = ref_weights(1024, "random", 0.05, 9)
    pos = np.abs(val)                 # non-negative: Perron root, converges in a few steps
    lam, iters = sa.spectral_radius_device(rp, ci, pos)
    assert bits(lam) == bits(po.csr_spectral_radius(rp, ci, pos))
    assert iters < 100
    n = 300
    rp0 = np.zeros(n + 1, np.int64)
    lam0, it0 = sa.spectral_radius_device(rp0, np.zeros(0, np.int32), np.zeros(0))
    assert lam0 == 0.0 and it0 == 1
    assert po.csr_spectral_radius(rp0, np.zeros(0, np.int32), np.zeros(0)) == 0.0


def test_power_iteration_rejects_oversize():
    n = 20000
    with pytest.raises(RuntimeError):
        sa.spectral_radius_device(np.zeros(n + 1, np.int64), np.zeros(0, np.int32), np.zeros(0))


@pytest.mark.parametrize("n,p,ei,seed", [(2048, 0.05, 0.8, 11), (1000, 0.2, 0.5, 12), (4096, 0.01, 1.0, 13),
                                         (777, 0.0, 0.8, 1), (300, 1.0, 0.3, 2), (1, 0.5, 0.8, 3)])
def test_random_er_bit_exact(n, p, ei, seed):
    rho = 0.9
    rp, ci, val, lam = sa.random_er(n, p, seed=seed, rho=rho, ei=ei)
    rp_e, ci_e, val_e, lam_e = er.random_er(n, p, ei, seed, rho)
    print(f"n={n} p={p} ei={ei}: nnz={len(ci)} lambda={lam!r}", flush=True)
    assert np.array_equal(rp, rp_e) and np.array_equal(ci, ci_e)
    assert bits(lam) == bits(lam_e)
    assert np.array_equal(val.view(np.uint64), val_e.view(np.uint64))
    if p == 1.0:
        assert len(ci) == n * (n - 1)


def test_create_device_reservoir():
    """A discrete-LIF N=512, ER p=0.1, rho 0.9 reservoir built
    on the GPU runs bit-exact against the oracle on its own weights, and its
    one-step Mackey-Glass ridge readout is as good as a rand()-built one's."""
    N, T = 512, 3000
    kw = dict(din=1, dout=1, rho=0.9, gain=1.0, conn=0.1, conn_type="random", neuron="lif_discrete",
              params=[0.0, 1.0, 0.1, 0.0])
    series = workloads.mackey_glass(T + 1, seed=0)
    u = series[:T].reshape(T, 1)
    y = series[1:T + 1].reshape(T, 1)
    r = sa.Reservoir(sa.make_config(N, **kw), device_seed=21)
    rp, ci, val = r.csr()
    W_in = r.W_in()
    rp_e, ci_e, val_e, _ = er.random_er(N, 0.1, 0.8, 21, 0.9)
    win_e, wout_e = er.io_weights(N, 1, 1, 0.8, 21)
    assert np.array_equal(rp, rp_e) and np.array_equal(ci, ci_e) and np.array_equal(val, val_e)
    assert np.array_equal(W_in, win_e) and np.array_equal(r.W_out().ravel(), wout_e)
    o = po.OracleReservoir(po.make_config(N, **kw), csr=(rp, ci, val), W_in=W_in)
    X, R, _ = r.run_record(u)
    Xo, Ro, _ = o.run_record(u, T)
    assert Ro.sum() > 0
    assert np.array_equal(R, Ro)
    assert np.array_equal(X.view(np.uint64), Xo.view(np.uint64))

    def nrmse(res):
        res.train_ridge(u, y, 1e-6)
        pred = np.asarray(res.run(u)).reshape(T, 1)
        return float(np.sqrt(np.mean((pred[T // 2:] - y[T // 2:]) ** 2)) / np.std(y[T // 2:]))

    e_dev = nrmse(sa.Reservoir(sa.make_config(N, **kw), device_seed=22))
    sa.srand(22)
    e_ref = nrmse(sa.Reservoir(sa.make_config(N, **kw)))
    print(f"one-step MG NRMSE: device-built {e_dev:.4g}, rand()-built {e_ref:.4g}", flush=True)
    assert np.isfinite(e_dev) and 0.5 * e_ref < e_dev < 2.0 * e_ref


@pytest.mark.parametrize("topo,n,conn,ei,seed", [
    ("small_world", 600, 20 / 599, 0.8, 31), ("small_world", 97, 0.5, 0.5, 32), ("small_world", 3, 0.5, 0.8, 33),
    ("small_world", 2, 0.5, 0.8, 34), ("scale_free", 500, 4 / 499, 0.8, 35), ("scale_free", 300, 0.05, 0.3, 36),
    ("scale_free", 2, 0.5, 0.8, 37), ("scale_free", 1, 0.5, 0.8, 38), ("random", 400, 0.05, 0.8, 39)])
def test_build_topology_bit_exact(topo, n, conn, ei, seed):
    """Device small-world and scale-free construction against the restatement
    in tests/sw_ba_restate.py (whose law tests/test_sw_ba_restate.py checks
    against the reference's construction): CSR, lambda and values bit for bit,
    the degenerate sizes (no edges below 3 or 2 neurons) included."""
    print(f"case: {topo} n={n} conn={conn:.4f}", flush=True)
    rp, ci, val, lam = sa.build_topology(topo, n, conn, seed=seed, rho=0.9, ei=ei)
    rp_e, ci_e, val_e, lam_e = sb.build(topo, n, conn, ei, seed, 0.9)
    print(f"  nnz={len(ci)} lambda={lam!r}", flush=True)
    assert np.array_equal(rp, rp_e) and np.array_equal(ci, ci_e)
    assert bits(lam) == bits(lam_e)
    assert np.array_equal(val.view(np.uint64), val_e.view(np.uint64))


@pytest.mark.parametrize("topo,k", [("small_world", 20), ("scale_free", 8)])
def test_create_device_sw_ba_reservoir(topo, k):
    """FLIF-GL reservoirs whose small-world or scale-free W is built on the
    device: the CSR equals the restatement, the run is bit-exact against the
    oracle on those weights."""
    N, L, T = 640, 64, 400
    params = [1.0, 0.0, 0.0, 20.0, 0.8, float(L), 0.05, 2.0]
    kw = dict(din=1, dout=1, rho=0.95, gain=2.0, conn=k / (N - 1), conn_type=topo, neuron="flif_gl", params=params)
    r = sa.Reservoir(sa.make_config(N, **kw), device_seed=41)
    rp, ci, val = r.csr()
    rp_e, ci_e, val_e, _ = sb.build(topo, N, k / (N - 1), 0.8, 41, 0.95)
    assert np.array_equal(rp, rp_e) and np.array_equal(ci, ci_e) and np.array_equal(val, val_e)
    u = np.random.default_rng(3).uniform(0, 0.5, (T, 1))
    o = po.OracleReservoir(po.make_config(N, **kw), csr=(rp, ci, val), W_in=r.W_in())
    X, R, _ = r.run_record(u)
    Xo, Ro, _ = o.run_record(u, T)
    print(f"{topo}: nnz {len(ci)}, spikes {int(R.sum())}", flush=True)
    assert Ro.sum() > 0
    assert np.array_equal(R, Ro)
    assert np.array_equal(X.view(np.uint64), Xo.view(np.uint64))
    r.close()


@pytest.mark.parametrize("topo,k", [("small_world", 20), ("scale_free", 8)])
def test_topology_construction_time(topo, k):
    """Device vs host (the reference's rand() order) construction at N=16384."""
    n = 16384
    t0 = time.perf_counter()
    rp, ci, val, lam = sa.build_topology(topo, n, k / (n - 1), seed=5, rho=0.9)
    t_dev = time.perf_counter() - t0
    sa.srand(5)
    t0 = time.perf_counter()
    rp_h, ci_h, _, _, _ = sa.build_weights(sa.make_config(n, conn=k / (n - 1), conn_type=topo, rho=0.9))
    t_host = time.perf_counter() - t0
    print(f"N={n} {topo}: device {t_dev:.3f} s (nnz {len(ci)}), host {t_host:.3f} s (nnz {len(ci_h)})", flush=True)
    assert len(ci) == len(ci_h)
    assert np.isfinite(lam) and lam > 0


def test_construction_time():
    """Device vs host construction of an N=16384, p=0.01 random reservoir."""
    kw = dict(din=1, dout=1, rho=0.9, conn=0.01, conn_type="random", neuron="lif_discrete",
              params=[0.0, 1.0, 0.1, 0.0])
    cfg = sa.make_config(16384, **kw)
    t0 = time.perf_counter()
    r = sa.Reservoir(cfg, device_seed=3)
    t_dev = time.perf_counter() - t0
    nnz = len(r.csr()[1])
    r.close()
    sa.srand(3)
    t0 = time.perf_counter()
    h = sa.Reservoir(cfg)
    t_host = time.perf_counter() - t0
    h.close()
    print(f"N=16384 p=0.01 (nnz {nnz}): device create {t_dev:.3f} s, host create {t_host:.3f} s", flush=True)
    assert t_dev < t_host
