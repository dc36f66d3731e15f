#!/usr/bin/env python3
"""Benchmark of the reservoir step on one MI355X.

Workload (FP64 exact mode, single reservoir): FLIF-GL N=4096, T_mem=512
(L=512), dt=1, Watts-Strogatz ring k=20 p=0.1 with N(0,1) weights rescaled to
rho=0.95, input gain 2.0, one input channel u ~ U[0, 0.5].  A "step" is one
reservoir time step over all N neurons (recurrent CSR pull over the spike
vector, FLIF-GL update with the 511-lag GL history sum, state row written to
HBM).  Inputs, weights and the state record are resident in HBM before the
timed region.  value = N * steps / wall time of the timed steps
(neuron-steps/s).

The CPU baseline is the oracle's C restatement of the reference step
(oracle/liboracle.so, OpenMP over neurons) on a bounded sample of the same
workload, timed on this host.

With --copy-calibration a plain device-to-device copy of 256 MiB runs first:
its FETCH_SIZE / WRITE_SIZE under rocprofv3 --pmc calibrate the counters that
give the step kernel's HBM traffic.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--copy-calibration]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "neuron-steps/s FLIF-GL N=4096 L=512; AGILE GPU-s to matched val NRMSE"
HBM_PEAK_GBS = 8000.0


class HipEvents:
    """HIP events recorded on the reservoir's own stream."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [C.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.hip.hipEventDestroy.argtypes = [C.c_void_p]

    def create(self):
        e = C.c_void_p()
        assert self.hip.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.hip.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.hip.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return ms.value


def algorithmic_bytes_per_neuron_step(n, nnz, L):
    """Bytes one neuron-step must move in the FP64 exact step: the L-1 history
    samples, the new ring sample, the CSR row (int32 col + f64 value per edge,
    int64 row pointer), W_in, the spike byte and the state-record write."""
    return 8 * (L - 1) + 8 + (nnz / n) * 12 + 8 + 8 + 1 + 8


def measured_traffic_per_step():
    """HBM bytes per step of the step kernel from the committed rocprofv3 --pmc
    summary (FETCH_SIZE / WRITE_SIZE, copy-calibrated) for this workload."""
    f = os.path.join(ROOT, "profiles", "k_persist_tp_flifgl_n4096_L512.json")
    if not os.path.exists(f):
        return None, None
    with open(f) as fh:
        d = json.load(fh)
    return d["hbm_read_bytes_per_step"] + d["hbm_write_bytes_per_step"], d["source"]


def cpu_baseline(cfg, csr, W_in, u, budget_s=12.0):
    import numpy as np
    from oracle import pyoracle as po
    po.build()
    n = cfg[0].num_neurons
    params = [cfg[1][i] for i in range(8)]
    o = po.OracleReservoir(po.make_config(n, din=1, dout=cfg[0].num_outputs, gain=cfg[0].input_strength,
                                          neuron="flif_gl", params=params), csr=csr, W_in=W_in)
    # bounded sample of the same series: steps 0..L-1 (history ramp, untimed), then timed chunks of
    # steady-state steps (full 511-lag history) until the budget is spent
    L = 512
    o.run_record(u[:L], L, states=False, raster=False)
    chunk, done = 25, 0
    u = u[L:]
    t0 = time.perf_counter()
    while True:
        o.run_record(u[done:done + chunk], chunk, states=True, raster=False)
        done += chunk
        if time.perf_counter() - t0 >= budget_s or done + chunk > u.shape[0]:
            break
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    return {"value": n * done / dt, "unit": "neuron-steps/s", "cores": cores, "kind": "port",
            "sample": f"steps {L}..{L + done - 1} of the same N={n} L=512 workload (oracle C restatement, "
                      f"OpenMP over neurons, {dt:.1f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 256 MiB device-to-device (counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs one reservoir per device and has no "
              "multi-GPU path (see DESIGN.md for how it would shard).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import spires_amd as sa
    from spires_amd import workloads

    if not os.path.exists(sa.LIB_PATH):
        sa.build()
    if args.copy_calibration:
        import torch
        nbytes = 256 << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": nbytes}), flush=True)
        del src, dst
    K, W = args.steps, args.warmup
    cfg, csr, W_in, u = workloads.config1(T=K + W)
    n = cfg[0].num_neurons
    L = 512
    r = sa.Reservoir(cfg, csr=csr, W_in=W_in)
    r.upload_inputs(u)
    r.reserve_records(max(K, W), sa.REC_STATE)
    ev = HipEvents()
    stream = r.stream()
    r.run_resident(0, W, sa.REC_STATE)
    r.sync()
    e0, e1 = ev.create(), ev.create()
    ev.record(e0, stream)
    t0 = time.perf_counter()
    r.run_resident(W, K, sa.REC_STATE)
    ev.record(e1, stream)
    r.sync()
    wall = time.perf_counter() - t0
    gpu_ms = ev.elapsed_ms(e0, e1)
    ms_per_step = gpu_ms / K
    value = n * K / (gpu_ms * 1e-3)
    X, _, _ = r.read_records(K - 1, 1, states=True)
    nnz = len(csr[2])
    bps = algorithmic_bytes_per_neuron_step(n, nnz, L)
    achieved = bps * n / (ms_per_step * 1e-3) / 1e9
    traffic_step, traffic_src = measured_traffic_per_step()
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "neuron-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "FLIF-GL reservoir, FP64 exact mode: N=4096, T_mem=512 (L=512), dt=1, "
                               "Watts-Strogatz k=20 p=0.1, N(0,1) weights rho=0.95, gain 2.0, u~U[0,0.5]",
                   "neurons": n, "memory_len": L, "nnz": nnz, "time_steps": K, "mode": "exact",
                   "record": "state V [T x N] f64"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_step * K if traffic_step else None,
                     "traffic_source": traffic_src,
                     "algorithmic_bytes_per_launch": bps * n * K,
                     "kernel": "k_persist_tp<32> (one launch of K steps)"},
        "wall_s": wall,
        "final_state_checksum": float(np.sum(X)),
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(cfg, csr, W_in, u)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
