"""Synthetic workloads of the benchmark configurations.

Each builder returns plain numpy arrays (recurrent CSR, W_in, input series)
plus the config tuple, so the same workload can be handed to the GPU engine
and, in tests and the CPU baseline, to the oracle.

Topologies come from the engine's own generator (the reference's
init_weights, reservoir.c:345-545, with glibc rand()); weight values are then
replaced by N(0,1) draws from a seeded numpy generator and rescaled to the
requested spectral radius with the reference's power-iteration rule
(sparse.c:93-127), because the reference itself can only draw +/-U[0,1]
weights.
"""
from __future__ import annotations

import numpy as np

from . import make_config, build_weights, spectral_radius, srand


# FLIF-GL parameter layout: [V_th, V_reset, V_rest, tau_m, alpha, T_mem, bias, t_ref]
def flif_gl_params(alpha=0.8, T_mem=512.0, bias=0.05, t_ref=2.0, tau_m=20.0, V_th=1.0):
    return [V_th, 0.0, 0.0, tau_m, alpha, float(T_mem), bias, t_ref]


def gaussian_reservoir(n, topology, conn, rho, params, seed, neuron="flif_gl", din=1, dout=1, gain=1.0,
                       dt=1.0, w_in_scale=1.0):
    """Topology from the reference generator, N(0,1) weights rescaled to rho,
    W_in ~ U[-w_in_scale, w_in_scale].  Returns (cfg_tuple, (rp, ci, val), W_in)."""
    srand(seed)
    rp, ci, _, _, _ = build_weights(make_config(n, din=din, dout=dout, conn=conn, conn_type=topology,
                                                neuron=neuron, params=params, dt=dt))
    rng = np.random.default_rng(seed)
    val = rng.standard_normal(len(ci))
    lam = spectral_radius(rp, ci, val)
    if lam > 1e-9:
        val = val * (rho / lam)
    W_in = rng.uniform(-w_in_scale, w_in_scale, (n, din))
    cfg = make_config(n, din=din, dout=dout, rho=rho, gain=gain, conn=conn, conn_type=topology, neuron=neuron,
                      params=params, dt=dt)
    return cfg, (rp, ci, val), W_in


def flifgl_small_world(n=4096, T=20000, T_mem=512, seed=1):
    """FP64 exact mode, single reservoir: FLIF-GL N=4096, L=512, Watts-Strogatz
    k=20 p=0.1, N(0,1) weights at rho=0.95, gain 2.0, input U[0,0.5],
    delayed-recall targets u[n-d], d=1..100."""
    params = flif_gl_params(alpha=0.8, T_mem=T_mem, bias=0.05, t_ref=2.0)
    cfg, csr, W_in = gaussian_reservoir(n, "small_world", 20.0 / (n - 1), 0.95, params, seed, din=1,
                                        dout=100, gain=2.0)
    return cfg, csr, W_in, flifgl_small_world_inputs(T, seed)


def flifgl_small_world_inputs(T, seed=1):
    """The input series of flifgl_small_world: u ~ U[0, 0.5], [T x 1].  Drawn
    sequentially from one generator, so a shorter series is a prefix of a
    longer one."""
    return np.random.default_rng(seed + 1000).uniform(0.0, 0.5, (T, 1))


def algorithmic_bytes_per_neuron_step(n, nnz, L):
    """Bytes one neuron-step must move in the FP64 exact step: the L-1 history
    samples, the new ring sample, the CSR row (int32 col + f64 value per edge,
    int64 row pointer), W_in, the spike byte and the state-record write."""
    return 8 * (L - 1) + 8 + (nnz / n) * 12 + 8 + 8 + 1 + 8


def delayed_recall_targets(u, delays):
    """y[t, j] = u[t - d_j] (zero before the series starts)."""
    T = u.shape[0]
    y = np.zeros((T, len(delays)))
    for j, d in enumerate(delays):
        y[d:, j] = u[:T - d, 0]
    return y


def mackey_glass(T, seed=0, tau=17.0, beta=0.2, gamma=0.1, n=10, dt=0.1, every=1.0, washout=1000):
    """Mackey-Glass dx/dt = beta x(t-tau)/(1+x(t-tau)^n) - gamma x, RK4 with
    step dt; the delayed term at the half step is the mean of its grid
    neighbours.  x0 = 1.2 + U[-0.1, 0.1] from the seed; history = x0.
    Sampled every `every` time units after a washout of `washout` samples."""
    rng = np.random.default_rng(seed)
    x0 = 1.2 + rng.uniform(-0.1, 0.1)
    lag = int(round(tau / dt))
    sub = int(round(every / dt))
    total = (T + washout) * sub + 1
    x = np.empty(total + lag)
    x[:lag + 1] = x0

    def f(xt, xd):
        return beta * xd / (1.0 + xd ** n) - gamma * xt

    for k in range(lag, lag + total - 1):
        xd0 = x[k - lag]
        xd1 = x[k - lag + 1]
        xdh = 0.5 * (xd0 + xd1)
        xt = x[k]
        k1 = f(xt, xd0)
        k2 = f(xt + 0.5 * dt * k1, xdh)
        k3 = f(xt + 0.5 * dt * k2, xdh)
        k4 = f(xt + dt * k3, xd1)
        x[k + 1] = xt + dt / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
    series = x[lag::sub][washout:washout + T]
    return series.copy()


def uneven_batch_specs(count=64, seed=2):
    """Hyperparameters of the batched workload: N log-uniform in [256, 8192],
    topology alternating Erdos-Renyi p=0.05 / Barabasi-Albert m=5, alpha
    U[0.5, 1], T_mem in {128, 256, 512, 1024}, rho U[0.7, 1.2]."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(round(float(np.exp(rng.uniform(np.log(256), np.log(8192))))))
        out.append(dict(n=n, topology="random" if k % 2 == 0 else "scale_free",
                        conn=0.05 if k % 2 == 0 else 5.0 / (n - 1), alpha=float(rng.uniform(0.5, 1.0)),
                        T_mem=int(rng.choice([128, 256, 512, 1024])), rho=float(rng.uniform(0.7, 1.2)),
                        seed=1000 + k))
    return out


def class_sequences(count=512, channels=13, classes=10, phi=0.9, offset_std=0.5, min_len=30, max_len=100, seed=3):
    """Sequences of AR(1) frames: per channel x_t = phi x_{t-1} + N(0,1), plus a
    fixed random mean offset (std offset_std) per class.  Returns the frames
    concatenated [T x channels], the sequence lengths and the labels."""
    rng = np.random.default_rng(seed)
    means = rng.normal(0.0, offset_std, (classes, channels))
    lens = rng.integers(min_len, max_len + 1, count)
    labels = rng.integers(0, classes, count)
    frames = []
    for L, c in zip(lens, labels):
        x = np.empty((L, channels))
        prev = np.zeros(channels)
        for t in range(L):
            prev = phi * prev + rng.standard_normal(channels)
            x[t] = prev
        frames.append(x + means[c])
    return np.concatenate(frames), lens, labels
