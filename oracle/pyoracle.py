"""ctypes binding of the CPU oracle (oracle/liboracle.so) and of the reference
subset build (oracle/_ref/libspires_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libspires_ref.so")
REFERENCE_ROOT = "/root/reference"

NEURON = {"lif_discrete": 0, "lif_bio": 1, "flif_caputo": 2, "flif_gl": 3, "flif_diffusive": 4}
SYNAPSE = {"simple": 0, "psc_homogeneous": 1, "psc_heterogeneous": 2, "fractional": 3,
           "fractional_multiexp": 4}
CONN = {"random": 0, "small_world": 1, "scale_free": 2}


class Config(C.Structure):
    """spires_reservoir_config (include/spires.h)."""
    _fields_ = [
        ("num_neurons", C.c_size_t), ("num_inputs", C.c_size_t), ("num_outputs", C.c_size_t),
        ("spectral_radius", C.c_double), ("ei_ratio", C.c_double), ("input_strength", C.c_double),
        ("connectivity", C.c_double), ("dt", C.c_double),
        ("connectivity_type", C.c_int), ("neuron_type", C.c_int),
        ("neuron_params", C.POINTER(C.c_double)),
        ("synapse_type", C.c_int), ("synapse_backend", C.c_int),
        ("synapse_params", C.POINTER(C.c_double)),
    ]


class AgileOpts(C.Structure):
    _fields_ = [("mu", C.c_double), ("eta_min", C.c_double), ("eta_max", C.c_double),
                ("patience0", C.c_double), ("patience_decay", C.c_double),
                ("refine_step", C.c_double), ("refine_max_iters", C.c_int),
                ("refine_no_improve", C.c_int), ("spsa_eps", C.c_double),
                ("reflect_at_bounds", C.c_int), ("seed", C.c_uint64), ("verbose", C.c_int)]


class AgileReport(C.Structure):
    _fields_ = [("best_loss", C.c_double), ("best_iter", C.c_int), ("total_evals", C.c_int),
                ("total_iters", C.c_int), ("exited_refinement", C.c_int)]


LOSS_FN = C.CFUNCTYPE(C.c_double, C.POINTER(C.c_double), C.c_int, C.c_void_p)

_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I64 = C.POINTER(C.c_int64)
_I32 = C.POINTER(C.c_int32)
_U8 = C.POINTER(C.c_uint8)
_SZ = C.c_size_t


def build(force: bool = False) -> None:
    """Compile the oracle (and the reference subset when /root/reference exists)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-B", "-C", HERE], check=True)
    if os.path.isdir(REFERENCE_ROOT) and (force or not os.path.exists(REF_PATH)):
        subprocess.run(["make", "-s", "-B", "-f", os.path.join(HERE, "Makefile.ref")], check=True)


def dptr(a):
    return None if a is None else a.ctypes.data_as(_D)


_lib = None
_libc = C.CDLL("libc.so.6")


def srand(seed: int) -> None:
    _libc.srand(C.c_uint(seed))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_create.restype = _P
        L.or_create.argtypes = [C.POINTER(Config)]
        L.or_create_csr.restype = _P
        L.or_create_csr.argtypes = [C.POINTER(Config), _SZ, _I64, _I32, _D, _D, _D]
        L.or_free.argtypes = [_P]
        L.or_reset.argtypes = [_P]
        L.or_step.argtypes = [_P, _D]
        L.or_compute_output.argtypes = [_P, _D]
        L.or_read_state.argtypes = [_P, _D]
        L.or_read_spikes.argtypes = [_P, _D]
        L.or_run_record.argtypes = [_P, _D, _SZ, _D, _U8, _D]
        L.or_train_ridge.argtypes = [_P, _D, _D, _SZ, C.c_double]
        L.or_ridge_solve.restype = C.c_int
        L.or_ridge_solve.argtypes = [_D, _D, _SZ, _SZ, _SZ, C.c_double, _D]
        L.or_train_rls.argtypes = [_P, _D, _D, _SZ, C.c_double, C.c_double]
        L.or_train_online.argtypes = [_P, _D, C.c_double]
        L.or_coarse_grain.restype = _P
        L.or_coarse_grain.argtypes = [_P, C.c_double]
        L.or_num_neurons.restype = _SZ
        L.or_num_neurons.argtypes = [_P]
        L.or_nnz.restype = _SZ
        L.or_nnz.argtypes = [_P]
        L.or_get_csr.argtypes = [_P, _I64, _I32, _D]
        L.or_get_dense.argtypes = [_P, _D]
        L.or_get_W_in.argtypes = [_P, _D]
        L.or_get_W_out.argtypes = [_P, _D]
        L.or_set_W_out.argtypes = [_P, _D]
        L.or_gl_coeffs.argtypes = [C.c_double, C.c_int, _D]
        L.or_csr_spectral_radius.restype = C.c_double
        L.or_csr_spectral_radius.argtypes = [_SZ, _I64, _I32, _D]
        L.or_init_weights_dense.restype = C.c_int
        L.or_init_weights_dense.argtypes = [_SZ, _SZ, C.c_int, C.c_double, C.c_double, _D, _D]
        L.or_lu_solve.restype = C.c_int
        L.or_lu_solve.argtypes = [_SZ, _D, _D, _D]
        L.or_df_seed.restype = C.c_uint64
        L.or_df_seed.argtypes = [_SZ, _D]
        L.or_probe_synapse.argtypes = [_P, _D, _SZ, _D]
        L.or_probe_csr_rows.argtypes = [_SZ, _I64, _I32, _D, _D, _D]
        L.or_probe_multiexp.restype = C.c_int
        L.or_probe_multiexp.argtypes = [_P, _D, _D]
        L.or_agile_optimize.restype = C.c_int
        L.or_agile_optimize.argtypes = [_D, _D, _D, C.c_int, LOSS_FN, C.c_void_p,
                                        C.POINTER(AgileOpts), _D, C.POINTER(AgileReport)]
        _lib = L
    return _lib


_ref = None


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        R = C.CDLL(REF_PATH)
        R.ref_neurons_run.restype = C.c_int
        R.ref_neurons_run.argtypes = [C.c_int, _D, C.c_double, _SZ, _SZ, _D, _D, _D,
                                      C.POINTER(C.c_uint64), _D, C.c_int]
        R.ref_psc_het_run.restype = C.c_int
        R.ref_psc_het_run.argtypes = [_D, _SZ, _D, C.c_int, C.c_double, C.c_double, _SZ, _D, _D, _D]
        R.ref_agile.restype = C.c_int
        R.ref_agile.argtypes = [_D, _D, _D, C.c_int, LOSS_FN, C.c_void_p, C.POINTER(AgileOpts),
                                _D, C.POINTER(AgileReport)]
        _ref = R
    return _ref


def make_config(n, din=1, dout=1, rho=0.9, ei=0.8, gain=1.0, conn=0.1, dt=1.0,
                conn_type="random", neuron="flif_gl", params=(), synapse="simple",
                backend=0, syn_params=None):
    """Config plus the owned parameter arrays (keep the returned tuple alive)."""
    np_arr = (C.c_double * max(1, len(params)))(*params)
    sp_arr = None
    if syn_params is not None:
        sp_arr = (C.c_double * max(1, len(syn_params)))(*syn_params)
    cfg = Config(n, din, dout, rho, ei, gain, conn, dt, CONN.get(conn_type, conn_type),
                 NEURON.get(neuron, neuron), C.cast(np_arr, _D),
                 SYNAPSE.get(synapse, synapse), backend,
                 C.cast(sp_arr, _D) if sp_arr is not None else None)
    return cfg, np_arr, sp_arr


class OracleReservoir:
    """Thin object wrapper over or_res."""

    def __init__(self, cfg_tuple, csr=None, W_in=None, W_out=None):
        self._keep = cfg_tuple
        cfg = cfg_tuple[0]
        L = lib()
        self.N, self.Din, self.Dout = cfg.num_neurons, cfg.num_inputs, cfg.num_outputs
        if csr is None:
            self.h = L.or_create(C.byref(cfg))
        else:
            rp, ci, val = csr
            rp = np.ascontiguousarray(rp, np.int64)
            ci = np.ascontiguousarray(ci, np.int32)
            val = np.ascontiguousarray(val, np.float64)
            W_in = np.ascontiguousarray(W_in, np.float64)
            wo = None if W_out is None else np.ascontiguousarray(W_out, np.float64)
            self.h = L.or_create_csr(C.byref(cfg), len(val), rp.ctypes.data_as(_I64),
                                     ci.ctypes.data_as(_I32), dptr(val), dptr(W_in), dptr(wo))
        if not self.h:
            raise RuntimeError("oracle create failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_free(self.h)
            self.h = None

    def reset(self):
        lib().or_reset(self.h)

    def step(self, u=None):
        lib().or_step(self.h, None if u is None else dptr(np.ascontiguousarray(u, np.float64)))

    def state(self):
        v = np.empty(self.N)
        lib().or_read_state(self.h, dptr(v))
        return v

    def spikes(self):
        s = np.empty(self.N)
        lib().or_read_spikes(self.h, dptr(s))
        return s

    def output(self):
        y = np.empty(self.Dout)
        lib().or_compute_output(self.h, dptr(y))
        return y

    def run_record(self, u, T, states=True, raster=True, outputs=False):
        u = None if u is None else np.ascontiguousarray(u, np.float64)
        if u is not None and u.size < T * self.Din:
            raise ValueError(f"run_record: {T} steps need {T * self.Din} input values, got {u.size}")
        X = np.empty((T, self.N)) if states else None
        R = np.empty((T, self.N), np.uint8) if raster else None
        Y = np.empty((T, self.Dout)) if outputs else None
        lib().or_run_record(self.h, dptr(u), T, dptr(X),
                            None if R is None else R.ctypes.data_as(_U8), dptr(Y))
        return X, R, Y

    def train_ridge(self, u, y, lam):
        u = np.ascontiguousarray(u, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        lib().or_train_ridge(self.h, dptr(u), dptr(y), len(u) // self.Din, lam)

    def train_rls(self, u, y, delta, lam):
        u = np.ascontiguousarray(u, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        lib().or_train_rls(self.h, dptr(u), dptr(y), len(u) // self.Din, delta, lam)

    def train_online(self, target, lr):
        lib().or_train_online(self.h, dptr(np.ascontiguousarray(target, np.float64)), lr)

    def csr(self):
        n = self.N
        nnz = lib().or_nnz(self.h)
        rp = np.empty(n + 1, np.int64)
        ci = np.empty(max(nnz, 1), np.int32)
        val = np.empty(max(nnz, 1))
        lib().or_get_csr(self.h, rp.ctypes.data_as(_I64), ci.ctypes.data_as(_I32), dptr(val))
        return rp, ci[:nnz], val[:nnz]

    def dense(self):
        W = np.empty((self.N, self.N))
        lib().or_get_dense(self.h, dptr(W))
        return W

    def W_in(self):
        W = np.empty((self.N, self.Din))
        lib().or_get_W_in(self.h, dptr(W))
        return W

    def W_out(self):
        W = np.empty((self.Dout, self.N))
        lib().or_get_W_out(self.h, dptr(W))
        return W

    def set_W_out(self, W):
        lib().or_set_W_out(self.h, dptr(np.ascontiguousarray(W, np.float64)))

    def coarse_grain(self, thr):
        q = OracleReservoir.__new__(OracleReservoir)
        q._keep = self._keep
        q.h = lib().or_coarse_grain(self.h, thr)
        q.N = lib().or_num_neurons(q.h)
        q.Din, q.Dout = self.Din, self.Dout
        return q


def probe_synapse(res, spikes):
    """Vector the row dot reads after each step's synapse_prepare, for a spike
    train [T x N] (the reservoir's own synapse state advances)."""
    spikes = np.ascontiguousarray(spikes, np.float64)
    out = np.empty_like(spikes)
    lib().or_probe_synapse(res.h, dptr(spikes), spikes.shape[0], dptr(out))
    return out


def probe_csr_rows(rp, ci, val, x):
    rp = np.ascontiguousarray(rp, np.int64)
    ci = np.ascontiguousarray(ci, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    out = np.empty(len(rp) - 1)
    lib().or_probe_csr_rows(len(rp) - 1, rp.ctypes.data_as(_I64), ci.ctypes.data_as(_I32), dptr(val), dptr(x),
                            dptr(out))
    return out


def probe_multiexp(res, cap=64):
    tau, c = np.empty(cap), np.empty(cap)
    k = lib().or_probe_multiexp(res.h, dptr(tau), dptr(c))
    return tau[:k], c[:k]


def gl_coeffs(alpha, n):
    c = np.empty(n)
    lib().or_gl_coeffs(alpha, n, dptr(c))
    return c


def init_weights_dense(n, din, conn_type, conn, ei):
    W = np.empty((n, n))
    Wi = np.empty((n, din))
    lib().or_init_weights_dense(n, din, CONN.get(conn_type, conn_type), conn, ei, dptr(W), dptr(Wi))
    return W, Wi


def ridge_solve(X, Y, lam):
    X = np.ascontiguousarray(X, np.float64)
    Y = np.ascontiguousarray(Y, np.float64)
    T, N = X.shape
    D = Y.shape[1]
    W = np.zeros((D, N))
    rc = lib().or_ridge_solve(dptr(X), dptr(Y), T, N, D, lam, dptr(W))
    return W, rc


def lu_solve(A, b):
    A = np.ascontiguousarray(A, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.empty(len(b))
    rc = lib().or_lu_solve(len(b), dptr(A), dptr(b), dptr(x))
    return x, rc


def csr_spectral_radius(rp, ci, val):
    rp = np.ascontiguousarray(rp, np.int64)
    ci = np.ascontiguousarray(ci, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    return lib().or_csr_spectral_radius(len(rp) - 1, rp.ctypes.data_as(_I64),
                                        ci.ctypes.data_as(_I32), dptr(val))


def df_seed(i, params):
    p = np.ascontiguousarray(params, np.float64)
    return lib().or_df_seed(i, dptr(p))
