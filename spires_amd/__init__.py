"""spires_amd -- MI355X (gfx950) reservoir engine behind the spires C ABI.

This module is a thin ctypes binding of ``libspires_amd.so`` (built from
``spires_amd/csrc``): the drop-in ``include/spires.h`` API plus the
extensions of ``include/spires_ext.h``.  There is no CPU fallback: if the
HIP library is missing, ``lib()`` raises.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import subprocess
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspires_amd.so")

NEURON = {"lif_discrete": 0, "lif_bio": 1, "flif_caputo": 2, "flif_gl": 3, "flif_diffusive": 4}
SYNAPSE = {"simple": 0, "psc_homogeneous": 1, "psc_heterogeneous": 2, "fractional": 3,
           "fractional_multiexp": 4}
CONN = {"random": 0, "small_world": 1, "scale_free": 2}
BACKEND = {"sparse": 0, "dense": 1}
STATUS = {0: "SPIRES_OK", 1: "SPIRES_ERR_INVALID_ARG", 2: "SPIRES_ERR_ALLOC", 3: "SPIRES_ERR_INTERNAL"}

REC_STATE, REC_RASTER, REC_OUTPUT, REC_APPEND = 1, 2, 4, 8
TB_ONE_LAUNCH = 16   # TimeBlock: every reservoir in one launch of 64-neuron one-wave slices
MODE_EXACT, MODE_FAST = 0, 1
PATH_AUTO, PATH_PER_STEP, PATH_PERSISTENT, PATH_PERSISTENT_HBM, PATH_PERSISTENT_LDS = 0, 1, 2, 3, 4


class SpiresError(RuntimeError):
    def __init__(self, status, msg=""):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


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


class OptBudget(C.Structure):
    _fields_ = [("data_fraction", C.c_double), ("num_seeds", C.c_int), ("time_limit_sec", C.c_double),
                ("_reserved_i0", C.c_int), ("_reserved_d0", C.c_double)]


class OptScore(C.Structure):
    _fields_ = [("lambda_var", C.c_double), ("lambda_cost", C.c_double), ("metric", C.c_int),
                ("_reserved_i0", C.c_int)]


class OptResult(C.Structure):
    _fields_ = [("best_config", Config), ("best_log10_ridge", C.c_double), ("best_score", C.c_double),
                ("metric_mean", C.c_double), ("metric_std", C.c_double)]


LOSS_FN = C.CFUNCTYPE(C.c_double, C.POINTER(C.c_double), C.c_int, C.c_void_p)

_D = C.POINTER(C.c_double)
_I64 = C.POINTER(C.c_int64)
_I32 = C.POINTER(C.c_int32)
_U8 = C.POINTER(C.c_uint8)
_F = C.POINTER(C.c_float)
_P = C.c_void_p
_SZ = C.c_size_t

# Every symbol of include/spires.h and include/spires_ext.h: name -> (restype, argtypes)
SIGNATURES = {
    "spires_reservoir_create": (C.c_int, [C.POINTER(Config), C.POINTER(_P)]),
    "spires_reservoir_destroy": (None, [_P]),
    "spires_reservoir_reset": (C.c_int, [_P]),
    "spires_step": (C.c_int, [_P, _D]),
    "spires_run": (_D, [_P, _D, _SZ]),
    "spires_train_online": (C.c_int, [_P, _D, C.c_double]),
    "spires_train_ridge": (C.c_int, [_P, _D, _D, _SZ, C.c_double]),
    "spires_train_rls": (C.c_int, [_P, _D, _D, _SZ, C.c_double, C.c_double]),
    "spires_coarse_grain": (C.c_int, [_P, C.c_double, C.POINTER(_P)]),
    "spires_copy_reservoir_state": (_D, [_P]),
    "spires_read_reservoir_state": (C.c_int, [_P, _D]),
    "spires_read_spike_state": (C.c_int, [_P, _D]),
    "spires_copy_weights": (_D, [_P]),
    "spires_read_weights": (C.c_int, [_P, _D]),
    "spires_compute_output": (C.c_int, [_P, _D]),
    "spires_copy_readout": (_D, [_P]),
    "spires_read_readout": (C.c_int, [_P, _D]),
    "spires_num_neurons": (_SZ, [_P]),
    "spires_num_inputs": (_SZ, [_P]),
    "spires_num_outputs": (_SZ, [_P]),
    "spires_optimize": (C.c_int, [C.POINTER(Config), C.POINTER(OptBudget), C.c_int, C.POINTER(OptScore),
                                  C.POINTER(OptResult), _D, _D, _SZ]),
    "spires_ext_create_device": (C.c_int, [C.POINTER(Config), C.c_uint64, C.POINTER(_P)]),
    "spires_ext_create_csr": (C.c_int, [C.POINTER(Config), _SZ, _I64, _I32, _D, _D, _D, C.POINTER(_P)]),
    "spires_ext_run_record": (C.c_int, [_P, _D, _SZ, _D, _U8, _D]),
    "spires_ext_upload_inputs": (C.c_int, [_P, _D, _SZ]),
    "spires_ext_reserve_records": (C.c_int, [_P, _SZ, C.c_uint]),
    "spires_ext_run_resident": (C.c_int, [_P, _SZ, _SZ, C.c_uint]),
    "spires_ext_read_records": (C.c_int, [_P, _SZ, _SZ, _D, _U8, _D]),
    "spires_ext_run_resident_batch": (C.c_int, [C.POINTER(_P), C.c_int, _SZ, _SZ, C.c_uint]),
    "spires_ext_fit_ridge_records": (C.c_int, [_P, _SZ, _SZ, _D, C.c_double, _D]),
    "spires_ext_batch_workgroups": (C.c_int, [C.POINTER(_P), C.c_int]),
    "spires_ext_kernel_plan": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "spires_ext_inject_exchange_timeout": (C.c_int, [_P]),
    "spires_ext_exchange_floor": (C.c_int, [C.c_int, _SZ, C.POINTER(C.c_double)]),
    "spires_ext_timeblock_create": (C.c_int, [C.POINTER(_P), C.c_int, _SZ, C.c_uint, C.POINTER(_P)]),
    "spires_ext_timeblock_run": (C.c_int, [_P, _SZ]),
    "spires_ext_timeblock_launches": (C.c_int, [_P]),
    "spires_ext_timeblock_destroy": (None, [_P]),
    "spires_ext_predict_records": (C.c_int, [_P, _SZ, _SZ, _D]),
    "spires_ext_fit_ridge_records_batch": (C.c_int, [C.POINTER(_P), C.c_int, _SZ, _SZ, C.POINTER(_D), _D, _D]),
    "spires_ext_train_ridge_washout": (C.c_int, [_P, _D, _D, _SZ, _SZ, C.c_double, _D]),
    "spires_ext_set_mode": (C.c_int, [_P, C.c_int, C.c_int]),
    "spires_ext_set_kernel_path": (C.c_int, [_P, C.c_int]),
    "spires_ext_set_state": (C.c_int, [_P, _D]),
    "spires_ext_sync": (C.c_int, [_P]),
    "spires_ext_stream": (_P, [_P]),
    "spires_ext_nnz": (_SZ, [_P]),
    "spires_ext_get_csr": (C.c_int, [_P, _I64, _I32, _D]),
    "spires_ext_get_input_weights": (C.c_int, [_P, _D]),
    "spires_ext_last_error": (C.c_char_p, []),
    "spires_ext_shutdown": (None, []),
    "spires_ext_agile_optimize": (C.c_int, [_D, _D, _D, C.c_int, LOSS_FN, _P, _D, _D]),
    "spires_ext_optimize_evaluate": (C.c_int, [C.POINTER(Config), C.POINTER(OptBudget), C.POINTER(OptScore), _D,
                                               _D, _D, _SZ, _D]),
    "spires_ext_spectral_radius": (C.c_double, [_SZ, _I64, _I32, _D]),
    "spires_ext_mlp_create": (C.c_int, [C.c_int, C.c_int, C.c_ulonglong, C.POINTER(_P)]),
    "spires_ext_mlp_destroy": (None, [_P]),
    "spires_ext_mlp_num_params": (C.c_int, [_P]),
    "spires_ext_mlp_get_params": (C.c_int, [_P, _F]),
    "spires_ext_mlp_set_params": (C.c_int, [_P, _F]),
    "spires_ext_mlp_predict": (C.c_int, [_P, _F, C.c_int, _F]),
    "spires_ext_mlp_gradient": (C.c_int, [_P, _F, _F, C.c_int, _F, _F]),
    "spires_ext_mlp_train": (C.c_int, [_P, _F, _F, C.c_int, C.c_int, C.c_float, _F]),
    "spires_ext_mlp_last_error": (C.c_char_p, []),
    "spires_ext_random_er": (C.c_int, [_SZ, C.c_double, C.c_double, C.c_uint64, C.c_double, C.POINTER(_SZ),
                                       C.POINTER(_I64),
                                       C.POINTER(_I32), C.POINTER(_D), C.POINTER(C.c_double)]),
    "spires_ext_build_topology": (C.c_int, [C.c_int, _SZ, C.c_double, C.c_double, C.c_uint64, C.c_double,
                                            C.POINTER(_SZ), C.POINTER(_I64), C.POINTER(_I32), C.POINTER(_D),
                                            C.POINTER(C.c_double)]),
    "spires_ext_spectral_radius_device": (C.c_int, [_SZ, _I64, _I32, _D, C.POINTER(C.c_double),
                                                    C.POINTER(C.c_int)]),
    "spires_ext_spectral_radius_device_noalloc": (C.c_int, [_SZ, _P, _P, _P, _P, _P, _P]),
    "spires_ext_build_weights": (C.c_int, [C.POINTER(Config), C.POINTER(_SZ), C.POINTER(_I64), C.POINTER(_I32),
                                           C.POINTER(_D), C.POINTER(_D), C.POINTER(_D)]),
}

_lib = None
_libc = C.CDLL("libc.so.6")
_libc.free.argtypes = [_P]

# Live handle owners (reservoirs, time blocks, surrogates), closed at exit
# before the engine releases its process-wide GPU resources: nothing of ours
# is left for the HIP runtime's own teardown.
_live = weakref.WeakSet()


def shutdown() -> None:
    """Close every live handle, then release the engine's per-device rocBLAS
    handle and readout workspace (spires_ext_shutdown).  Runs at exit."""
    for obj in list(_live):
        try:
            obj.close()
        except Exception:
            pass
    if _lib is not None:
        _lib.spires_ext_shutdown()


atexit.register(shutdown)


def build(force: bool = True) -> None:
    """Compile libspires_amd.so for gfx950 (hipcc cross-compiles without a GPU)."""
    args = ["make", "-s", "-j8", "-C", HERE]
    if force:
        args.insert(2, "-B")
    subprocess.run(args, check=True)


def lib():
    """Load the HIP library (raises if it was never built: there is no CPU path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: build it with spires_amd.build() "
                               "(the engine has no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def srand(seed: int) -> None:
    """glibc srand(): the reference's (and this library's) construction RNG."""
    _libc.srand(C.c_uint(seed))


def last_error() -> str:
    m = lib().spires_ext_last_error()
    return m.decode() if m else ""


def _check(status):
    if status != 0:
        raise SpiresError(status, last_error())


def _dp(a):
    return None if a is None else a.ctypes.data_as(_D)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def make_config(n, din=1, dout=1, rho=0.9, ei=0.8, gain=1.0, conn=0.1, dt=1.0,
                conn_type="random", neuron="flif_gl", params=(), synapse="simple",
                backend="sparse", syn_params=None):
    """(Config, owned arrays): the arrays must outlive any reservoir made from
    the config, because the library retains the parameter pointers (as the
    reference does)."""
    np_arr = (C.c_double * max(1, len(params)))(*params)
    sp_arr = None
    if syn_params is not None:
        sp_arr = (C.c_double * max(1, len(syn_params)))(*syn_params)
    cfg = Config(n, din, dout, rho, ei, gain, conn, dt, CONN.get(conn_type, conn_type),
                 NEURON.get(neuron, neuron), C.cast(np_arr, _D), SYNAPSE.get(synapse, synapse),
                 BACKEND.get(backend, backend), C.cast(sp_arr, _D) if sp_arr is not None else None)
    return cfg, np_arr, sp_arr


def spectral_radius(rp, ci, val):
    """Reference power-iteration estimate of rho(W) for a CSR matrix."""
    rp = np.ascontiguousarray(rp, np.int64)
    ci = np.ascontiguousarray(ci, np.int32)
    val = _f64(val)
    return lib().spires_ext_spectral_radius(len(rp) - 1, rp.ctypes.data_as(_I64), ci.ctypes.data_as(_I32),
                                            _dp(val))


def spectral_radius_device(rp, ci, val):
    """The reference power iteration on the GPU: (lambda, sparse products run);
    lambda is bit-identical to spectral_radius()."""
    rp = np.ascontiguousarray(rp, np.int64)
    ci = np.ascontiguousarray(ci, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    lam, it = C.c_double(), C.c_int()
    _check(lib().spires_ext_spectral_radius_device(len(rp) - 1, rp.ctypes.data_as(_I64), ci.ctypes.data_as(_I32),
                                                   val.ctypes.data_as(_D), C.byref(lam), C.byref(it)))
    return lam.value, it.value


def _take_csr(n, nnz, rp, ci, val, lam):
    k = nnz.value
    try:
        out = (np.ctypeslib.as_array(rp, shape=(n + 1,)).copy(),
               np.ctypeslib.as_array(ci, shape=(k,)).copy() if k else np.zeros(0, np.int32),
               np.ctypeslib.as_array(val, shape=(k,)).copy() if k else np.zeros(0, np.float64),
               lam.value)
    finally:
        for q in (rp, ci, val):
            _libc.free(C.cast(q, _P))
    return out


def random_er(n, p, seed, rho, ei=0.8):
    """Erdos-Renyi recurrent weights built on the GPU: (row_ptr, col_idx,
    values, lambda) with values scaled to spectral radius rho (lambda is the
    power-iteration estimate before scaling)."""
    nnz, lam = _SZ(), C.c_double()
    rp, ci, val = _I64(), _I32(), _D()
    _check(lib().spires_ext_random_er(n, float(p), float(ei), C.c_uint64(seed), float(rho), C.byref(nnz), C.byref(rp),
                                      C.byref(ci), C.byref(val), C.byref(lam)))
    return _take_csr(n, nnz, rp, ci, val, lam)


def build_topology(topology, n, conn, seed, rho, ei=0.8):
    """Recurrent weights of any topology ("random", "small_world",
    "scale_free") built on the GPU (spires_ext_build_topology): (row_ptr,
    col_idx, values, lambda), values scaled to spectral radius rho."""
    nnz, lam = _SZ(), C.c_double()
    rp, ci, val = _I64(), _I32(), _D()
    _check(lib().spires_ext_build_topology(CONN[topology], n, float(conn), float(ei), C.c_uint64(seed), float(rho),
                                           C.byref(nnz), C.byref(rp), C.byref(ci), C.byref(val), C.byref(lam)))
    return _take_csr(n, nnz, rp, ci, val, lam)


def build_weights(cfg_tuple):
    """(row_ptr, col_idx, values, W_in, W_out) that spires_reservoir_create
    would build for this config (consumes glibc rand() the same way)."""
    cfg = cfg_tuple[0]
    nnz = _SZ()
    rp, ci, val, wi, wo = _I64(), _I32(), _D(), _D(), _D()
    _check(lib().spires_ext_build_weights(C.byref(cfg), C.byref(nnz), C.byref(rp), C.byref(ci), C.byref(val),
                                          C.byref(wi), C.byref(wo)))
    n, k = cfg.num_neurons, nnz.value
    out = (np.ctypeslib.as_array(rp, shape=(n + 1,)).copy(),
           np.ctypeslib.as_array(ci, shape=(max(k, 1),))[:k].copy(),
           np.ctypeslib.as_array(val, shape=(max(k, 1),))[:k].copy(),
           np.ctypeslib.as_array(wi, shape=(max(n * cfg.num_inputs, 1),))[:n * cfg.num_inputs].copy()
           .reshape(n, cfg.num_inputs),
           np.ctypeslib.as_array(wo, shape=(max(n * cfg.num_outputs, 1),))[:n * cfg.num_outputs].copy()
           .reshape(cfg.num_outputs, n))
    for p in (rp, ci, val, wi, wo):
        _libc.free(C.cast(p, _P))
    return out


def exchange_floor(workgroups, steps=20000):
    """Microseconds per step of the persistent kernels' spike exchange alone
    among `workgroups` workgroups (spires_ext_exchange_floor)."""
    us = C.c_double()
    _check(lib().spires_ext_exchange_floor(int(workgroups), int(steps), C.byref(us)))
    return us.value


def run_resident_batch(reservoirs, t0, T, flags=None):
    """Advance several reservoirs over steps [t0, t0+T) of their uploaded
    inputs in one device launch (spires_ext_run_resident_batch)."""
    flags = REC_STATE if flags is None else flags
    arr = (_P * len(reservoirs))(*[r.h for r in reservoirs])
    _check(lib().spires_ext_run_resident_batch(arr, len(reservoirs), t0, T, flags))


def fit_ridge_records_batch(reservoirs, row0, row1, Ys, lams):
    """Ridge W_out of each reservoir on its recorded rows [row0, row1) against
    its own Y and lambda, factorised together (spires_ext_fit_ridge_records_batch).
    Returns [count x 2]: (0 Cholesky / 1 LU / -1 singular, relative residual)."""
    k = len(reservoirs)
    arr = (_P * k)(*[r.h for r in reservoirs])
    Ys = [_f64(Y).reshape(row1 - row0, -1) for Y in Ys]
    yp = (_D * k)(*[_dp(Y) for Y in Ys])
    lam = _f64(np.asarray(lams, dtype=np.float64).reshape(k))
    info = np.zeros((k, 2))
    _check(lib().spires_ext_fit_ridge_records_batch(arr, k, row0, row1, yp, _dp(lam), _dp(info)))
    return info


def batch_workgroups(reservoirs):
    """Workgroups a batched run of these reservoirs would use (0: cannot batch)."""
    arr = (_P * len(reservoirs))(*[r.h for r in reservoirs])
    return lib().spires_ext_batch_workgroups(arr, len(reservoirs))


class TimeBlock:
    """K-step time blocks of several exact-mode FLIF-GL reservoirs replayed from
    one HIP graph (spires_ext_timeblock_*).  Inputs are read and records written
    at row = macro step since reset; reserve the records before creating it."""

    def __init__(self, reservoirs, K, flags=None):
        self._res = list(reservoirs)
        flags = REC_STATE if flags is None else flags
        arr = (_P * len(self._res))(*[r.h for r in self._res])
        self.h = _P()
        _check(lib().spires_ext_timeblock_create(arr, len(self._res), K, flags, C.byref(self.h)))
        self.K = K
        _live.add(self)

    @property
    def launches(self):
        return lib().spires_ext_timeblock_launches(self.h)

    def run(self, blocks=1):
        _check(lib().spires_ext_timeblock_run(self.h, blocks))

    def close(self):
        if getattr(self, "h", None):
            lib().spires_ext_timeblock_destroy(self.h)
            self.h = None

    __del__ = close


class BlockRunner:
    """Advances exact-mode FLIF-GL reservoirs together in time blocks of K steps,
    each block one replay of a HIP graph (TimeBlock), in one of two tile shapes:

      "slices"   every reservoir in one cooperative launch of 64-neuron one-wave
                 slices (TB_ONE_LAUNCH, k_mux_gl);
      "products" workgroups of 16 or 32 neurons with product waves, packed per
                 neurons-per-workgroup class into launches of at most one
                 workgroup per CU (k_persist_tp_batch); a reservoir whose product
                 rows do not fit in LDS runs as a launch of its own per block.

    Inputs are read and records written at row = macro step since reset; reserve
    the records (`flags`: REC_STATE and/or REC_RASTER) before creating it.
    run(t, blocks) advances every reservoir from step t by blocks * K steps."""

    SHAPES = ("slices", "products")

    def __init__(self, reservoirs, K, shape="slices", flags=REC_STATE):
        if shape not in self.SHAPES:
            raise ValueError(f"tile shape {shape!r} not in {self.SHAPES}")
        self.res, self.K, self.shape, self.flags = list(reservoirs), int(K), shape, flags
        if shape == "slices":
            graph, self.singles = self.res, []
        else:
            plans = [r.kernel_plan() for r in self.res]
            graph = [r for r, p in zip(self.res, plans) if p["batch_nb"] > 0]
            self.singles = [r for r, p in zip(self.res, plans) if p["batch_nb"] == 0]
        self.tb = TimeBlock(graph, self.K, flags | (TB_ONE_LAUNCH if shape == "slices" else 0)) if graph else None

    @property
    def launches(self):
        return (self.tb.launches if self.tb else 0) + len(self.singles)

    def run(self, t, blocks=1):
        for b in range(blocks):
            if self.tb:
                self.tb.run(1)
            for r in self.singles:
                r.run_resident(t + b * self.K, self.K, self.flags | REC_APPEND)

    def close(self):
        if self.tb:
            self.tb.close()
            self.tb = None


def autotune_blocks(reservoirs, t, Ks=(64, 256, 1024), shapes=BlockRunner.SHAPES, flags=REC_STATE):
    """Picks the tile shape and block length K for BlockRunner by timing one
    block of each candidate from step t on (the blocks advance the reservoirs,
    so no step is run twice).  Returns (shape, K, {(shape, K): us per step}, t
    after the timed blocks)."""
    import time as _time
    table = {}
    for shape in shapes:
        for K in Ks:
            br = BlockRunner(reservoirs, K, shape, flags)
            try:
                synchronize(reservoirs)
                t0 = _time.perf_counter()
                br.run(t)
                synchronize(reservoirs)
                table[(shape, K)] = (_time.perf_counter() - t0) / K * 1e6
            finally:
                br.close()
            t += K
    shape, K = min(table, key=table.get)
    return shape, K, table, t


def synchronize(reservoirs):
    """Waits for the device work queued for these reservoirs."""
    for r in reservoirs:
        r.sync()


class Surrogate:
    """The AGILE surrogate MLP in_dim -> hidden -> hidden -> 1 on the GPU
    (spires_ext_mlp_*): forward, backward and Adam run as HIP kernels."""

    def __init__(self, in_dim=7, hidden=64, seed=1):
        import numpy as np
        self._np = np
        self.h = C.c_void_p()
        rc = lib().spires_ext_mlp_create(in_dim, hidden, seed, C.byref(self.h))
        if rc:
            raise RuntimeError("spires_ext_mlp_create: " + lib().spires_ext_mlp_last_error().decode())
        self.in_dim = in_dim
        self.count = lib().spires_ext_mlp_num_params(self.h)
        _live.add(self)

    def _f32(self, a):
        return self._np.ascontiguousarray(a, dtype=self._np.float32)

    def _ck(self, rc, what):
        if rc:
            raise RuntimeError(what + ": " + lib().spires_ext_mlp_last_error().decode())

    def params(self):
        p = self._np.empty(self.count, dtype=self._np.float32)
        self._ck(lib().spires_ext_mlp_get_params(self.h, p.ctypes.data_as(_F)), "mlp_get_params")
        return p

    def set_params(self, p):
        p = self._f32(p)
        assert p.size == self.count
        self._ck(lib().spires_ext_mlp_set_params(self.h, p.ctypes.data_as(_F)), "mlp_set_params")

    def predict(self, X):
        X = self._f32(X).reshape(-1, self.in_dim)
        y = self._np.empty(X.shape[0], dtype=self._np.float32)
        self._ck(lib().spires_ext_mlp_predict(self.h, X.ctypes.data_as(_F), X.shape[0], y.ctypes.data_as(_F)),
                 "mlp_predict")
        return y

    def gradient(self, X, T):
        X = self._f32(X).reshape(-1, self.in_dim)
        T = self._f32(T).reshape(-1)
        g = self._np.empty(self.count, dtype=self._np.float32)
        loss = C.c_float()
        self._ck(lib().spires_ext_mlp_gradient(self.h, X.ctypes.data_as(_F), T.ctypes.data_as(_F), X.shape[0],
                                               g.ctypes.data_as(_F), C.byref(loss)), "mlp_gradient")
        return g, loss.value

    def train(self, X, T, epochs=200, lr=1e-3):
        X = self._f32(X).reshape(-1, self.in_dim)
        T = self._f32(T).reshape(-1)
        loss = C.c_float()
        self._ck(lib().spires_ext_mlp_train(self.h, X.ctypes.data_as(_F), T.ctypes.data_as(_F), X.shape[0],
                                            epochs, C.c_float(lr), C.byref(loss)), "mlp_train")
        return loss.value

    def close(self):
        if self.h:
            lib().spires_ext_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Reservoir:
    """Object view of a spires_reservoir handle (spires.h names)."""

    def __init__(self, cfg_tuple, csr=None, W_in=None, W_out=None, device_seed=None):
        """csr=(row_ptr, col_idx, values) with W_in/W_out: given weights;
        device_seed: recurrent weights (any topology) built on the GPU from that seed
        (spires_ext_create_device); neither: spires_reservoir_create."""
        self._keep = cfg_tuple
        cfg = cfg_tuple[0]
        self.N, self.Din, self.Dout = cfg.num_neurons, cfg.num_inputs, cfg.num_outputs
        h = _P()
        L = lib()
        if device_seed is not None:
            _check(L.spires_ext_create_device(C.byref(cfg), C.c_uint64(device_seed), C.byref(h)))
        elif csr is None:
            _check(L.spires_reservoir_create(C.byref(cfg), C.byref(h)))
        else:
            rp, ci, val = csr
            self._rp = np.ascontiguousarray(rp, np.int64)
            self._ci = np.ascontiguousarray(ci, np.int32)
            self._val = _f64(val)
            wi = _f64(W_in)
            wo = _f64(W_out)
            _check(L.spires_ext_create_csr(C.byref(cfg), len(self._val), self._rp.ctypes.data_as(_I64),
                                           self._ci.ctypes.data_as(_I32), _dp(self._val), _dp(wi), _dp(wo),
                                           C.byref(h)))
        self.h = h
        _live.add(self)

    @classmethod
    def _wrap(cls, h, keep, N, Din, Dout):
        r = cls.__new__(cls)
        r.h, r._keep, r.N, r.Din, r.Dout = h, keep, N, Din, Dout
        _live.add(r)
        return r

    def close(self):
        if getattr(self, "h", None):
            lib().spires_reservoir_destroy(self.h)
            self.h = None

    __del__ = close

    def reset(self):
        _check(lib().spires_reservoir_reset(self.h))

    def step(self, u=None):
        u = _f64(u)
        _check(lib().spires_step(self.h, _dp(u)))

    def run(self, u):
        u = _f64(u)
        T = len(u) // max(1, self.Din) if u.ndim == 1 else u.shape[0]
        p = lib().spires_run(self.h, _dp(u), T)
        if not p:
            raise SpiresError(3, last_error())
        y = np.ctypeslib.as_array(p, shape=(T * self.Dout,)).copy().reshape(T, self.Dout)
        _libc.free(p)
        return y

    def run_record(self, u, T=None, states=True, raster=True, outputs=False):
        u = _f64(u)
        if T is None:
            T = u.shape[0] if u is not None else 0
        X = np.empty((T, self.N)) if states else None
        R = np.empty((T, self.N), np.uint8) if raster else None
        Y = np.empty((T, self.Dout)) if outputs else None
        _check(lib().spires_ext_run_record(self.h, _dp(u), T, _dp(X),
                                           None if R is None else R.ctypes.data_as(_U8), _dp(Y)))
        return X, R, Y

    def upload_inputs(self, u):
        u = _f64(u)
        _check(lib().spires_ext_upload_inputs(self.h, _dp(u), u.shape[0]))

    def reserve_records(self, rows, flags):
        _check(lib().spires_ext_reserve_records(self.h, rows, flags))

    def run_resident(self, t0, T, flags=0):
        _check(lib().spires_ext_run_resident(self.h, t0, T, flags))

    def fit_ridge_records(self, row0, row1, Y, lam):
        """Ridge W_out on recorded state rows [row0, row1) against Y; returns
        (flags, relative residual): flags bit 0 = LU fallback, bit 1 = fast
        mode fell back to the FP64 Gram."""
        Y = _f64(Y).reshape(row1 - row0, -1)
        info = np.zeros(2)
        _check(lib().spires_ext_fit_ridge_records(self.h, row0, row1, _dp(Y), lam, _dp(info)))
        return int(info[0]), float(info[1])

    def predict_records(self, row0, rows):
        """Readout outputs [rows x Dout] of recorded state rows under the current W_out."""
        y = np.empty((rows, self.Dout))
        _check(lib().spires_ext_predict_records(self.h, row0, rows, _dp(y)))
        return y

    def read_records(self, row0, rows, states=False, raster=False, outputs=False):
        X = np.empty((rows, self.N)) if states else None
        R = np.empty((rows, self.N), np.uint8) if raster else None
        Y = np.empty((rows, self.Dout)) if outputs else None
        _check(lib().spires_ext_read_records(self.h, row0, rows, _dp(X),
                                             None if R is None else R.ctypes.data_as(_U8), _dp(Y)))
        return X, R, Y

    def sync(self):
        _check(lib().spires_ext_sync(self.h))

    def stream(self):
        return lib().spires_ext_stream(self.h)

    def set_mode(self, mode, block_k=0):
        _check(lib().spires_ext_set_mode(self.h, mode, block_k))

    def kernel_plan(self):
        """Exact-mode launch plan (spires_ext_kernel_plan) as a dict."""
        info = (C.c_int * 7)()
        _check(lib().spires_ext_kernel_plan(self.h, info))
        return dict(zip(("kernel", "nb", "workgroups", "row_pull", "weights_lds", "lds_bytes", "batch_nb"),
                        list(info)))

    def set_kernel_path(self, path):
        """0 automatic, 1 one launch per micro-step, 2 persistent kernel only."""
        _check(lib().spires_ext_set_kernel_path(self.h, path))

    def train_ridge(self, u, y, lam):
        u, y = _f64(u), _f64(y)
        _check(lib().spires_train_ridge(self.h, _dp(u), _dp(y), u.shape[0], lam))

    def train_ridge_washout(self, u, y, washout, lam):
        u, y = _f64(u), _f64(y)
        info = np.zeros(2)
        _check(lib().spires_ext_train_ridge_washout(self.h, _dp(u), _dp(y), u.shape[0], washout, lam, _dp(info)))
        return info

    def train_rls(self, u, y, delta, lam):
        u, y = _f64(u), _f64(y)
        _check(lib().spires_train_rls(self.h, _dp(u), _dp(y), u.shape[0], delta, lam))

    def train_online(self, target, lr):
        _check(lib().spires_train_online(self.h, _dp(_f64(target)), lr))

    def coarse_grain(self, thr):
        h = _P()
        _check(lib().spires_coarse_grain(self.h, thr, C.byref(h)))
        return Reservoir._wrap(h, self._keep, lib().spires_num_neurons(h), self.Din, self.Dout)

    def state(self):
        v = np.empty(self.N)
        _check(lib().spires_read_reservoir_state(self.h, _dp(v)))
        return v

    def set_state(self, V):
        _check(lib().spires_ext_set_state(self.h, _dp(_f64(V))))

    def spikes(self):
        s = np.empty(self.N)
        _check(lib().spires_read_spike_state(self.h, _dp(s)))
        return s

    def output(self):
        y = np.empty(self.Dout)
        _check(lib().spires_compute_output(self.h, _dp(y)))
        return y

    def weights(self):
        W = np.empty((self.N, self.N))
        _check(lib().spires_read_weights(self.h, _dp(W)))
        return W

    def csr(self):
        nnz = lib().spires_ext_nnz(self.h)
        rp = np.empty(self.N + 1, np.int64)
        ci = np.empty(max(nnz, 1), np.int32)
        val = np.empty(max(nnz, 1))
        _check(lib().spires_ext_get_csr(self.h, rp.ctypes.data_as(_I64), ci.ctypes.data_as(_I32), _dp(val)))
        return rp, ci[:nnz], val[:nnz]

    def W_in(self):
        W = np.empty((self.N, self.Din))
        _check(lib().spires_ext_get_input_weights(self.h, _dp(W)))
        return W

    def W_out(self):
        W = np.empty((self.Dout, self.N))
        _check(lib().spires_read_readout(self.h, _dp(W)))
        return W
