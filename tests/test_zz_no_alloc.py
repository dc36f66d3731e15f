"""spires_ext_spectral_radius_device_noalloc (no device allocation; work
buffers in a fixed per-device array) against spires_ext_spectral_radius_device
on the same matrices: lambda bit for bit and the same iteration count, on two
calls in a row, the second with a smaller matrix (the array keeps the first
call's contents)."""
import ctypes as C

import numpy as np
import pytest

import spires_amd as sa

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        h = C.CDLL("libamdhip64.so")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipDeviceSynchronize.argtypes = []
        _hip = h
    return _hip


class DevBuf:
    """Device copy of a host array (the test allocates; the entry point does not)."""

    def __init__(self, a=None, nbytes=None):
        self.nbytes = a.nbytes if a is not None else nbytes
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), max(8, self.nbytes)) == 0
        if a is not None and self.nbytes:
            assert hip().hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), self.nbytes, 1) == 0   # H2D

    def read(self, dtype, count):
        out = np.empty(count, dtype)
        assert hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, out.nbytes, 2) == 0     # D2H
        return out

    def free(self):
        if self.p:
            hip().hipFree(self.p)
            self.p = C.c_void_p()


def _matrix(n, topo, conn, seed):
    sa.srand(seed)
    rp, ci, val, _, _ = sa.build_weights(sa.make_config(n, conn=conn, conn_type=topo, params=[1, 0, 0, 20, 0.8, 8, 0, 2]))
    return rp, ci, val


def _noalloc(rp, ci, val):
    n = len(rp) - 1
    bufs = [DevBuf(np.ascontiguousarray(rp, np.int64)), DevBuf(np.ascontiguousarray(ci, np.int32)),
            DevBuf(np.ascontiguousarray(val, np.float64)), DevBuf(nbytes=8), DevBuf(nbytes=4)]
    try:
        rc = sa.lib().spires_ext_spectral_radius_device_noalloc(n, *[b.p for b in bufs], None)
        assert rc == 0, sa.last_error()
        assert hip().hipDeviceSynchronize() == 0
        return float(bufs[3].read(np.float64, 1)[0]), int(bufs[4].read(np.int32, 1)[0])
    finally:
        for b in bufs:
            b.free()


def test_noalloc_power_iteration_matches_allocating_entry_point():
    cases = [(4096, "scale_free", 4.0 / 4095, 21), (1000, "random", 0.05, 22), (300, "small_world", 10.0 / 299, 23)]
    for n, topo, conn, seed in cases:      # each call smaller than the one before
        print(f"case: noalloc power iteration n={n} {topo}", flush=True)
        rp, ci, val = _matrix(n, topo, conn, seed)
        lam_ref, it_ref = sa.spectral_radius_device(rp, ci, val)
        lam, it = _noalloc(rp, ci, val)
        print(f"lambda {lam!r} ({it} products) vs {lam_ref!r} ({it_ref})", flush=True)
        assert np.float64(lam).view(np.uint64) == np.float64(lam_ref).view(np.uint64)
        assert it == it_ref


def test_noalloc_power_iteration_rejects_oversize():
    p = C.c_void_p(1)
    rc = sa.lib().spires_ext_spectral_radius_device_noalloc(19969, p, p, p, p, p, None)
    assert rc == 1 and "19968" in sa.last_error()
    assert sa.lib().spires_ext_spectral_radius_device_noalloc(16, None, p, p, p, p, None) == 1
