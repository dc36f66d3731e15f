"""numpy restatement of the device reservoir construction (test helper).

The device Erdos-Renyi generator (spires_ext_random_er) hashes every ordered
pair (i, j) with SplitMix64: u1 = unit0(key(i, j, 0)) decides the edge (u1 < p,
i != j), mag = unit1(key(i, j, 1)) in (0, 1] and u3 = unit0(key(i, j, 2))
< ei the sign, following generate_weight's distribution (reference
reservoir.c:345-545).  The matrix is then rescaled by rho / lambda with lambda
from the reference power iteration (here the oracle's).  spires_ext_create_device
draws W_in then W_out from one sequential SplitMix64 stream of the seed, two
draws per weight (1 - unit0 for the magnitude, unit0 < ei for the sign).
Every step is exact integer or IEEE arithmetic, so the device output must
match this bit for bit."""
import numpy as np

from oracle import pyoracle as po

_U = np.uint64
GOLDEN = _U(0x9E3779B97F4A7C15)
INV53 = 1.0 / 9007199254740992.0


def mix64(z):
    z = np.asarray(z, np.uint64) + GOLDEN
    z = (z ^ (z >> _U(30))) * _U(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> _U(27))) * _U(0x94D049BB133111EB)
    return z ^ (z >> _U(31))


def pair_key(seed, n, i, j, stream):
    idx = ((np.asarray(i, np.uint64) * _U(n) + np.asarray(j, np.uint64)) << _U(2)) | _U(stream)
    return mix64(_U(seed) ^ mix64(idx))


def unit0(h):
    return (h >> _U(11)).astype(np.float64) * INV53


def unit1(h):
    return ((h >> _U(11)).astype(np.float64) + 1.0) * INV53


def er_unscaled(n, p, ei, seed, block=256):
    """CSR (row_ptr, col_idx, values) before the spectral rescale."""
    rows, cols = [], []
    j = np.arange(n, dtype=np.int64)[None, :]
    with np.errstate(over="ignore"):
        for i0 in range(0, n, block):
            i = np.arange(i0, min(n, i0 + block), dtype=np.int64)[:, None]
            e = (unit0(pair_key(seed, n, i, j, 0)) < p) & (i != j)
            ii, jj = np.nonzero(e)
            rows.append(ii + i0)
            cols.append(jj)
        ri = np.concatenate(rows) if rows else np.zeros(0, np.int64)
        cj = np.concatenate(cols) if cols else np.zeros(0, np.int64)
        mag = unit1(pair_key(seed, n, ri, cj, 1))
        u3 = unit0(pair_key(seed, n, ri, cj, 2))
    val = np.where(u3 < ei, mag, -mag)
    rp = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(ri, minlength=n), out=rp[1:])
    return rp, cj.astype(np.int32), val


def random_er(n, p, ei, seed, rho):
    """(row_ptr, col_idx, values, lambda) as spires_ext_random_er returns them."""
    rp, ci, val = er_unscaled(n, p, ei, seed)
    lam = po.csr_spectral_radius(rp, ci, val)
    if lam > 1e-9:
        val = val * (rho / lam)
    return rp, ci, val, lam


def io_weights(n, din, dout, ei, seed):
    """(W_in [n x din], W_out flat [n * dout]) of spires_ext_create_device."""
    k = 2 * n * (din + dout)
    with np.errstate(over="ignore"):
        z = (_U(seed) ^ _U(0xD1B54A32D192ED03)) + GOLDEN * np.arange(k, dtype=np.uint64)
        u = unit0(mix64(z))
    mag, sgn = 1.0 - u[0::2], u[1::2]
    m = n * din
    w_in = np.where(sgn[:m] < ei, mag[:m], -mag[:m]).reshape(n, din)
    w_out = np.where(sgn[m:] < 0.5, mag[m:], -mag[m:])
    return w_in, w_out
