// Spike exchange and CSR row pull shared by the persistent step kernels.
//
// Workgroup g publishes, after each micro-step st, wpg words
//   [tag = st + 1 : 32 | spike bits of 32 of its neurons : 32]
// into parity slot st & 1 with relaxed agent-scope atomic stores.  A reader
// polls the words of micro-step st - 1 until every tag matches; the bits come
// in the same 64-bit word as the tag, so no fence is needed.  Because each
// workgroup reads all words every step, no workgroup is ever more than one
// step ahead of another, so two parity slots suffice.
#pragma once
#include "dev_internal.hpp"

namespace sdev {

constexpr int kMaxWordsPerLane = 8;     // exchange words polled per lane (G * wpg <= 512)
constexpr long long kExchangeSpinLimit = 1ll << 22;   // ~seconds of polling before a launch aborts

__device__ __forceinline__ void publish_spikes(unsigned long long* xch, int nw, int g, int wpg, long long st,
                                               int lane, unsigned long long ballot_mask) {
    if (lane < wpg) {
        const unsigned bits = (unsigned)(ballot_mask >> (32 * lane));
        const unsigned long long word = ((unsigned long long)(unsigned)(st + 1) << 32) | bits;
        __hip_atomic_store(xch + (size_t)(st & 1) * nw + (size_t)g * wpg + lane, word, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Spike bits of micro-step st-1 into bits_s[nw] (one wave).  At the launch's
// first step they come from the u8 spike vector S left in HBM by the previous
// launch.  Returns false (after raising the abort flags) on timeout or when
// another workgroup aborted.
__device__ __forceinline__ bool gather_spikes(const unsigned long long* xch, int nw, int nb, int wpg, int n,
                                              long long st, long long st0, const uint8_t* S, unsigned* bits_s,
                                              unsigned* abort_flag, int* abort_s, int lane) {
    if (st == st0) {
        for (int w = lane; w < nw; w += 64) {
            const int gg = w / wpg, part = w % wpg;
            unsigned b = 0;
            for (int j = 0; j < 32; ++j) {
                const int nj = gg * nb + part * 32 + j;
                if (part * 32 + j < nb && nj < n && S[nj]) b |= 1u << j;
            }
            bits_s[w] = b;
        }
        return true;
    }
    const unsigned long long* slot = xch + (size_t)((st - 1) & 1) * nw;
    const unsigned tag = (unsigned)st;
    unsigned long long word[kMaxWordsPerLane];
    bool ready[kMaxWordsPerLane];
#pragma unroll
    for (int q = 0; q < kMaxWordsPerLane; ++q) { ready[q] = lane + 64 * q >= nw; word[q] = 0; }
    long long spins = 0;
    bool ok = true;
    while (true) {
#pragma unroll
        for (int q = 0; q < kMaxWordsPerLane; ++q)
            if (!ready[q]) word[q] = __hip_atomic_load(&slot[lane + 64 * q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        bool all_here = true;
#pragma unroll
        for (int q = 0; q < kMaxWordsPerLane; ++q) {
            if (!ready[q]) ready[q] = (unsigned)(word[q] >> 32) == tag;
            all_here = all_here && ready[q];
        }
        if (__all(all_here)) break;
        const bool give_up = (++spins > kExchangeSpinLimit) ||
                             __hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
        if (__any(give_up)) {
            ok = false;
            if (lane == 0) {
                *abort_s = 1;
                __hip_atomic_store(abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
#pragma unroll
    for (int q = 0; q < kMaxWordsPerLane; ++q)
        if (lane + 64 * q < nw) bits_s[lane + 64 * q] = (unsigned)word[q];
    return ok;
}

// Pre-decoded CSR column: (exchange word << 5) | bit.
__device__ __forceinline__ unsigned encode_col(int j, int nb, int wpg) {
    const int gg = j / nb, r = j - gg * nb;
    return ((unsigned)(gg * wpg + (r >> 5)) << 5) | (unsigned)(r & 31);
}

// rec = sum over the row's edges e in [rs, re) in CSR order (sparse.c:68-76).
// SIMPLE: spiking columns only (bit-identical to the full sum with 0/1
// inputs).  PSC (stype != SIMPLE): per-edge trace tr = tr*d + (1-d)*s_j,
// advanced in place, then w * tr.  Edge e of the row sits at global index
// E0 + e; eidx/eval are the LDS copies when csr_lds.
__device__ __forceinline__ double row_pull(int stype, bool csr_lds, int rs, int re, int64_t E0,
                                           const unsigned* __restrict__ eidx, const double* __restrict__ eval,
                                           const int32_t* __restrict__ col, const double* __restrict__ val,
                                           const unsigned* __restrict__ bits, int nb, int wpg,
                                           const double* __restrict__ edge_decay, double* __restrict__ edge_tr) {
    double s = 0.0;
    if (stype == ST_SIMPLE && csr_lds) {
        int e = rs;
        for (; e + 8 <= re; e += 8) {   // all loads of a batch issued before the in-order sum
            unsigned pk[8], bw[8];
            double wv[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) { pk[q] = eidx[e + q]; wv[q] = eval[e + q]; }
#pragma unroll
            for (int q = 0; q < 8; ++q) bw[q] = bits[pk[q] >> 5];
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if ((bw[q] >> (pk[q] & 31)) & 1u) s = s + wv[q];
        }
        for (; e < re; ++e) {
            const unsigned pk = eidx[e];
            if ((bits[pk >> 5] >> (pk & 31)) & 1u) s = s + eval[e];
        }
        return s;
    }
    for (int e = rs; e < re; ++e) {
        const int64_t k = E0 + e;
        const unsigned pk = csr_lds ? eidx[e] : encode_col(col[k], nb, wpg);
        const double w = csr_lds ? eval[e] : val[k];
        const unsigned sbit = (bits[pk >> 5] >> (pk & 31)) & 1u;
        if (stype == ST_SIMPLE) {
            if (sbit) s = s + w;
        } else {
            const double d = edge_decay[k];
            const double tr = edge_tr[k] * d + (1.0 - d) * (double)sbit;
            edge_tr[k] = tr;
            s = s + w * tr;
        }
    }
    return s;
}

}  // namespace sdev
