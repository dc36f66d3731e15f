// Spike exchange and CSR row pull shared by the persistent step kernels.
//
// Workgroup g publishes, after each micro-step st, wpg words
//   [tag = st + 1 : 32 | spike bits of 32 of its neurons : 32]
// into parity slot st & 1 with relaxed agent-scope atomic stores.  A reader
// polls the words of micro-step st - 1 until every tag matches; the bits come
// in the same 64-bit word as the tag, so no fence is needed.  Because each
// workgroup reads all words every step, no workgroup is ever more than one
// step ahead of another, so two parity slots suffice.
//
// Each word sits in a 64-byte line of its own (kXchStride words apart): every
// workgroup polls every word, and words packed into a few lines make those
// lines a hot spot for the far-memory loads (MI355X, 256 workgroups, exchange
// alone: 2.13 us per round with spread words, 2.2-2.5 us with packed words).
#pragma once
#include "dev_internal.hpp"

namespace sdev {

constexpr int kMaxWordsPerLane = 8;     // exchange words polled per lane (G * wpg <= 512)
constexpr long long kExchangeSpinLimit = 1ll << 22;   // ~seconds of polling before a launch aborts
constexpr int kXchStride = 8;           // u64 words from one exchange word to the next

__device__ __forceinline__ void publish_spikes(unsigned long long* xch, int nw, int g, int wpg, long long st,
                                               int lane, unsigned long long ballot_mask) {
    if (lane < wpg) {
        const unsigned bits = (unsigned)(ballot_mask >> (32 * lane));
        const unsigned long long word = ((unsigned long long)(unsigned)(st + 1) << 32) | bits;
        __hip_atomic_store(xch + ((size_t)(st & 1) * nw + (size_t)g * wpg + lane) * kXchStride, word,
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Spike bits of micro-step st-1 into bits_s[nw] (one wave).  At the launch's
// first step they come from the u8 spike vector S left in HBM by the previous
// launch; neuron r of workgroup gg is gg * nb + r, or gg + ileave * r when the
// workgroups interleave neurons (ileave = workgroup count).  Returns false
// (after raising the abort flags) on timeout or when another workgroup aborted.
// Each poll reloads all nq = ceil(nw / 64) words of a lane (no per-word
// branches); one poll is in flight at a time, so none is left outstanding for
// the barrier that follows.  The shared abort flag is read every 64 polls.
__device__ __forceinline__ bool gather_spikes(const unsigned long long* xch, int nw, int nb, int wpg, int n,
                                              long long st, long long st0, const uint8_t* S, unsigned* bits_s,
                                              unsigned* abort_flag, int* abort_s, int lane, int ileave = 0) {
    if (st == st0) {
        for (int w = lane; w < nw; w += 64) {
            const int gg = w / wpg, part = w % wpg;
            unsigned b = 0;
            for (int j = 0; j < 32; ++j) {
                const int r = part * 32 + j;
                const int nj = ileave ? gg + ileave * r : gg * nb + r;
                if (r < nb && nj < n && S[nj]) b |= 1u << j;
            }
            bits_s[w] = b;
        }
        return true;
    }
    const unsigned long long* slot = xch + (size_t)((st - 1) & 1) * nw * kXchStride;
    const unsigned tag = (unsigned)st;
    const int nq = (nw + 63) >> 6;                     // uniform
    unsigned word[kMaxWordsPerLane];
    unsigned rdy = 0;                                  // bit q: word q of this lane seen (or absent)
#pragma unroll
    for (int q = 0; q < kMaxWordsPerLane; ++q) {
        word[q] = 0;
        if (q >= nq || lane + 64 * q >= nw) rdy |= 1u << q;
    }
    const unsigned full = (1u << kMaxWordsPerLane) - 1u;
    long long spins = 0;
    bool ok = true;
    while (true) {
        unsigned long long w[kMaxWordsPerLane];
#pragma unroll
        for (int q = 0; q < kMaxWordsPerLane; ++q)
            if (q < nq)
                w[q] = __hip_atomic_load(&slot[(size_t)min(lane + 64 * q, nw - 1) * kXchStride], __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
        for (int q = 0; q < kMaxWordsPerLane; ++q)
            if (q < nq && (unsigned)(w[q] >> 32) == tag) {
                word[q] = (unsigned)w[q];
                rdy |= 1u << q;
            }
        if (__all(rdy == full)) break;
        if ((++spins & 63) == 0) {
            const bool give_up = spins > kExchangeSpinLimit ||
                                 __hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
            if (__any(give_up)) {
                ok = false;
                if (lane == 0) {
                    *abort_s = 1;
                    __hip_atomic_store(abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                break;
            }
        }
        __builtin_amdgcn_s_sleep(1);
    }
#pragma unroll
    for (int q = 0; q < kMaxWordsPerLane; ++q)
        if (q < nq && lane + 64 * q < nw) bits_s[lane + 64 * q] = word[q];
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
    if (stype == ST_SIMPLE) {   // CSR slice in HBM: batches of 16 edges, loads issued before the in-order sum
        const int32_t* cr = col + E0;
        const double* vr = val + E0;
        int e = rs;
        for (; e + 16 <= re; e += 16) {
            int32_t cj[16];
            double wv[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) { cj[q] = cr[e + q]; wv[q] = vr[e + q]; }
            unsigned hit = 0;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const unsigned pk = encode_col(cj[q], nb, wpg);
                hit |= ((bits[pk >> 5] >> (pk & 31)) & 1u) << q;
            }
#pragma unroll
            for (int q = 0; q < 16; ++q)
                if ((hit >> q) & 1u) s = s + wv[q];
        }
        for (; e < re; ++e) {
            const unsigned pk = encode_col(cr[e], nb, wpg);
            if ((bits[pk >> 5] >> (pk & 31)) & 1u) s = s + vr[e];
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

// Row pull over column bitmaps (SIMPLE synapses, dense rows).  Row r of a
// workgroup keeps bm[r][w], the set of its columns among neurons 32w..32w+31,
// and the spikes come as words of the same layout; a row's CSR edges are its
// bitmap's set bits in ascending (word, bit) order, which is ascending column
// order.  The pull runs in two passes of one wave:
//   scan  all 64 lanes, lane = part * nb + r: the part's words [part*cw,
//         part*cw + cw) of row r are ANDed with the spike words and the
//         block-relative indices of the active edges are listed in place
//         (slot <= edge index, so parts never overlap);
//   sum   lane r adds the listed weights of its parts in order, the loads of
//         each batch issued before its adds.
// The sum therefore visits the spiking columns in ascending order, as the
// edge-list pull does.  pb[lane] is the block-relative index of the first
// edge of the lane's part (fixed for a launch).
__device__ __forceinline__ void bitmap_scan(const unsigned* __restrict__ bm_s, const unsigned* __restrict__ bits_s,
                                            unsigned short* __restrict__ list_s, int* __restrict__ cnt_s,
                                            const int* __restrict__ pb_s, int nw, int cw, int nb, bool live,
                                            int lane) {
    const int r = lane & (nb - 1), part = lane / nb;
    const int w_lo = part * cw, w_hi = min(nw, w_lo + cw);
    const unsigned* row = bm_s + r * nw;
    const int first = pb_s[lane];
    int pos = first, c = first;
    if (live) {
        for (int w0 = w_lo; w0 < w_hi; w0 += 8) {
            unsigned m[8], sw[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const bool ok = w0 + q < w_hi;
                m[q] = ok ? row[w0 + q] : 0u;
                sw[q] = ok ? bits_s[w0 + q] : 0u;
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                unsigned act = m[q] & sw[q];
                while (act) {
                    const unsigned b = (unsigned)__builtin_ctz(act);
                    list_s[c++] = (unsigned short)(pos + __builtin_popcount(m[q] & ((1u << b) - 1u)));
                    act &= act - 1u;
                }
                pos += __builtin_popcount(m[q]);
            }
        }
    }
    cnt_s[lane] = c - first;
}

// B weights loaded per batch; the list holds >= B valid indices past any
// part's count (stale entries or the initial zeros), so every load is in range.
template <int B>
__device__ __forceinline__ double bitmap_sum(const unsigned short* __restrict__ list_s, const int* __restrict__ cnt_s,
                                             const int* __restrict__ pb_s, const double* __restrict__ w, int nb,
                                             int lane) {
    double s = 0.0;
    for (int part = 0; part < 64 / nb; ++part) {
        const int c = cnt_s[part * nb + lane], b0 = pb_s[part * nb + lane];
        for (int j = 0; j < c; j += B) {
            unsigned idx[B];
            double wv[B];
#pragma unroll
            for (int q = 0; q < B; ++q) idx[q] = list_s[b0 + j + q];
#pragma unroll
            for (int q = 0; q < B; ++q) wv[q] = w[idx[q]];
#pragma unroll
            for (int q = 0; q < B; ++q)
                if (j + q < c) s = s + wv[q];
        }
    }
    return s;
}

}  // namespace sdev
