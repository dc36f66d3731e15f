// Persistent FP64 exact-mode step kernel.
//
// One cooperative launch runs many reservoir steps.  Workgroup g owns the
// contiguous neurons [g*NB, g*NB+NB) and keeps for the whole launch:
//   - their FLIF-GL voltage history ring in LDS ([L][NB], lane-contiguous), or
//     streams it from HBM when N*L*8 bytes exceed the chip's LDS (RL = false);
//   - the GL coefficients in LDS (broadcast reads);
//   - the CSR rows of its neurons in LDS, each column pre-decoded to
//     (exchange word, bit), when they fit.
// Per micro-step (two waves, one workgroup barrier):
//   wave 0  sums the GL history H = sum_{k=1..lim} c_k V[n-k] in ascending k
//           with one dependent FP64 add per lag, the reference's order
//           (flif_gl.c:66-72), from software-pipelined LDS loads;
//   wave 1  meanwhile collects the previous step's spikes of every workgroup
//           from step-tagged exchange words, pulls its neurons' CSR rows over
//           them (sparse.c:68-76; PSC traces are kept per edge and advanced in
//           the pull) and forms ext = gain * W_in u (reservoir.c:132-136);
//   barrier;
//   wave 0  applies the neuron update with total = ext + rec
//           (reservoir.c:158-160), writes its ring sample and the state
//           record, and publishes its spikes.
// Per-step scratch (rec, ext) is double-buffered by step parity, so one
// barrier per step orders the two waves.  The exchange word of a workgroup is
// [tag:32 | spike bits:32], written with a relaxed agent-scope atomic store;
// the tag (micro-step + 1) travels in the same 64-bit word as the data, so a
// reader that sees the tag sees the bits and no cache-maintenance fence is
// needed.  Every workgroup reads every word each step, so no workgroup gets
// more than one step ahead of another and two parity slots suffice.  Waits
// are bounded: on timeout the launch aborts and reports an error.
#include "dev_internal.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

namespace sdev {

constexpr int kPersistThreads = 128;
constexpr long long kSpinLimit = 1ll << 22;   // ~seconds of polling before the launch aborts

struct PersistArgs {
    int n, din, L, G, wpg, micro, stype;
    int ne_lds;                   // LDS capacity for CSR edges per workgroup (0: read CSR from HBM)
    long long st0;                // micro-step index since reset at launch start
    long long t0, T;              // macro steps [t0, t0+T) of the resident input series
    double gain, dt;
    double np[16];
    const double* W_in;
    const double* u;
    const int64_t* rp; const int32_t* col; const double* val;
    const double* edge_decay; double* edge_tr;
    double* V; const uint8_t* S; uint8_t* S_out; double* tprev; double* ring; const double* coef;
    unsigned long long* xch;      // [2][G*wpg]
    unsigned* abort_flag;
    double* rec_V; uint8_t* rec_S;
};

// h += sum_{q<cnt} c[kb+q] * ring[(sb-q)*STRIDE + lane]   (descending slots, ascending lags)
// Blocks of 16 lags in two register sets used alternately (no copies), so the
// loads of the next block are in flight while the current block's dependent
// adds run.
template <int STRIDE>
__device__ __forceinline__ void gl_load16(const double* __restrict__ ring, const double* __restrict__ c, int kb,
                                          int sb, int lane, double* v, double* w) {
    const double* rb = ring + (sb - 15) * STRIDE + lane;
#pragma unroll
    for (int j = 0; j < 16; ++j) { v[j] = rb[(15 - j) * STRIDE]; w[j] = c[kb + j]; }
}
__device__ __forceinline__ double gl_acc16(double h, const double* v, const double* w) {
#pragma unroll
    for (int j = 0; j < 16; ++j) h = h + w[j] * v[j];
    return h;
}
template <int STRIDE>
__device__ __forceinline__ double gl_seg_lds(double h, const double* __restrict__ ring, const double* __restrict__ c,
                                             int kb, int sb, int cnt, int lane) {
    const int nblk = cnt >> 4;
    int b = 0;
    if (nblk > 0) {
        double va[16], wa[16], vb[16], wb[16];
        gl_load16<STRIDE>(ring, c, kb, sb, lane, va, wa);
        while (b + 2 <= nblk) {
            gl_load16<STRIDE>(ring, c, kb + 16 * (b + 1), sb - 16 * (b + 1), lane, vb, wb);
            h = gl_acc16(h, va, wa);
            if (b + 2 < nblk) gl_load16<STRIDE>(ring, c, kb + 16 * (b + 2), sb - 16 * (b + 2), lane, va, wa);
            h = gl_acc16(h, vb, wb);
            b += 2;
        }
        if (b < nblk) { h = gl_acc16(h, va, wa); ++b; }
    }
    for (int q = 16 * nblk; q < cnt; ++q) h = h + c[kb + q] * ring[(sb - q) * STRIDE + lane];
    return h;
}

__device__ __forceinline__ double gl_seg_hbm(double h, const double* __restrict__ ring, const double* __restrict__ c,
                                             int kb, int sb, int cnt, size_t stride) {
    constexpr int B = 16;
    int q = 0;
    for (; q + B <= cnt; q += B) {
        double v[B];
        const double* rb = ring + (size_t)(sb - q) * stride;
#pragma unroll
        for (int j = 0; j < B; ++j) v[j] = rb[-(ptrdiff_t)(j * stride)];
#pragma unroll
        for (int j = 0; j < B; ++j) h = h + c[kb + q + j] * v[j];
    }
    for (; q < cnt; ++q) h = h + c[kb + q] * ring[(size_t)(sb - q) * stride];
    return h;
}

__device__ __forceinline__ unsigned long long ld_word(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int NT, int NB, bool RL>
__global__ void __launch_bounds__(kPersistThreads) k_persist(PersistArgs a) {
    extern __shared__ double lds[];
    constexpr bool GL = (NT == NT_FLIF_GL);
    const int g = blockIdx.x, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int i0 = g * NB;
    const int nloc = min(NB, a.n - i0);
    const int L = a.L;
    const int wpg = a.wpg;
    const int nw = a.G * wpg;
    // ---- LDS layout
    double* ring_s = lds;                                   // [L][NB]      (RL)
    double* coef_s = ring_s + (RL ? (size_t)L * NB : 0);    // [L]          (GL)
    double* rec_s = coef_s + (GL ? L : 0);                  // [2][NB]
    double* ext_s = rec_s + 2 * NB;                         // [2][NB]
    double* eval_s = ext_s + 2 * NB;                        // [ne_lds]
    unsigned* bits_s = (unsigned*)(eval_s + a.ne_lds);      // [nw]
    unsigned* eidx_s = bits_s + nw;                         // [ne_lds]
    __shared__ int abort_s;
    if (tid == 0) abort_s = 0;
    if (GL) {
        for (int k = tid; k < L; k += kPersistThreads) coef_s[k] = a.coef[k];
        if (RL)
            for (int e = tid; e < L * NB; e += kPersistThreads) {
                const int s = e / NB, j = e % NB;
                ring_s[e] = (j < nloc) ? a.ring[(size_t)s * a.n + i0 + j] : 0.0;
            }
    }
    const int64_t E0 = a.rp[i0];
    const int ne = (int)(a.rp[i0 + nloc] - E0);
    const bool csr_lds = a.ne_lds > 0;
    if (csr_lds)
        for (int e = tid; e < ne; e += kPersistThreads) {
            const int j = a.col[E0 + e];
            const int gg = j / NB, r = j - gg * NB;
            eidx_s[e] = ((unsigned)(gg * wpg + (r >> 5)) << 5) | (unsigned)(r & 31);
            eval_s[e] = a.val[E0 + e];
        }
    // ---- per-lane state: wave 0 lane = neuron, wave 1 lane = row puller
    const bool own = (wave == 0) && (lane < nloc);
    const int i = i0 + lane;
    double V = own ? a.V[i] : 0.0;
    double tp = (own && GL) ? a.tprev[i] : 0.0;
    bool spk = own ? (a.S[i] != 0) : false;
    int rs = 0, re = 0;
    constexpr int kMaxDinReg = 16;
    double win[kMaxDinReg];
    const bool win_reg = a.din <= kMaxDinReg;
#pragma unroll
    for (int j = 0; j < kMaxDinReg; ++j) win[j] = 0.0;
    if (wave == 1 && lane < nloc) {
        rs = (int)(a.rp[i] - E0);
        re = (int)(a.rp[i + 1] - E0);
        if (win_reg)
            for (int j = 0; j < a.din; ++j) win[j] = a.W_in[(size_t)i * a.din + j];
    }
    const double* p = a.np;
    __syncthreads();

    long long st = a.st0;
    for (long long t = 0; t < a.T; ++t) {
        const long long tt = a.t0 + t;
        const int tb = (int)(t & 1);
        for (int mi = 0; mi < a.micro; ++mi, ++st) {
            const int sb = (int)(st & 1);
            const int head = GL ? (int)(st % L) : 0;
            if (wave == 0) {
                // ---------------- GL history (overlaps the exchange on wave 1)
                double H = 0.0;
                if (GL && own && tp > p[6]) {
                    const int lim = st < L ? (int)st : L - 1;
                    const int n1 = lim < head ? lim : head;
                    if (RL) {
                        H = gl_seg_lds<NB>(H, ring_s, coef_s, 1, head - 1, n1, lane);
                        H = gl_seg_lds<NB>(H, ring_s, coef_s, n1 + 1, head - n1 - 1 + L, lim - n1, lane);
                    } else {
                        H = gl_seg_hbm(H, a.ring + i, coef_s, 1, head - 1, n1, (size_t)a.n);
                        H = gl_seg_hbm(H, a.ring + i, coef_s, n1 + 1, head - n1 - 1 + L, lim - n1, (size_t)a.n);
                    }
                }
                __syncthreads();   // rec / ext of this step ready
                if (abort_s) break;
                int ring_slot_write = 0;
                if (own) {
                    const double x = ext_s[tb * NB + lane] + rec_s[sb * NB + lane];
                    spk = false;
                    if constexpr (NT == NT_LIF_DISCRETE) {        // lif_discrete.c:22-34
                        V = (1.0 - p[2]) * V + x + p[3];
                        if (V > p[1]) { V = p[0]; spk = true; }
                    } else if constexpr (NT == NT_LIF_BIO) {      // lif_bio.c:16-29
                        const double dV = (-V + x + p[3]) / p[2];
                        V = V + dV * a.dt;
                        if (V >= p[1]) { V = p[0]; spk = true; }
                    } else {                                       // flif_gl.c:50-100
                        const int sp = head == 0 ? L - 1 : head - 1;
                        if (tp > p[6]) {
                            const double Vp = RL ? ring_s[sp * NB + lane] : a.ring[(size_t)sp * a.n + i];
                            const double rhs = (-(Vp - p[2]) / p[3]) + x + p[5];
                            V = p[4] * rhs - H;
                            if (V >= p[0]) { V = p[1]; spk = true; tp = 0.0; }
                        } else {
                            V = p[1];
                        }
                        ring_slot_write = 1;
                    }
                }
                // publish first (the other workgroups wait on it): word = [tag | bits of 32 neurons]
                const unsigned long long m = __ballot(own && spk);
                if (lane < wpg) {
                    const unsigned bits = (unsigned)(m >> (32 * lane));
                    const unsigned long long word = ((unsigned long long)(unsigned)(st + 1) << 32) | bits;
                    __hip_atomic_store(a.xch + (size_t)sb * nw + (size_t)g * wpg + lane, word, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                }
                if (own) {
                    if (GL && ring_slot_write) {
                        if (RL) ring_s[head * NB + lane] = V;
                        else a.ring[(size_t)head * a.n + i] = V;
                        tp = tp + a.dt;
                    }
                    if (mi == a.micro - 1) {
                        if (a.rec_V) a.rec_V[(size_t)t * a.n + i] = V;
                        if (a.rec_S) a.rec_S[(size_t)t * a.n + i] = spk ? 1 : 0;
                    }
                }
            } else {
                // prefetch this macro step's input (its latency overlaps the exchange wait)
                double uin[kMaxDinReg];
                if (mi == 0 && win_reg) {
#pragma unroll
                    for (int j = 0; j < kMaxDinReg; ++j) uin[j] = (j < a.din) ? a.u[tt * a.din + j] : 0.0;
                }
                // ---------------- spikes of micro-step st-1
                bool aborted = false;
                if (st == a.st0) {
                    for (int w = lane; w < nw; w += 64) {   // left in HBM by the previous launch
                        const int gg = w / wpg, part = w % wpg;
                        unsigned b = 0;
                        for (int j = 0; j < 32; ++j) {
                            const int nj = gg * NB + part * 32 + j;
                            if (part * 32 + j < NB && nj < a.n && a.S[nj]) b |= 1u << j;
                        }
                        bits_s[w] = b;
                    }
                } else {
                    const unsigned long long* slot = a.xch + (size_t)(sb ^ 1) * nw;
                    const unsigned tag = (unsigned)st;       // tag of micro-step st-1
                    constexpr int kMaxWordsPerLane = 8;      // nw <= 512
                    unsigned long long word[kMaxWordsPerLane];
                    bool ready[kMaxWordsPerLane];
#pragma unroll
                    for (int q = 0; q < kMaxWordsPerLane; ++q) { ready[q] = lane + 64 * q >= nw; word[q] = 0; }
                    long long spins = 0;
                    while (true) {
#pragma unroll
                        for (int q = 0; q < kMaxWordsPerLane; ++q)
                            if (!ready[q]) word[q] = ld_word(&slot[lane + 64 * q]);
                        bool all_here = true;
#pragma unroll
                        for (int q = 0; q < kMaxWordsPerLane; ++q) {
                            if (!ready[q]) ready[q] = (unsigned)(word[q] >> 32) == tag;
                            all_here = all_here && ready[q];
                        }
                        if (__all(all_here)) break;
                        const bool give_up = (++spins > kSpinLimit) ||
                            __hip_atomic_load(a.abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
                        if (__any(give_up)) {
                            aborted = true;
                            if (lane == 0) {
                                abort_s = 1;
                                __hip_atomic_store(a.abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            }
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                    }
#pragma unroll
                    for (int q = 0; q < kMaxWordsPerLane; ++q)
                        if (lane + 64 * q < nw) bits_s[lane + 64 * q] = (unsigned)word[q];
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                // ---------------- CSR row pull and external input
                if (!aborted && lane < nloc) {
                    double s = 0.0;
                    if (a.stype == ST_SIMPLE) {
                        if (csr_lds) {
                            // batches of 8 edges: all LDS loads issued, then the in-order sum
                            int e = rs;
                            for (; e + 8 <= re; e += 8) {
                                unsigned pk[8], bw[8];
                                double wv[8];
#pragma unroll
                                for (int q = 0; q < 8; ++q) { pk[q] = eidx_s[e + q]; wv[q] = eval_s[e + q]; }
#pragma unroll
                                for (int q = 0; q < 8; ++q) bw[q] = bits_s[pk[q] >> 5];
#pragma unroll
                                for (int q = 0; q < 8; ++q)
                                    if ((bw[q] >> (pk[q] & 31)) & 1u) s = s + wv[q];
                            }
                            for (; e < re; ++e) {
                                const unsigned pk = eidx_s[e];
                                if ((bits_s[pk >> 5] >> (pk & 31)) & 1u) s = s + eval_s[e];
                            }
                        } else {
                            for (int64_t k = E0 + rs; k < E0 + re; ++k) {
                                const int j = a.col[k];
                                const int gg = j / NB, r = j - gg * NB;
                                if ((bits_s[gg * wpg + (r >> 5)] >> (r & 31)) & 1u) s = s + a.val[k];
                            }
                        }
                    } else {   // PSC traces per edge (homogeneous: uniform decay)
                        for (int e = rs; e < re; ++e) {
                            const int64_t k = E0 + e;
                            unsigned word_i, bit;
                            double w;
                            if (csr_lds) {
                                const unsigned pk = eidx_s[e];
                                word_i = pk >> 5; bit = pk & 31; w = eval_s[e];
                            } else {
                                const int j = a.col[k];
                                const int gg = j / NB, r = j - gg * NB;
                                word_i = (unsigned)(gg * wpg + (r >> 5)); bit = (unsigned)(r & 31); w = a.val[k];
                            }
                            const double sj = (double)((bits_s[word_i] >> bit) & 1u);
                            const double d = a.edge_decay[k];
                            const double tr = a.edge_tr[k] * d + (1.0 - d) * sj;
                            a.edge_tr[k] = tr;
                            s = s + w * tr;
                        }
                    }
                    rec_s[sb * NB + lane] = s;
                    if (mi == 0) {
                        double su = 0.0;
                        if (win_reg) {
#pragma unroll
                            for (int j = 0; j < kMaxDinReg; ++j)
                                if (j < a.din) su = su + win[j] * uin[j];
                        } else {
                            for (int j = 0; j < a.din; ++j) su = su + a.W_in[(size_t)i * a.din + j] * a.u[tt * a.din + j];
                        }
                        ext_s[tb * NB + lane] = a.gain * su;
                    }
                }
                __syncthreads();   // rec / ext of this step ready
                if (abort_s) break;
            }
        }
        if (abort_s) break;
    }
    // write back the persistent state
    if (own) {
        a.V[i] = V;
        a.S_out[i] = spk ? 1 : 0;
        if (GL) a.tprev[i] = tp;
    }
    if (GL && RL) {
        __syncthreads();
        for (int e = tid; e < L * NB; e += kPersistThreads) {
            const int s = e / NB, j = e % NB;
            if (j < nloc) a.ring[(size_t)s * a.n + i0 + j] = ring_s[e];
        }
    }
}

namespace {

struct Plan {
    int nb = 0, G = 0, ring_lds = 0, ne_lds = 0;
    size_t lds = 0;
};

bool persist_plan(spires_dev* g, int ncu, Plan* pl) {
    if (g->backend != BK_SPARSE) return false;
    if (!(g->ntype == NT_FLIF_GL || g->ntype == NT_LIF_DISCRETE || g->ntype == NT_LIF_BIO)) return false;
    if (!(g->stype == ST_SIMPLE || g->stype == ST_PSC_HET)) return false;   // PSC-hom runs as per-edge PSC
    const int n = g->n;
    const bool gl = g->ntype == NT_FLIF_GL;
    const int L = gl ? g->L : 0;
    const size_t cap = 160 * 1024 - 2048;
    auto max_edges = [&](int nb) {
        int64_t m = 0;
        for (int i0 = 0; i0 < n; i0 += nb) m = std::max<int64_t>(m, g->rp_host[std::min(n, i0 + nb)] - g->rp_host[i0]);
        return m;
    };
    auto fill = [&](int nb, int rl) -> bool {
        const int G = (n + nb - 1) / nb;
        const int wpg = (nb + 31) / 32;
        if (G > ncu || G * wpg > 512) return false;
        size_t base = (rl ? (size_t)L * nb * 8 : 0) + (gl ? (size_t)L * 8 : 0) + 4 * (size_t)nb * 8 +
                      (size_t)G * wpg * 4 + 64;
        if (base > cap) return false;
        const int64_t me = max_edges(nb);
        pl->ne_lds = (base + (size_t)me * 12 <= cap) ? (int)std::max<int64_t>(me, 1) : 0;
        pl->nb = nb; pl->G = G; pl->ring_lds = rl;
        pl->lds = base + (size_t)pl->ne_lds * 12;
        return true;
    };
    if (gl) {
        for (int nb : {32, 16})
            if (fill(nb, 1)) return true;
        for (int nb : {32, 64})
            if (fill(nb, 0)) return true;
        return false;
    }
    for (int nb : {32, 64})
        if (fill(nb, 0)) return true;
    return false;
}

template <int NT>
const void* pick_kernel(int nb, int rl) {
    if (rl) {
        if (nb == 32) return (const void*)k_persist<NT, 32, true>;
        return (const void*)k_persist<NT, 16, true>;
    }
    if (nb == 32) return (const void*)k_persist<NT, 32, false>;
    return (const void*)k_persist<NT, 64, false>;
}

int num_cus() {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return ncu;
}

}  // namespace

int persist_supported(spires_dev* g) {
    Plan pl;
    return persist_plan(g, num_cus(), &pl) ? 1 : 0;
}

int launch_persist(spires_dev* g, const double* u_series, long long t0, long long T, double* rec_V,
                   uint8_t* rec_S) {
    Plan pl;
    if (!persist_plan(g, num_cus(), &pl)) { set_error("persistent path unsupported"); return SPIRES_DEV_EINVAL; }
    const int wpg = (pl.nb + 31) / 32;
    const size_t nw = (size_t)pl.G * wpg;
    if (g->xch_words < 2 * nw) {
        dfree(g->xch);
        int rc = dalloc(&g->xch, 2 * nw);
        if (rc) return rc;
        g->xch_words = 2 * nw;
        g->xch_plan_G = 0;
    }
    if (g->xch_plan_G != pl.G) {
        SDEV_CK(hipMemsetAsync(g->xch, 0, g->xch_words * 8, g->stream));
        g->xch_plan_G = pl.G;
    }
    if (!g->abort_flag) {
        int rc = dalloc(&g->abort_flag, 1);
        if (rc) return rc;
    }
    SDEV_CK(hipMemsetAsync(g->abort_flag, 0, 4, g->stream));
    PersistArgs a;
    memset(&a, 0, sizeof(a));
    a.n = g->n; a.din = g->din; a.L = g->L; a.G = pl.G; a.wpg = wpg; a.micro = g->micro;
    a.stype = g->stype; a.ne_lds = pl.ne_lds;
    a.st0 = g->step; a.t0 = t0; a.T = T;
    a.gain = g->d.gain; a.dt = g->d.dt;
    memcpy(a.np, g->d.np, sizeof(a.np));
    a.W_in = g->W_in; a.u = u_series;
    a.rp = g->rp; a.col = g->col; a.val = g->val;
    a.edge_decay = g->edge_decay; a.edge_tr = g->edge_tr;
    a.V = g->V; a.S = g->S[g->cur]; a.S_out = g->S[g->cur ^ 1]; a.tprev = g->tprev; a.ring = g->ring;
    a.coef = g->coef;
    a.xch = g->xch; a.abort_flag = g->abort_flag;
    a.rec_V = rec_V; a.rec_S = rec_S;
    void* args[] = {&a};
    const void* fn = nullptr;
    switch (g->ntype) {
    case NT_LIF_DISCRETE: fn = pick_kernel<NT_LIF_DISCRETE>(pl.nb, pl.ring_lds); break;
    case NT_LIF_BIO: fn = pick_kernel<NT_LIF_BIO>(pl.nb, pl.ring_lds); break;
    default: fn = pick_kernel<NT_FLIF_GL>(pl.nb, pl.ring_lds); break;
    }
    SDEV_CK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pl.lds));
    SDEV_CK(hipLaunchCooperativeKernel(fn, dim3(pl.G), dim3(kPersistThreads), args, (unsigned)pl.lds, g->stream));
    g->step += (long long)T * g->micro;
    g->msteps += (size_t)T;
    g->cur ^= 1;
    unsigned flag = 0;
    SDEV_CK(hipMemcpyAsync(&flag, g->abort_flag, 4, hipMemcpyDeviceToHost, g->stream));
    SDEV_CK(hipStreamSynchronize(g->stream));
    if (flag) { set_error("persistent step kernel timed out waiting for the spike exchange"); return SPIRES_DEV_EINTERNAL; }
    return 0;
}

}  // namespace sdev
