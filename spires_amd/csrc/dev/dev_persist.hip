// Persistent FP64 exact-mode step kernel.
//
// One cooperative launch runs many reservoir steps.  Workgroup g owns the
// contiguous neurons [g*nb, g*nb+nb) and keeps, for the whole launch,
//   - their FLIF-GL voltage history ring in LDS ([L][nb], lane-contiguous), or
//     streams it from HBM when N*L*8 bytes exceed the chip's LDS,
//   - the GL coefficients in LDS (read as a broadcast).
// Per micro-step:
//   wave 0  sums the GL history H = sum_{k=1..lim} c_k V[n-k] in ascending k
//           (one dependent FP64 add per lag, the reference's order,
//           flif_gl.c:66-72);
//   wave 1  meanwhile gathers the previous step's spikes of every workgroup
//           from step-tagged exchange words and pulls the CSR rows of the
//           workgroup's neurons over them (sparse.c:68-76; PSC traces are kept
//           per edge and advanced inside the pull, psc_*.c prepare);
//   then wave 0 applies the neuron update with total = ext + rec
//   (reservoir.c:158-160), writes the new ring sample and the state record,
//   and publishes its spikes.
// The exchange word of workgroup g is [tag:32 | spike bits:32] written with a
// relaxed agent-scope atomic store; the tag (micro-step + 1) travels in the
// same 64-bit word as the data, so a reader that sees the tag sees the bits
// and no cache-maintenance fence is needed.  Every workgroup reads every word
// each step, so no workgroup can be more than one step ahead of another and
// two parity slots suffice.  Waits are bounded; on timeout the launch aborts
// and reports an error instead of hanging.
#include "dev_internal.hpp"

#include <cstring>

namespace sdev {

constexpr int kPersistThreads = 128;      // wave 0: neuron updates, wave 1: exchange + row pull
constexpr long long kSpinLimit = 1ll << 22;   // ~seconds of polling before the launch aborts

struct PersistArgs {
    int n, din, L, nb, G, wpg, micro, ntype, stype;
    int ring_lds;                 // 1: ring in LDS, 0: ring in HBM ([L][n])
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

__device__ __forceinline__ unsigned long long ld_word(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_word(unsigned long long* p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// GL history from a lane-contiguous ring [L][nb] (LDS or HBM with stride).
template <typename RingPtr>
__device__ __forceinline__ double gl_sum(RingPtr ring, size_t stride, const double* __restrict__ c, int L,
                                         int head, int lim, int lane) {
    double h = 0.0;
    int k = 1;
    const int r1 = lim < head ? lim : head;
    constexpr int B = 16;
    for (; k + B - 1 <= r1; k += B) {
        double v[B], cc[B];
#pragma unroll
        for (int j = 0; j < B; ++j) { v[j] = ring[(size_t)(head - k - j) * stride + lane]; cc[j] = c[k + j]; }
#pragma unroll
        for (int j = 0; j < B; ++j) h = h + cc[j] * v[j];
    }
    for (; k <= r1; ++k) h = h + c[k] * ring[(size_t)(head - k) * stride + lane];
    for (; k + B - 1 <= lim; k += B) {
        double v[B], cc[B];
#pragma unroll
        for (int j = 0; j < B; ++j) { v[j] = ring[(size_t)(head - k - j + L) * stride + lane]; cc[j] = c[k + j]; }
#pragma unroll
        for (int j = 0; j < B; ++j) h = h + cc[j] * v[j];
    }
    for (; k <= lim; ++k) h = h + c[k] * ring[(size_t)(head - k + L) * stride + lane];
    return h;
}

template <int NT>
__global__ void __launch_bounds__(kPersistThreads) k_persist(PersistArgs a) {
    extern __shared__ double lds[];
    const int g = blockIdx.x;
    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    const int lane = tid & 63;
    const int nb = a.nb;
    const int i0 = g * nb;
    const int nloc = min(nb, a.n - i0);
    const int L = a.L;
    const int nw = a.G * a.wpg;
    // LDS layout
    double* ring_s = lds;                                               // [L][nb] if ring_lds
    double* coef_s = ring_s + (a.ring_lds ? (size_t)L * nb : 0);        // [L]
    double* rec_s = coef_s + (NT == NT_FLIF_GL ? L : 0);                // [nb]
    double* ext_s = rec_s + nb;                                         // [nb]
    double* H_s = ext_s + nb;                                           // [nb]
    unsigned* bits_s = (unsigned*)(H_s + nb);                           // [nw]
    __shared__ int abort_s;
    if (tid == 0) abort_s = 0;
    if (NT == NT_FLIF_GL) {
        for (int k = tid; k < L; k += kPersistThreads) coef_s[k] = a.coef[k];
        if (a.ring_lds)
            for (int e = tid; e < L * nb; e += kPersistThreads) {
                const int s = e / nb, j = e % nb;
                ring_s[e] = (j < nloc) ? a.ring[(size_t)s * a.n + i0 + j] : 0.0;
            }
    }
    // per-lane neuron state (wave 0 lanes own neurons)
    const bool own = (wave == 0) && (lane < nloc);
    const int i = i0 + lane;
    double V = own ? a.V[i] : 0.0;
    double tp = (own && (NT == NT_FLIF_GL)) ? a.tprev[i] : 0.0;
    bool spk = own ? (a.S[i] != 0) : false;
    const double* p = a.np;
    __syncthreads();

    long long st = a.st0;
    for (long long t = 0; t < a.T; ++t) {
        const long long tt = a.t0 + t;
        for (int mi = 0; mi < a.micro; ++mi, ++st) {
            const int head = (NT == NT_FLIF_GL) ? (int)(st % L) : 0;
            const int lim = (NT == NT_FLIF_GL) ? (st < L ? (int)st : L - 1) : 0;
            if (wave == 0) {
                if (NT == NT_FLIF_GL && own && tp > p[6]) {
                    if (a.ring_lds) H_s[lane] = gl_sum(ring_s, (size_t)nb, coef_s, L, head, lim, lane);
                    else H_s[lane] = gl_sum(a.ring + i0, (size_t)a.n, coef_s, L, head, lim, lane);
                }
            } else {
                // ---- gather spikes of micro-step st-1 ----
                if (st == a.st0) {
                    // first step of the launch: spikes left in HBM by the previous launch
                    for (int w = lane; w < nw; w += 64) {
                        const int gg = w / a.wpg, part = w % a.wpg;
                        unsigned b = 0;
                        for (int j = 0; j < 32; ++j) {
                            const int nj = gg * nb + part * 32 + j;
                            if (part * 32 + j < nb && nj < a.n && a.S[nj]) b |= 1u << j;
                        }
                        bits_s[w] = b;
                    }
                } else {
                    const unsigned long long* slot = a.xch + (size_t)((st - 1) & 1) * nw;
                    const unsigned tag = (unsigned)st;          // tag of micro-step st-1
                    for (int w0 = 0; w0 < nw; w0 += 64) {
                        const int w = w0 + lane;
                        bool ready = w >= nw;
                        unsigned long long word = 0;
                        long long spins = 0;
                        bool aborted = false;
                        while (true) {
                            if (!ready) {
                                word = ld_word(&slot[w]);
                                ready = (unsigned)(word >> 32) == tag;
                            }
                            if (__all(ready)) break;
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
                        if (w < nw) bits_s[w] = (unsigned)word;
                        if (aborted) break;
                    }
                }
                // external input once per macro step (reservoir.c:132-136)
                if (mi == 0 && lane < nloc) {
                    double su = 0.0;
                    for (int j = 0; j < a.din; ++j) su = su + a.W_in[(size_t)(i0 + lane) * a.din + j] * a.u[tt * a.din + j];
                    ext_s[lane] = a.gain * su;
                }
            }
            __syncthreads();   // bits gathered (wave 1), ext ready
            if (abort_s) break;
            if (wave == 1 && lane < nloc) {
                // ---- CSR row pull over the spike bits ----
                const int ii = i0 + lane;
                double s = 0.0;
                const int64_t k0 = a.rp[ii], k1 = a.rp[ii + 1];
                if (a.stype == ST_SIMPLE) {
                    for (int64_t k = k0; k < k1; ++k) {
                        const int j = a.col[k];
                        const int gg = j / nb, r = j - gg * nb;
                        if ((bits_s[gg * a.wpg + (r >> 5)] >> (r & 31)) & 1u) s = s + a.val[k];
                    }
                } else {  // PSC traces per edge (homogeneous: uniform decay)
                    for (int64_t k = k0; k < k1; ++k) {
                        const int j = a.col[k];
                        const int gg = j / nb, r = j - gg * nb;
                        const double sj = (double)((bits_s[gg * a.wpg + (r >> 5)] >> (r & 31)) & 1u);
                        const double d = a.edge_decay[k];
                        const double tr = a.edge_tr[k] * d + (1.0 - d) * sj;
                        a.edge_tr[k] = tr;
                        s = s + a.val[k] * tr;
                    }
                }
                rec_s[lane] = s;
            }
            __syncthreads();   // rec ready, H ready
            if (wave == 0) {
                if (own) {
                    const double x = ext_s[lane] + rec_s[lane];
                    spk = false;
                    if constexpr (NT == NT_LIF_DISCRETE) {
                        V = (1.0 - p[2]) * V + x + p[3];
                        if (V > p[1]) { V = p[0]; spk = true; }
                    } else if constexpr (NT == NT_LIF_BIO) {
                        const double dV = (-V + x + p[3]) / p[2];
                        V = V + dV * a.dt;
                        if (V >= p[1]) { V = p[0]; spk = true; }
                    } else {  // FLIF-GL, flif_gl.c:50-100
                        const int sp = head == 0 ? L - 1 : head - 1;
                        if (tp > p[6]) {
                            const double Vp = a.ring_lds ? ring_s[(size_t)sp * nb + lane] : a.ring[(size_t)sp * a.n + i];
                            const double rhs = (-(Vp - p[2]) / p[3]) + x + p[5];
                            V = p[4] * rhs - H_s[lane];
                            if (V >= p[0]) { V = p[1]; spk = true; tp = 0.0; }
                        } else {
                            V = p[1];
                        }
                        if (a.ring_lds) ring_s[(size_t)head * nb + lane] = V;
                        else a.ring[(size_t)head * a.n + i] = V;
                        tp = tp + a.dt;
                    }
                    if (mi == a.micro - 1) {
                        if (a.rec_V) a.rec_V[(size_t)t * a.n + i] = V;
                        if (a.rec_S) a.rec_S[(size_t)t * a.n + i] = spk ? 1 : 0;
                    }
                }
                // publish this step's spikes: word = [tag | bits of 32 neurons]
                const unsigned long long m = __ballot(own && spk);
                if (lane < a.wpg) {
                    const unsigned bits = (unsigned)(m >> (32 * lane));
                    const unsigned long long word = ((unsigned long long)(unsigned)(st + 1) << 32) | bits;
                    st_word(a.xch + (size_t)(st & 1) * nw + (size_t)g * a.wpg + lane, word);
                }
            }
            __syncthreads();   // ring sample written before the next history sum
        }
        if (abort_s) break;
    }
    // write back the persistent state
    if (own) {
        a.V[i] = V;
        a.S_out[i] = spk ? 1 : 0;
        if (NT == NT_FLIF_GL) a.tprev[i] = tp;
    }
    if (NT == NT_FLIF_GL && a.ring_lds)
        for (int e = tid; e < L * nb; e += kPersistThreads) {
            const int s = e / nb, j = e % nb;
            if (j < nloc) a.ring[(size_t)s * a.n + i0 + j] = ring_s[e];
        }
}

// Chooses the workgroup split; returns false if the persistent path cannot
// run this reservoir (then the per-step path is used).
static bool persist_plan(const spires_dev* g, int ncu, int* nb, int* G, int* ring_lds, size_t* lds_bytes) {
    if (g->backend != BK_SPARSE) return false;
    if (!(g->ntype == NT_FLIF_GL || g->ntype == NT_LIF_DISCRETE || g->ntype == NT_LIF_BIO)) return false;
    if (!(g->stype == ST_SIMPLE || g->stype == ST_PSC_HET)) return false;   // PSC-hom runs as per-edge PSC
    const int n = g->n;
    const int L = g->ntype == NT_FLIF_GL ? g->L : 0;
    const size_t lds_cap = 160 * 1024 - 4096;
    for (int cand : {32, 16}) {
        const int GG = (n + cand - 1) / cand;
        const size_t need = (size_t)L * cand * 8 + (size_t)L * 8 + 3 * (size_t)cand * 8 + (size_t)GG * 4 + 64;
        if (GG <= ncu && need <= lds_cap && L > 0) {
            *nb = cand; *G = GG; *ring_lds = 1; *lds_bytes = need;
            return true;
        }
    }
    // history streamed from HBM: up to 64 neurons per workgroup
    int cand = (n + ncu - 1) / ncu;
    if (cand < 32) cand = 32;
    if (cand > 64) return false;
    const int GG = (n + cand - 1) / cand;
    const int wpg = (cand + 31) / 32;
    *nb = cand; *G = GG; *ring_lds = 0;
    *lds_bytes = (size_t)L * 8 + 3 * (size_t)cand * 8 + (size_t)GG * wpg * 4 + 64;
    return true;
}

int persist_supported(spires_dev* g) {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    int nb, G, rl;
    size_t lb;
    return persist_plan(g, ncu, &nb, &G, &rl, &lb) ? 1 : 0;
}

int launch_persist(spires_dev* g, const double* u_series, long long t0, long long T, double* rec_V,
                   uint8_t* rec_S) {
    int dev = 0, ncu = 0;
    SDEV_CK(hipGetDevice(&dev));
    SDEV_CK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    int nb, G, ring_lds;
    size_t lds;
    if (!persist_plan(g, ncu, &nb, &G, &ring_lds, &lds)) { set_error("persistent path unsupported"); return SPIRES_DEV_EINVAL; }
    const int wpg = (nb + 31) / 32;
    const size_t nw = (size_t)G * wpg;
    if (g->xch_words < 2 * nw) {
        dfree(g->xch);
        int rc = dalloc(&g->xch, 2 * nw);
        if (rc) return rc;
        g->xch_words = 2 * nw;
        SDEV_CK(hipMemsetAsync(g->xch, 0, 2 * nw * 8, g->stream));
        g->xch_plan_G = G;
    }
    if (g->xch_plan_G != G) {
        SDEV_CK(hipMemsetAsync(g->xch, 0, g->xch_words * 8, g->stream));
        g->xch_plan_G = G;
    }
    if (!g->abort_flag) {
        int rc = dalloc(&g->abort_flag, 1);
        if (rc) return rc;
    }
    SDEV_CK(hipMemsetAsync(g->abort_flag, 0, 4, g->stream));
    PersistArgs a;
    memset(&a, 0, sizeof(a));
    a.n = g->n; a.din = g->din; a.L = g->L; a.nb = nb; a.G = G; a.wpg = wpg; a.micro = g->micro;
    a.ntype = g->ntype; a.stype = g->stype; a.ring_lds = ring_lds;
    a.st0 = g->step; a.t0 = t0; a.T = T;
    a.gain = g->d.gain; a.dt = g->d.dt;
    memcpy(a.np, g->d.np, sizeof(a.np));
    a.W_in = g->W_in; a.u = u_series;
    a.rp = g->rp; a.col = g->col; a.val = g->val;
    a.edge_decay = g->edge_decay; a.edge_tr = g->edge_tr;
    a.V = g->V; a.S = g->S[g->cur]; a.S_out = g->S[g->cur ^ 1]; a.tprev = g->tprev; a.ring = g->ring; a.coef = g->coef;
    a.xch = g->xch; a.abort_flag = g->abort_flag;
    a.rec_V = rec_V; a.rec_S = rec_S;
    void* args[] = {&a};
    const void* fn = nullptr;
    switch (g->ntype) {
    case NT_LIF_DISCRETE: fn = (const void*)k_persist<NT_LIF_DISCRETE>; break;
    case NT_LIF_BIO: fn = (const void*)k_persist<NT_LIF_BIO>; break;
    default: fn = (const void*)k_persist<NT_FLIF_GL>; break;
    }
    SDEV_CK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    SDEV_CK(hipLaunchCooperativeKernel(fn, dim3(G), dim3(kPersistThreads), args, (unsigned)lds, g->stream));
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
