// FP32 fast mode for FLIF-GL: time-blocked history contraction on MFMA.
//
// For a block of K micro-steps starting at n0, the GL history of step n0+j
// (flif_gl.c:66-72) splits into
//   H[n0+j] = sum_{m=1}^{min(n0, L-1-j)} c_{m+j} V[n0-m]      (samples before the block)
//           + sum_{k=1}^{j} c_k V[n0+j-k]                      (in-block triangle)
// The first part, for all j at once, is the [nb x L] . [L x K] product of the
// workgroup's FP32 voltage history (ring [L][N] in HBM, slot = time mod L)
// with a Toeplitz coefficient matrix B[s][j] = c_{m(s)+j}, m(s) = (n0-s) mod L,
// generated on the fly from the coefficients in LDS.  Four waves compute it
// with v_mfma_f32_16x16x4f32 at each block start; then the K steps run like
// the exact persistent kernel (spike exchange + CSR pull on wave 1) with only
// the j-term triangle summed step-wise from LDS.  The membrane update,
// threshold comparison and reset stay FP64 (V = pow(dt,alpha) rhs - H).
#include "dev_internal.hpp"
#include "dev_exchange.hpp"

#include <algorithm>
#include <cstring>

namespace sdev {

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kFastThreads = 256;     // 4 waves: GEMM on all, steps on waves 0-1
constexpr int kFastNb = 64;           // neurons per workgroup (one wave of lanes)

struct FastArgs {
    int n, din, L, nb, G, wpg, micro, K, stype;
    int ne_lds;                   // LDS capacity for CSR edges per workgroup (0: CSR from HBM)
    long long st0, t0, T;
    double gain, dt;
    double np[16];
    const double* W_in;
    const double* u;
    const int64_t* rp; const int32_t* col; const double* val;
    const double* edge_decay; double* edge_tr;
    double* V; const uint8_t* S; uint8_t* S_out; double* tprev;
    float* ring;                  // [L][n] FP32
    const double* coef;           // FP64 coefficients (converted to FP32 in LDS)
    unsigned long long* xch;
    unsigned* abort_flag;
    double* rec_V; uint8_t* rec_S;
};

__global__ void __launch_bounds__(kFastThreads) k_fast(FastArgs a) {
    extern __shared__ __align__(16) float flds[];   // 16-byte aligned: b128 LDS reads
    // wave index and abort flag are read through readfirstlane so the compiler
    // treats them as wave-uniform (scalar branches, scalar loop counters).
    const int g = blockIdx.x, tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nb = a.nb, i0 = g * nb, nloc = min(nb, a.n - i0), L = a.L, K = a.K;
    const int wpg = a.wpg, nw = a.G * wpg;
    // LDS: coefficients (FP32, padded to even length), pre-block contraction O [K][nb],
    // in-block voltages Vb [K][nb], then 8-byte data: rec [2][nb], ext [2][nb],
    // CSR values [ne_lds], then 4-byte data: spike bits [nw], CSR columns [ne_lds]
    float* c_s = flds;
    float* O_s = c_s + (L + (L & 1));
    float* Vb_s = O_s + K * nb;
    double* rec_s = (double*)(Vb_s + K * nb);
    double* ext_s = rec_s + 2 * nb;
    double* eval_s = ext_s + 2 * nb;
    unsigned* bits_s = (unsigned*)(eval_s + a.ne_lds);
    unsigned* eidx_s = bits_s + nw;
    __shared__ int abort_s;
    if (tid == 0) abort_s = 0;
    for (int k = tid; k < L; k += kFastThreads) c_s[k] = (float)a.coef[k];
    const int64_t E0 = a.rp[i0];
    const int ne = (int)(a.rp[i0 + nloc] - E0);
    const bool csr_lds = a.ne_lds > 0;
    if (csr_lds)
        for (int e = tid; e < ne; e += kFastThreads) {
            eidx_s[e] = encode_col(a.col[E0 + e], nb, wpg);
            eval_s[e] = a.val[E0 + e];
        }
    const bool own = wave == 0 && lane < nloc;
    const int i = i0 + lane;
    double V = own ? a.V[i] : 0.0;
    double tp = own ? a.tprev[i] : 0.0;
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
    long long t = 0;
    int mi = 0;
    while (t < a.T) {
        // ---------------- block start: O = Hist . B on MFMA (all 4 waves) ----------------
        const long long n0 = st;
        {
            // wave w owns rows [16w, 16w+16) of the nb = 64 neurons, all K columns (K/16 tiles)
            const int row = 16 * wave + (lane & 15);
            const int kq = lane >> 4;                       // 0..3: k within the 4-deep MFMA step
            floatx4 acc[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[q] = floatx4{0.f, 0.f, 0.f, 0.f};
            const int ntile = K / 16;
            const int gi = i0 + row;
            for (int s0 = 0; s0 < L; s0 += 4) {
                const int s = s0 + kq;
                float av = 0.f;
                long long mm = 0;
                if (s < L) {
                    av = (gi < a.n) ? a.ring[(size_t)s * a.n + gi] : 0.f;
                    mm = (n0 - s) % L;
                    if (mm < 0) mm += L;
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (q < ntile) {
                        const int j = 16 * q + (lane & 15);
                        const long long idx = mm + j;
                        float bv = 0.f;
                        if (s < L && mm >= 1 && mm <= n0 && idx <= L - 1) bv = c_s[idx];
                        acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc[q], 0, 0, 0);
                    }
                }
            }
            // D layout: lane l, reg r -> row 4*(l/16)+r of the tile, column l%16
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (q < ntile) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int rr = 16 * wave + 4 * (lane >> 4) + r;
                        const int cc = 16 * q + (lane & 15);
                        O_s[cc * nb + rr] = acc[q][r];
                    }
                }
            }
        }
        __syncthreads();
        // ---------------- K steps (one workgroup barrier each) ----------------
        for (int j = 0; j < K && t < a.T; ++j) {
            const long long tt = a.t0 + t;
            const int head = (int)(st % L);
            const int sb = (int)(st & 1), tb = (int)(t & 1);
            if (wave == 0) {
                double H = 0.0;
                if (own && tp > p[6]) {
                    float h = O_s[j * nb + lane];
                    for (int k = 1; k <= j; ++k) h = h + c_s[k] * Vb_s[(j - k) * nb + lane];
                    H = (double)h;
                }
                __syncthreads();
                if (__builtin_amdgcn_readfirstlane(abort_s)) break;
                if (own) {
                    const double x = ext_s[tb * nb + lane] + rec_s[sb * nb + lane];
                    spk = false;
                    if (tp > p[6]) {
                        const double rhs = (-(V - p[2]) / p[3]) + x + p[5];
                        V = p[4] * rhs - H;
                        if (V >= p[0]) { V = p[1]; spk = true; tp = 0.0; }
                    } else {
                        V = p[1];
                    }
                }
                publish_spikes(a.xch, nw, g, wpg, st, lane, __ballot(own && spk));
                if (own) {
                    tp = tp + a.dt;
                    const float vf = (float)V;
                    Vb_s[j * nb + lane] = vf;
                    a.ring[(size_t)head * a.n + i] = vf;
                    if (mi == a.micro - 1) {
                        if (a.rec_V) a.rec_V[(size_t)t * a.n + i] = V;
                        if (a.rec_S) a.rec_S[(size_t)t * a.n + i] = spk ? 1 : 0;
                    }
                }
            } else if (wave == 1) {
                double uin[kMaxDinReg];
                if (mi == 0 && win_reg) {
#pragma unroll
                    for (int q = 0; q < kMaxDinReg; ++q) uin[q] = (q < a.din) ? a.u[tt * a.din + q] : 0.0;
                }
                const bool aborted = !gather_spikes(a.xch, nw, nb, wpg, a.n, st, a.st0, a.S, bits_s, a.abort_flag,
                                                    &abort_s, lane);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                if (!aborted && lane < nloc) {
                    rec_s[sb * nb + lane] = row_pull(a.stype, csr_lds, rs, re, E0, eidx_s, eval_s, a.col, a.val,
                                                     bits_s, nb, wpg, a.edge_decay, a.edge_tr);
                    if (mi == 0) {
                        double su = 0.0;
                        if (win_reg) {
#pragma unroll
                            for (int q = 0; q < kMaxDinReg; ++q)
                                if (q < a.din) su = su + win[q] * uin[q];
                        } else {
                            for (int q = 0; q < a.din; ++q) su = su + a.W_in[(size_t)i * a.din + q] * a.u[tt * a.din + q];
                        }
                        ext_s[tb * nb + lane] = a.gain * su;
                    }
                }
                __syncthreads();
                if (__builtin_amdgcn_readfirstlane(abort_s)) break;
            } else {
                __syncthreads();
                if (__builtin_amdgcn_readfirstlane(abort_s)) break;
            }
            ++st;
            if (++mi == a.micro) { mi = 0; ++t; }
        }
        if (__builtin_amdgcn_readfirstlane(abort_s)) break;
        __syncthreads();   // this block's voltages (Vb, ring) precede the next contraction
    }
    if (own) {
        a.V[i] = V;
        a.S_out[i] = spk ? 1 : 0;
        a.tprev[i] = tp;
    }
}

__global__ void k_ring_f64_to_f32(size_t cnt, const double* src, float* dst) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < cnt) dst[q] = (float)src[q];
}
__global__ void k_ring_f32_to_f64(size_t cnt, const float* src, double* dst) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < cnt) dst[q] = (double)src[q];
}

int fast_supported(spires_dev* g) {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    if (g->ntype != NT_FLIF_GL || g->backend != BK_SPARSE) return 0;
    if (!(g->stype == ST_SIMPLE || g->stype == ST_PSC_HET)) return 0;
    if (g->L < 16) return 0;
    return (g->n + kFastNb - 1) / kFastNb <= ncu ? 1 : 0;
}

// Converts the voltage history between the exact (FP64) and fast (FP32) rings
// when the mode changes.
int fast_sync_rings(spires_dev* g, bool to_fast) {
    const size_t cnt = (size_t)g->L * g->n;
    if (cnt == 0) return 0;
    if (to_fast) {
        if (g->ring32_cap < cnt) {
            dfree(g->ring32);
            int rc = dalloc(&g->ring32, cnt);
            if (rc) return rc;
            g->ring32_cap = cnt;
        }
        hipLaunchKernelGGL(k_ring_f64_to_f32, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, g->stream, cnt,
                           g->ring, g->ring32);
    } else if (g->ring32) {
        hipLaunchKernelGGL(k_ring_f32_to_f64, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, g->stream, cnt,
                           g->ring32, g->ring);
    }
    SDEV_CK(hipGetLastError());
    return 0;
}

int launch_fast(spires_dev* g, const double* u_series, long long t0, long long T, double* rec_V,
                uint8_t* rec_S) {
    const int nb = kFastNb, G = (g->n + nb - 1) / nb, wpg = 2;
    const size_t nw = (size_t)G * wpg;
    int K = g->block_k > 0 ? g->block_k : 64;
    if (K > 64) K = 64;
    K = (K / 16) * 16;
    if (K < 16) K = 16;
    while (K > g->L - 1 && K > 16) K -= 16;
    if (g->xch_words < 2 * nw || g->xch_plan_G != -G) {
        if (g->xch_words < 2 * nw) {
            dfree(g->xch);
            int rc = dalloc(&g->xch, 2 * nw);
            if (rc) return rc;
            g->xch_words = 2 * nw;
        }
        SDEV_CK(hipMemsetAsync(g->xch, 0, g->xch_words * 8, g->stream));
        g->xch_plan_G = -G;   // negative: plan of the fast kernel
    }
    if (!g->abort_flag) {
        int rc = dalloc(&g->abort_flag, 1);
        if (rc) return rc;
    }
    SDEV_CK(hipMemsetAsync(g->abort_flag, 0, 4, g->stream));
    if (!g->ring32_valid) {
        int rc = fast_sync_rings(g, true);
        if (rc) return rc;
        g->ring32_valid = 1;
    }
    FastArgs a;
    memset(&a, 0, sizeof(a));
    a.n = g->n; a.din = g->din; a.L = g->L; a.nb = nb; a.G = G; a.wpg = wpg; a.micro = g->micro; a.K = K;
    a.stype = g->stype; a.st0 = g->step; a.t0 = t0; a.T = T;
    a.gain = g->d.gain; a.dt = g->d.dt;
    memcpy(a.np, g->d.np, sizeof(a.np));
    a.W_in = g->W_in; a.u = u_series;
    a.rp = g->rp; a.col = g->col; a.val = g->val;
    a.edge_decay = g->edge_decay; a.edge_tr = g->edge_tr;
    a.V = g->V; a.S = g->S[g->cur]; a.S_out = g->S[g->cur ^ 1]; a.tprev = g->tprev;
    a.ring = g->ring32; a.coef = g->coef;
    a.xch = g->xch; a.abort_flag = g->abort_flag;
    a.rec_V = rec_V; a.rec_S = rec_S;
    const int Le = g->L + (g->L & 1);
    const size_t base = (size_t)Le * 4 + (size_t)2 * K * nb * 4 + 4 * (size_t)nb * 8 + nw * 4 + 64;
    int64_t me = 0;
    for (int i0 = 0; i0 < g->n; i0 += nb)
        me = std::max<int64_t>(me, g->rp_host[std::min(g->n, i0 + nb)] - g->rp_host[i0]);
    const size_t cap = 160 * 1024 - 2048;
    const int ne_lds = (base + (size_t)me * 12 <= cap) ? (int)std::max<int64_t>(me, 1) : 0;
    const size_t lds = base + (size_t)ne_lds * 12;
    a.ne_lds = ne_lds;
    void* args[] = {&a};
    SDEV_CK(hipFuncSetAttribute((const void*)k_fast, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    SDEV_CK(hipLaunchCooperativeKernel((const void*)k_fast, dim3(G), dim3(kFastThreads), args, (unsigned)lds,
                                       g->stream));
    g->step += (long long)T * g->micro;
    g->msteps += (size_t)T;
    g->cur ^= 1;
    g->ring64_valid = 0;
    unsigned flag = 0;
    SDEV_CK(hipMemcpyAsync(&flag, g->abort_flag, 4, hipMemcpyDeviceToHost, g->stream));
    SDEV_CK(hipStreamSynchronize(g->stream));
    if (flag) { set_error("fast step kernel timed out waiting for the spike exchange"); return SPIRES_DEV_EINTERNAL; }
    return 0;
}

}  // namespace sdev
