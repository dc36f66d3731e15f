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
#include "dev_exchange.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

namespace sdev {

constexpr int kPersistThreads = 128;

struct PersistArgs {
    int n, din, L, G, wpg, micro, stype;
    int ne_lds;                   // LDS capacity for CSR edges per workgroup (0: read CSR from HBM)
    int rpm, cw;                  // row pull: 0 edge list, 1 spike-word bitmaps (cw words per scan part)
    int win;                      // k_persist_tp: product window in 32-lag chunks (0: whole rows)
    int hl;                       // k_persist_tp: the workgroup's history rows kept in LDS (whole rows only)
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
    const unsigned* rbm;          // [n][ceil(n/32)] row bitmaps in column order (rpm 1)
    // Device step clock (replayed time blocks, k_persist_tp only): when set,
    // st0 = clk[0], t0 = clk[1], spikes are read from S (clk[2] even) or S_out
    // (odd) and written to the other, and records start at row t0.
    const long long* clk;
};

// FP64 multiply and add with their issue order pinned: each is one IEEE
// round-to-nearest operation (the same as the compiler's v_mul_f64 /
// v_add_f64), but the compiler may not move a product next to the add that
// consumes it, which would expose the multiply latency on the add chain.
__device__ __forceinline__ void add_f64_ordered(double& h, double p) {
    asm volatile("v_add_f64 %0, %0, %1" : "+v"(h) : "v"(p));
}
// Sixteen dependent adds h += b[0].x, b[0].y, ..., b[7].y in one statement:
// back-to-back dependent v_add_f64 need no wait state, but the compiler puts
// one between consecutive inline-asm statements, which costs about as much as
// the add itself on a latency-bound chain.
__device__ __forceinline__ void add16_f64_ordered(double& h, const double2* b) {
    asm volatile(
        "v_add_f64 %0, %0, %1\n\tv_add_f64 %0, %0, %2\n\tv_add_f64 %0, %0, %3\n\tv_add_f64 %0, %0, %4\n\t"
        "v_add_f64 %0, %0, %5\n\tv_add_f64 %0, %0, %6\n\tv_add_f64 %0, %0, %7\n\tv_add_f64 %0, %0, %8\n\t"
        "v_add_f64 %0, %0, %9\n\tv_add_f64 %0, %0, %10\n\tv_add_f64 %0, %0, %11\n\tv_add_f64 %0, %0, %12\n\t"
        "v_add_f64 %0, %0, %13\n\tv_add_f64 %0, %0, %14\n\tv_add_f64 %0, %0, %15\n\tv_add_f64 %0, %0, %16"
        : "+v"(h)
        : "v"(b[0].x), "v"(b[0].y), "v"(b[1].x), "v"(b[1].y), "v"(b[2].x), "v"(b[2].y), "v"(b[3].x), "v"(b[3].y),
          "v"(b[4].x), "v"(b[4].y), "v"(b[5].x), "v"(b[5].y), "v"(b[6].x), "v"(b[6].y), "v"(b[7].x), "v"(b[7].y));
}
__device__ __forceinline__ void mul_f64_ordered(double& r, double a, double b) {
    asm volatile("v_mul_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
}

// One 16-lag block of a history row: 8 two-lag ring reads (.y lag q+2j,
// .x lag q+2j+1) and the matching coefficient pairs.  sb - q is odd, so the
// pair of lags (q+2j, q+2j+1) is the aligned slot pair (pb-j)*2, (pb-j)*2+1
// with pb = (sb-q-1)/2: one base address per block, immediate offsets per pair.
__device__ __forceinline__ void gl_block_load(const double2* __restrict__ row2, const double* __restrict__ cs, int kb,
                                              int sb, int q, double2* v, double2* c) {
    const double2* rb = row2 + ((sb - q - 1) >> 1) - 7;
    const double* cb = cs + kb + q;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        v[j] = rb[7 - j];
        c[j] = make_double2(cb[2 * j], cb[2 * j + 1]);
    }
}
// The final, partial block of a segment: pair indices below zero (lags past
// the segment) are clamped to slot pair 0 and their products are never added.
__device__ __forceinline__ void gl_block_load_clamped(const double2* __restrict__ row2, const double* __restrict__ cs,
                                                      int kb, int sb, int q, double2* v, double2* c) {
    const int pb = (sb - q - 1) >> 1;
    const double* cb = cs + kb + q;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        v[j] = row2[max(pb - j, 0)];
        c[j] = make_double2(cb[2 * j], cb[2 * j + 1]);
    }
}
__device__ __forceinline__ void gl_block_products(double* pr, const double2* v, const double2* c) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        mul_f64_ordered(pr[2 * j], c[j].x, v[j].y);
        mul_f64_ordered(pr[2 * j + 1], c[j].y, v[j].x);
    }
}
// h += pr (block b, in lag order) while the products of block b+1 replace them.
__device__ __forceinline__ void gl_block_step(double& h, double* pr, const double2* v, const double2* c) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        add_f64_ordered(h, pr[2 * j]);
        mul_f64_ordered(pr[2 * j], c[j].x, v[j].y);
        add_f64_ordered(h, pr[2 * j + 1]);
        mul_f64_ordered(pr[2 * j + 1], c[j].y, v[j].x);
    }
}

// h += sum_{q<cnt} c[kb+q] * row[sb-q]  over one neuron's LDS history row
// (lag-contiguous layout [NB][LP], LP even): slots descend as lags ascend, so
// one 16-byte read returns two consecutive lags.  The sum keeps one dependent
// add per lag in ascending lag order (flif_gl.c:66-72).  Blocks of 16 lags are
// software-pipelined: the ring and coefficient reads of block b+1 are in
// flight while block b is multiplied, and the products of block b+1 are
// formed while the adds of block b run.  cs is the coefficient table in LDS,
// zero-padded by 16 entries past L-1 for the clamped final block.
__device__ __forceinline__ double gl_seg_row(double h, const double* __restrict__ row, const double* __restrict__ cs,
                                             int kb, int sb, int cnt) {
    if (cnt <= 0) return h;
    int q = 0;
    if ((sb & 1) == 0) {   // peel one lag so that pairs (s-1, s) start at odd s
        h = h + cs[kb] * row[sb];
        q = 1;
    }
    const double2* row2 = reinterpret_cast<const double2*>(row);
    const int nblk = (cnt - q) >> 4;
    double pr[16];
    double2 vA[8], cA[8], vB[8], cB[8];
    if (nblk > 0) {
        gl_block_load(row2, cs, kb, sb, q, vA, cA);
        if (nblk == 1) {
            gl_block_products(pr, vA, cA);
        } else {
            gl_block_load(row2, cs, kb, sb, q + 16, vB, cB);
            gl_block_products(pr, vA, cA);                      // pr: block 0, B: block 1
            int b = 1;
            for (; b + 2 <= nblk - 1; b += 2) {
                gl_block_load(row2, cs, kb, sb, q + 16 * (b + 1), vA, cA);
                gl_block_step(h, pr, vB, cB);                   // adds block b-1, products block b
                gl_block_load(row2, cs, kb, sb, q + 16 * (b + 2), vB, cB);
                gl_block_step(h, pr, vA, cA);                   // adds block b, products block b+1
            }
            if (b + 1 <= nblk - 1) {
                gl_block_load(row2, cs, kb, sb, q + 16 * (b + 1), vA, cA);
                gl_block_step(h, pr, vB, cB);
                gl_block_step(h, pr, vA, cA);
            } else {
                gl_block_step(h, pr, vB, cB);
            }
        }
        q += 16 * nblk;
    }
    const int last = cnt - q;                                    // lags of the partial block, 0..15
    if (last > 0) gl_block_load_clamped(row2, cs, kb, sb, q, vA, cA);   // in flight under the adds below
    if (nblk > 0) {
#pragma unroll
        for (int j = 0; j < 16; ++j) add_f64_ordered(h, pr[j]);
    }
    if (last > 0) {
        gl_block_products(pr, vA, cA);
#pragma unroll
        for (int j = 0; j < 15; ++j)
            if (j < last) add_f64_ordered(h, pr[j]);
    }
    return h;
}

// h += sum_{q<cnt} c[kb+q] * ring[(sb-q)*stride], ascending q, one multiply and
// one add per lag (flif_gl.c:66-72), over a history ring in HBM.  The ring is
// read in blocks of 16 lags with kHbmBlocks blocks in flight (the wave keeps
// 16 * kHbmBlocks samples in registers), so the step waits on HBM bandwidth
// rather than on one load latency per block; the final partial block is
// loaded as a whole before its adds.
constexpr int kHbmBlocks = 4;
__device__ __forceinline__ double gl_seg_hbm(double h, const double* __restrict__ ring, const double* __restrict__ c,
                                             int kb, int sb, int cnt, size_t stride) {
    constexpr int B = 16;
    const int nblk = cnt / B;
    double v[kHbmBlocks][B];
    auto load = [&](double* buf, int blk) {
        const double* rb = ring + (size_t)(sb - blk * B) * stride;
#pragma unroll
        for (int j = 0; j < B; ++j) buf[j] = rb[-(ptrdiff_t)(j * stride)];
    };
#pragma unroll
    for (int u = 0; u < kHbmBlocks; ++u)
        if (u < nblk) load(v[u], u);
    for (int b0 = 0; b0 < nblk; b0 += kHbmBlocks) {
#pragma unroll
        for (int u = 0; u < kHbmBlocks; ++u) {
            const int blk = b0 + u;
            if (blk < nblk) {
#pragma unroll
                for (int j = 0; j < B; ++j) h = h + c[kb + blk * B + j] * v[u][j];
                if (blk + kHbmBlocks < nblk) load(v[u], blk + kHbmBlocks);
            }
        }
    }
    const int q = nblk * B, rem = cnt - q;
    if (rem > 0) {
        double t[B];
#pragma unroll
        for (int j = 0; j < B; ++j) t[j] = j < rem ? ring[(size_t)(sb - q - j) * stride] : 0.0;
#pragma unroll
        for (int j = 0; j < B; ++j)
            if (j < rem) h = h + c[kb + q + j] * t[j];
    }
    return h;
}

template <int NT, int NB, bool RL>
__global__ void __launch_bounds__(kPersistThreads) __attribute__((amdgpu_waves_per_eu(1, 1))) k_persist(PersistArgs a) {
    extern __shared__ __align__(16) double lds[];   // 16-byte aligned: b128 ring-pair reads
    constexpr bool GL = (NT == NT_FLIF_GL);
    // wave index and abort flag are read through readfirstlane so the compiler
    // treats them as wave-uniform (scalar branches, scalar loop counters).
    const int g = blockIdx.x, tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int i0 = g * NB;
    const int nloc = min(NB, a.n - i0);
    const int L = a.L;
    const int wpg = a.wpg;
    const int nw = a.G * wpg;
    // ---- LDS layout
    const int LP = L + (L & 1) + 2;                         // padded row length (even)
    double* ring_s = lds;                                   // [NB][LP]     (RL)
    double* coef_s = ring_s + (RL ? (size_t)LP * NB : 0);   // [L + 16]     (GL; zero past L-1)
    double* rec_s = coef_s + (GL ? L + 16 : 0);             // [2][NB]
    double* ext_s = rec_s + 2 * NB;                         // [2][NB]
    double* eval_s = ext_s + 2 * NB;                        // [ne_lds]
    unsigned* bits_s = (unsigned*)(eval_s + a.ne_lds);      // [nw]
    unsigned* eidx_s = bits_s + nw;                         // [ne_lds]
    __shared__ int abort_s;
    if (tid == 0) abort_s = 0;
    if (GL) {
        for (int k = tid; k < L + 16; k += kPersistThreads) coef_s[k] = k < L ? a.coef[k] : 0.0;
        if (RL)
            for (int e = tid; e < L * NB; e += kPersistThreads) {
                const int s = e / NB, j = e % NB;
                ring_s[j * LP + s] = (j < nloc) ? a.ring[(size_t)s * a.n + i0 + j] : 0.0;
            }
    }
    const int64_t E0 = a.rp[i0];
    const int ne = (int)(a.rp[i0 + nloc] - E0);
    const bool csr_lds = a.ne_lds > 0;
    if (csr_lds)
        for (int e = tid; e < ne; e += kPersistThreads) {
            eidx_s[e] = encode_col(a.col[E0 + e], NB, wpg);
            eval_s[e] = a.val[E0 + e];
        }
    // ---- per-lane state: wave 0 lane = neuron, wave 1 lane = row puller
    const bool own = (wave == 0) && (lane < nloc);
    const int i = i0 + lane;
    double V = own ? a.V[i] : 0.0;
    double tp = (own && GL) ? a.tprev[i] : 0.0;
    bool spk = own ? (a.S[i] != 0) : false;
    int rs = 0, re = 0;
    constexpr int kMaxDinReg = 4;     // W_in rows / inputs held in registers (larger din reads memory)
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
    // Model parameters in VGPRs (uniform values, but scalar registers are the
    // scarce resource in this kernel: keeping them there forces spills).
    double p[7];
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        p[j] = a.np[j];
        asm volatile("" : "+v"(p[j]));
    }
    __syncthreads();

    long long st = a.st0;
    // ring head (st mod L) and history length min(st, L-1), advanced per micro-step
    int head = GL ? (int)(st % L) : 0;
    int lim = GL ? (st < L ? (int)st : L - 1) : 0;
    for (long long t = 0; t < a.T; ++t) {
        const long long tt = a.t0 + t;
        const int tb = (int)(t & 1);
        for (int mi = 0; mi < a.micro; ++mi, ++st) {
            const int sb = (int)(st & 1);
            if (wave == 0) {
                // ---------------- GL history (overlaps the exchange on wave 1)
                // Every lane runs the sum in uniform control flow (lanes without
                // a live neuron read a valid row and discard the result), so the
                // coefficient loads stay scalar and the loop stays unmasked.
                double H = 0.0;
                if (GL) {
                    const int n1 = lim < head ? lim : head;
                    if (RL) {
                        const double* row = ring_s + (lane < NB ? lane : NB - 1) * LP;
                        H = gl_seg_row(H, row, coef_s, 1, head - 1, n1);
                        H = gl_seg_row(H, row, coef_s, n1 + 1, head - n1 - 1 + L, lim - n1);
                    } else {
                        const double* col = a.ring + i0 + (lane < nloc ? lane : nloc - 1);
                        H = gl_seg_hbm(H, col, coef_s, 1, head - 1, n1, (size_t)a.n);
                        H = gl_seg_hbm(H, col, coef_s, n1 + 1, head - n1 - 1 + L, lim - n1, (size_t)a.n);
                    }
                }
                __syncthreads();   // rec / ext of this step ready
                if (__builtin_amdgcn_readfirstlane(abort_s)) break;
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
                            const double Vp = RL ? ring_s[lane * LP + sp] : a.ring[(size_t)sp * a.n + i];
                            const double rhs = (-(Vp - p[2]) / p[3]) + x + p[5];
                            V = p[4] * rhs - H;
                            if (V >= p[0]) { V = p[1]; spk = true; tp = 0.0; }
                        } else {
                            V = p[1];
                        }
                        ring_slot_write = 1;
                    }
                }
                // publish first (the other workgroups wait on it)
                publish_spikes(a.xch, nw, g, wpg, st, lane, __ballot(own && spk));
                if (own) {
                    if (GL && ring_slot_write) {
                        if (RL) ring_s[lane * LP + head] = V;
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
                const bool aborted = !gather_spikes(a.xch, nw, NB, wpg, a.n, st, a.st0, a.S, bits_s,
                                                    a.abort_flag, &abort_s, lane);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                // ---------------- CSR row pull and external input
                if (!aborted && lane < nloc) {
                    rec_s[sb * NB + lane] = row_pull(a.stype, csr_lds, rs, re, E0, eidx_s, eval_s, a.col, a.val,
                                                     bits_s, NB, wpg, a.edge_decay, a.edge_tr);
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
                if (__builtin_amdgcn_readfirstlane(abort_s)) break;
            }
            if (GL) {
                head = head + 1 == L ? 0 : head + 1;
                if (lim < L - 1) ++lim;
            }
        }
        if (__builtin_amdgcn_readfirstlane(abort_s)) break;
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
            if (j < nloc) a.ring[(size_t)s * a.n + i0 + j] = ring_s[j * LP + s];
        }
    }
}


// ---------------------------------------------------------------------------
// FLIF-GL persistent kernel with a trailing product stage.
//
// The GL history of micro-step m is H_m = sum_{k=1..lim} c_k V_{m-k}, summed
// in ascending k with one dependent add per lag (flif_gl.c:66-72).  The
// products are independent of each other, so they are formed off the chain:
//   wave 0        (chain)     H = c_1 V_{m-1}, then adds P_m[2..lim] from LDS in
//                             lag order, then the neuron tail (update, publish,
//                             history write);
//   waves 1,2,3,5 (products)  during step m write P_{m+1}[k] = c_k V_{m+1-k}
//                             into the same LDS rows, each 32-lag chunk only
//                             after the chain of step m has consumed it;
//   wave 4        (exchange)  spike exchange, CSR row pull and external input.
// The products of step m+1 need V_{m-1} (lag 2; handed over through a
// two-slot LDS stage by the tail of step m-1) and V_{m-2} and older (read from
// the history ring in HBM, [L][N], written by the tail two or more steps
// earlier and ordered by the workgroup barrier).  The chain wave never waits
// on the producers except at the one barrier per step; the producers follow
// its progress through an LDS counter.  Products and sums are plain IEEE
// double operations in the reference's order, so the result is bit-identical
// to the single-wave sum.
constexpr int kTpThreads = 384;
constexpr int kTpChunk = 32;
constexpr int kTpProducers = 4;
// Wave roles: 0 the chain, kTpXchWave the exchange, the other four the products.
// Waves w and w+4 share a SIMD.  The chain issues dependent adds back to back,
// so a wave beside it gets few issue slots: the exchange wave sits there rather
// than a product wave, whose last chunks would then finish after the chain.
constexpr int kTpXchWave = 4;
constexpr int kTpLoadDepth = 8;      // ring-load buffers of a product wave (chunks in flight + 1)
constexpr int kTpProdUnit = 2;       // chunks a product wave writes per poll of the chain's progress
constexpr int kRegEdges = 12;        // row pull from registers (16 neurons per workgroup): rows of at most 12 edges
constexpr long long kTpSpinLimit = 1ll << 24;
constexpr int kBitmapBatch = 16;     // weights loaded per batch of the bitmap pull from HBM (half from LDS)

// Chain progress: 64 * (micro-step in launch) + chunks done; compared modulo 2^32.
__device__ __forceinline__ unsigned lds_counter_load(const unsigned* f) {
    return __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_counter_store(unsigned* f, unsigned v) {
    __hip_atomic_store(f, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// Product-row stride in doubles: every chunk of 32 lags starting at lag 2 fits,
// and 2*PS = 4 (mod 32) dwords keeps the rows' 16-byte reads conflict-free.
__host__ __device__ __forceinline__ int tp_row_stride(int L) { return 32 * ((L - 2 + 31) / 32) + 2; }
// History row in LDS (k_persist_tp, whole rows): slot s of the ring at entry
// L - 1 - s, so ascending lags are ascending entries, and the entries past L - 1
// repeat entries 0, 1, ... (mod L), so that up to five consecutive lags read
// from consecutive entries without a wrap.  The stride is odd: with rows a
// multiple of 8 doubles apart the product waves' reads of 16 rows pile onto a
// few LDS banks (N=4096, L=512: 4.1 us per step at stride L+8, 3.23 at L+7).
__host__ __device__ __forceinline__ int tp_hist_stride(int L) { return (L + 5) | 1; }

// The chain of one product row: h += P[k] for k = 2 .. 2 + 32 * nch - 1 in
// ascending order, one dependent v_add_f64 per lag (flif_gl.c:66-72); entries
// past the step's last lag are +0.0, which leaves h unchanged (h is never -0.0:
// it starts as 0.0 + c_1 V).  A wave issues nothing else while one of its FP64
// adds is in flight, so every other instruction of the chain wave lengthens the
// chain: the 16-byte reads of lag pairs go out in bursts of 8 right after a wait
// (a read issued between two adds costs about twice as much as one in a burst),
// 14 pairs ahead of the adds, into a 32-pair ring in v[64:191].  The waits
// count reads only, which is conservative when a counter store is outstanding,
// and at most 15 LDS operations are ever in flight.  The last bursts read up to
// 224 bytes past the row (the next row, or the coefficient table after the last
// one) and are waited for before the statement ends.  Once every read of chunk
// c (32 lags) has completed, before its last adds, the LDS counter at byte
// address ca is set to cv + c: the producers may then overwrite the chunk.
// va is the byte address of P[2] in LDS (16-byte aligned).
__device__ __forceinline__ double chain_chunks(double h, unsigned va, int nch, unsigned ca, unsigned cv) {
    asm volatile(
        "s_waitcnt lgkmcnt(0)\n\t"
        "ds_read_b128 v[64:67], %[va] offset:0\n\t"
        "ds_read_b128 v[68:71], %[va] offset:16\n\t"
        "ds_read_b128 v[72:75], %[va] offset:32\n\t"
        "ds_read_b128 v[76:79], %[va] offset:48\n\t"
        "ds_read_b128 v[80:83], %[va] offset:64\n\t"
        "ds_read_b128 v[84:87], %[va] offset:80\n\t"
        "ds_read_b128 v[88:91], %[va] offset:96\n\t"
        "ds_read_b128 v[92:95], %[va] offset:112\n\t"
        "ds_read_b128 v[96:99], %[va] offset:128\n\t"
        "ds_read_b128 v[100:103], %[va] offset:144\n\t"
        "ds_read_b128 v[104:107], %[va] offset:160\n\t"
        "ds_read_b128 v[108:111], %[va] offset:176\n\t"
        "ds_read_b128 v[112:115], %[va] offset:192\n\t"
        "ds_read_b128 v[116:119], %[va] offset:208\n\t"
        "L_chain_loop_%=:\n\t"
        "s_cmp_lt_u32 %[n], 2\n\t"
        "s_cbranch_scc1 L_chain_half_%=\n\t"
        "s_waitcnt lgkmcnt(6)\n\t"
        "ds_read_b128 v[120:123], %[va] offset:224\n\t"
        "ds_read_b128 v[124:127], %[va] offset:240\n\t"
        "ds_read_b128 v[128:131], %[va] offset:256\n\t"
        "ds_read_b128 v[132:135], %[va] offset:272\n\t"
        "ds_read_b128 v[136:139], %[va] offset:288\n\t"
        "ds_read_b128 v[140:143], %[va] offset:304\n\t"
        "ds_read_b128 v[144:147], %[va] offset:320\n\t"
        "ds_read_b128 v[148:151], %[va] offset:336\n\t"
        "v_add_f64 %[h], %[h], v[64:65]\n\t"
        "v_add_f64 %[h], %[h], v[66:67]\n\t"
        "v_add_f64 %[h], %[h], v[68:69]\n\t"
        "v_add_f64 %[h], %[h], v[70:71]\n\t"
        "v_add_f64 %[h], %[h], v[72:73]\n\t"
        "v_add_f64 %[h], %[h], v[74:75]\n\t"
        "v_add_f64 %[h], %[h], v[76:77]\n\t"
        "v_add_f64 %[h], %[h], v[78:79]\n\t"
        "v_add_f64 %[h], %[h], v[80:81]\n\t"
        "v_add_f64 %[h], %[h], v[82:83]\n\t"
        "v_add_f64 %[h], %[h], v[84:85]\n\t"
        "v_add_f64 %[h], %[h], v[86:87]\n\t"
        "v_add_f64 %[h], %[h], v[88:89]\n\t"
        "v_add_f64 %[h], %[h], v[90:91]\n\t"
        "v_add_f64 %[h], %[h], v[92:93]\n\t"
        "v_add_f64 %[h], %[h], v[94:95]\n\t"
        "s_waitcnt lgkmcnt(6)\n\t"
        "ds_write_b32 %[ca], %[cv]\n\t"
        "v_add_u32 %[cv], 1, %[cv]\n\t"
        "ds_read_b128 v[152:155], %[va] offset:352\n\t"
        "ds_read_b128 v[156:159], %[va] offset:368\n\t"
        "ds_read_b128 v[160:163], %[va] offset:384\n\t"
        "ds_read_b128 v[164:167], %[va] offset:400\n\t"
        "ds_read_b128 v[168:171], %[va] offset:416\n\t"
        "ds_read_b128 v[172:175], %[va] offset:432\n\t"
        "ds_read_b128 v[176:179], %[va] offset:448\n\t"
        "ds_read_b128 v[180:183], %[va] offset:464\n\t"
        "v_add_f64 %[h], %[h], v[96:97]\n\t"
        "v_add_f64 %[h], %[h], v[98:99]\n\t"
        "v_add_f64 %[h], %[h], v[100:101]\n\t"
        "v_add_f64 %[h], %[h], v[102:103]\n\t"
        "v_add_f64 %[h], %[h], v[104:105]\n\t"
        "v_add_f64 %[h], %[h], v[106:107]\n\t"
        "v_add_f64 %[h], %[h], v[108:109]\n\t"
        "v_add_f64 %[h], %[h], v[110:111]\n\t"
        "v_add_f64 %[h], %[h], v[112:113]\n\t"
        "v_add_f64 %[h], %[h], v[114:115]\n\t"
        "v_add_f64 %[h], %[h], v[116:117]\n\t"
        "v_add_f64 %[h], %[h], v[118:119]\n\t"
        "v_add_f64 %[h], %[h], v[120:121]\n\t"
        "v_add_f64 %[h], %[h], v[122:123]\n\t"
        "v_add_f64 %[h], %[h], v[124:125]\n\t"
        "v_add_f64 %[h], %[h], v[126:127]\n\t"
        "s_waitcnt lgkmcnt(6)\n\t"
        "ds_read_b128 v[184:187], %[va] offset:480\n\t"
        "ds_read_b128 v[188:191], %[va] offset:496\n\t"
        "ds_read_b128 v[64:67], %[va] offset:512\n\t"
        "ds_read_b128 v[68:71], %[va] offset:528\n\t"
        "ds_read_b128 v[72:75], %[va] offset:544\n\t"
        "ds_read_b128 v[76:79], %[va] offset:560\n\t"
        "ds_read_b128 v[80:83], %[va] offset:576\n\t"
        "ds_read_b128 v[84:87], %[va] offset:592\n\t"
        "v_add_f64 %[h], %[h], v[128:129]\n\t"
        "v_add_f64 %[h], %[h], v[130:131]\n\t"
        "v_add_f64 %[h], %[h], v[132:133]\n\t"
        "v_add_f64 %[h], %[h], v[134:135]\n\t"
        "v_add_f64 %[h], %[h], v[136:137]\n\t"
        "v_add_f64 %[h], %[h], v[138:139]\n\t"
        "v_add_f64 %[h], %[h], v[140:141]\n\t"
        "v_add_f64 %[h], %[h], v[142:143]\n\t"
        "v_add_f64 %[h], %[h], v[144:145]\n\t"
        "v_add_f64 %[h], %[h], v[146:147]\n\t"
        "v_add_f64 %[h], %[h], v[148:149]\n\t"
        "v_add_f64 %[h], %[h], v[150:151]\n\t"
        "v_add_f64 %[h], %[h], v[152:153]\n\t"
        "v_add_f64 %[h], %[h], v[154:155]\n\t"
        "v_add_f64 %[h], %[h], v[156:157]\n\t"
        "v_add_f64 %[h], %[h], v[158:159]\n\t"
        "s_waitcnt lgkmcnt(6)\n\t"
        "ds_write_b32 %[ca], %[cv]\n\t"
        "v_add_u32 %[cv], 1, %[cv]\n\t"
        "ds_read_b128 v[88:91], %[va] offset:608\n\t"
        "ds_read_b128 v[92:95], %[va] offset:624\n\t"
        "ds_read_b128 v[96:99], %[va] offset:640\n\t"
        "ds_read_b128 v[100:103], %[va] offset:656\n\t"
        "ds_read_b128 v[104:107], %[va] offset:672\n\t"
        "ds_read_b128 v[108:111], %[va] offset:688\n\t"
        "ds_read_b128 v[112:115], %[va] offset:704\n\t"
        "ds_read_b128 v[116:119], %[va] offset:720\n\t"
        "v_add_f64 %[h], %[h], v[160:161]\n\t"
        "v_add_f64 %[h], %[h], v[162:163]\n\t"
        "v_add_f64 %[h], %[h], v[164:165]\n\t"
        "v_add_f64 %[h], %[h], v[166:167]\n\t"
        "v_add_f64 %[h], %[h], v[168:169]\n\t"
        "v_add_f64 %[h], %[h], v[170:171]\n\t"
        "v_add_f64 %[h], %[h], v[172:173]\n\t"
        "v_add_f64 %[h], %[h], v[174:175]\n\t"
        "v_add_f64 %[h], %[h], v[176:177]\n\t"
        "v_add_f64 %[h], %[h], v[178:179]\n\t"
        "v_add_f64 %[h], %[h], v[180:181]\n\t"
        "v_add_f64 %[h], %[h], v[182:183]\n\t"
        "v_add_f64 %[h], %[h], v[184:185]\n\t"
        "v_add_f64 %[h], %[h], v[186:187]\n\t"
        "v_add_f64 %[h], %[h], v[188:189]\n\t"
        "v_add_f64 %[h], %[h], v[190:191]\n\t"
        "v_add_u32 %[va], 0x200, %[va]\n\t"
        "s_sub_u32 %[n], %[n], 2\n\t"
        "s_branch L_chain_loop_%=\n\t"
        "L_chain_half_%=:\n\t"
        "s_cmp_eq_u32 %[n], 0\n\t"
        "s_cbranch_scc1 L_chain_end_%=\n\t"
        "s_waitcnt lgkmcnt(6)\n\t"
        "ds_read_b128 v[120:123], %[va] offset:224\n\t"
        "ds_read_b128 v[124:127], %[va] offset:240\n\t"
        "ds_read_b128 v[128:131], %[va] offset:256\n\t"
        "ds_read_b128 v[132:135], %[va] offset:272\n\t"
        "ds_read_b128 v[136:139], %[va] offset:288\n\t"
        "ds_read_b128 v[140:143], %[va] offset:304\n\t"
        "ds_read_b128 v[144:147], %[va] offset:320\n\t"
        "ds_read_b128 v[148:151], %[va] offset:336\n\t"
        "v_add_f64 %[h], %[h], v[64:65]\n\t"
        "v_add_f64 %[h], %[h], v[66:67]\n\t"
        "v_add_f64 %[h], %[h], v[68:69]\n\t"
        "v_add_f64 %[h], %[h], v[70:71]\n\t"
        "v_add_f64 %[h], %[h], v[72:73]\n\t"
        "v_add_f64 %[h], %[h], v[74:75]\n\t"
        "v_add_f64 %[h], %[h], v[76:77]\n\t"
        "v_add_f64 %[h], %[h], v[78:79]\n\t"
        "v_add_f64 %[h], %[h], v[80:81]\n\t"
        "v_add_f64 %[h], %[h], v[82:83]\n\t"
        "v_add_f64 %[h], %[h], v[84:85]\n\t"
        "v_add_f64 %[h], %[h], v[86:87]\n\t"
        "v_add_f64 %[h], %[h], v[88:89]\n\t"
        "v_add_f64 %[h], %[h], v[90:91]\n\t"
        "v_add_f64 %[h], %[h], v[92:93]\n\t"
        "v_add_f64 %[h], %[h], v[94:95]\n\t"
        "s_waitcnt lgkmcnt(6)\n\t"
        "ds_write_b32 %[ca], %[cv]\n\t"
        "v_add_u32 %[cv], 1, %[cv]\n\t"
        "ds_read_b128 v[152:155], %[va] offset:352\n\t"
        "ds_read_b128 v[156:159], %[va] offset:368\n\t"
        "ds_read_b128 v[160:163], %[va] offset:384\n\t"
        "ds_read_b128 v[164:167], %[va] offset:400\n\t"
        "ds_read_b128 v[168:171], %[va] offset:416\n\t"
        "ds_read_b128 v[172:175], %[va] offset:432\n\t"
        "ds_read_b128 v[176:179], %[va] offset:448\n\t"
        "ds_read_b128 v[180:183], %[va] offset:464\n\t"
        "v_add_f64 %[h], %[h], v[96:97]\n\t"
        "v_add_f64 %[h], %[h], v[98:99]\n\t"
        "v_add_f64 %[h], %[h], v[100:101]\n\t"
        "v_add_f64 %[h], %[h], v[102:103]\n\t"
        "v_add_f64 %[h], %[h], v[104:105]\n\t"
        "v_add_f64 %[h], %[h], v[106:107]\n\t"
        "v_add_f64 %[h], %[h], v[108:109]\n\t"
        "v_add_f64 %[h], %[h], v[110:111]\n\t"
        "v_add_f64 %[h], %[h], v[112:113]\n\t"
        "v_add_f64 %[h], %[h], v[114:115]\n\t"
        "v_add_f64 %[h], %[h], v[116:117]\n\t"
        "v_add_f64 %[h], %[h], v[118:119]\n\t"
        "v_add_f64 %[h], %[h], v[120:121]\n\t"
        "v_add_f64 %[h], %[h], v[122:123]\n\t"
        "v_add_f64 %[h], %[h], v[124:125]\n\t"
        "v_add_f64 %[h], %[h], v[126:127]\n\t"
        "L_chain_end_%=:\n\t"
        "s_waitcnt lgkmcnt(0)\n\t"
        : [h] "+v"(h), [va] "+v"(va), [n] "+s"(nch), [cv] "+v"(cv)
        : [ca] "v"(ca)
        : "memory", "scc",
          "v64", "v65", "v66", "v67", "v68", "v69", "v70", "v71", "v72", "v73", "v74", "v75", "v76", "v77", "v78", "v79",
          "v80", "v81", "v82", "v83", "v84", "v85", "v86", "v87", "v88", "v89", "v90", "v91", "v92", "v93", "v94", "v95",
          "v96", "v97", "v98", "v99", "v100", "v101", "v102", "v103", "v104", "v105", "v106", "v107", "v108", "v109", "v110", "v111",
          "v112", "v113", "v114", "v115", "v116", "v117", "v118", "v119", "v120", "v121", "v122", "v123", "v124", "v125", "v126", "v127",
          "v128", "v129", "v130", "v131", "v132", "v133", "v134", "v135", "v136", "v137", "v138", "v139", "v140", "v141", "v142", "v143",
          "v144", "v145", "v146", "v147", "v148", "v149", "v150", "v151", "v152", "v153", "v154", "v155", "v156", "v157", "v158", "v159",
          "v160", "v161", "v162", "v163", "v164", "v165", "v166", "v167", "v168", "v169", "v170", "v171", "v172", "v173", "v174", "v175",
          "v176", "v177", "v178", "v179", "v180", "v181", "v182", "v183", "v184", "v185", "v186", "v187", "v188", "v189", "v190", "v191");
    return h;
}
// Byte address of an LDS object (its offset in the workgroup's LDS).
__device__ __forceinline__ unsigned lds_offset(const void* p) {
    return (unsigned)(uintptr_t)(const __attribute__((address_space(3))) char*)p;
}

template <int R_>
struct RoleTag { static constexpr int value = R_; };

template <int NB, bool WIN>
__device__ __forceinline__ void tp_body(const PersistArgs& a, const int g) {
    constexpr int C = kTpChunk;
    constexpr int PARTS = 64 * kTpProducers / NB;   // producer lanes per neuron
    constexpr int PER = C / PARTS;                  // lags per producer lane per chunk
    static_assert(PER >= 2 && PER % 2 == 0, "producer lanes write lag pairs");
    extern __shared__ __align__(16) double lds[];   // 16-byte aligned: b128 product reads
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int i0 = g * NB;
    const int nloc = min(NB, a.n - i0);
    const int n = a.n, L = a.L, wpg = a.wpg, nw = a.G * wpg;
    const int PS = tp_row_stride(L);
    const int W = WIN ? a.win : 0;                     // product window (chunks); 0: whole rows
    const int PSr = W ? 32 * W + 2 : PS;               // product row stride (2*PSr = 4 mod 32 dwords)
    // ---- LDS layout
    double* P = lds;                                   // [NB][PSr]  P[r][k], or chunk slots when W > 0
    double* coef_s = P + (size_t)NB * PSr;             // [PS]       zero past L-1
    const bool HL = !WIN && a.hl;                      // history rows in LDS
    const int HR = tp_hist_stride(L);
    double* hist_s = coef_s + PS;                      // [NB][HR]   when HL
    double* stage = hist_s + (HL ? (size_t)NB * HR : 0);   // [2][NB]    V of the last two steps
    double* rec_s = stage + 2 * NB;                    // [2][NB]
    double* ext_s = rec_s + 2 * NB;                    // [2][NB]
    double* eval_s = ext_s + 2 * NB;                   // [ne_lds]
    unsigned* bits_s = (unsigned*)(eval_s + a.ne_lds); // [nw]
    unsigned* eidx_s = bits_s + nw;                    // edge list: [ne_lds]
    const int nwp = (n + 31) >> 5;                     // spike words in column order
    unsigned* bm_s = bits_s + nw;                      // bitmaps: [NB][nwp]
    unsigned* pk_s = NB == 32 ? bits_s : bm_s + NB * nwp;   //     [nwp] spikes in column order
    int* cnt_s = (int*)(bm_s + NB * nwp + (NB == 32 ? 0 : nwp));   // [64]
    int* pb_s = cnt_s + 64;                            //          [64]
    unsigned short* list_s = (unsigned short*)(pb_s + 64);   //    [ne + kBitmapBatch]
    __shared__ int abort_s;
    __shared__ unsigned cons_s;
    // windowed products: chunks produced by each producer wave and tails done
    // (cons_s then counts chunks consumed by the chain, over all steps)
    __shared__ unsigned prod_s[kTpProducers];
    __shared__ unsigned tail_s;
    if (tid == 0) { abort_s = 0; cons_s = 0u; tail_s = 0u; }
    long long st0 = a.st0, t0 = a.t0;
    const uint8_t* S_in = a.S;
    uint8_t* S_out = a.S_out;
    double* rec_V = a.rec_V;
    uint8_t* rec_S = a.rec_S;
    if (a.clk) {
        const long long c0 = a.clk[0], c1 = a.clk[1], c2 = a.clk[2];
        st0 = (long long)(unsigned)__builtin_amdgcn_readfirstlane((int)c0) |
              ((long long)__builtin_amdgcn_readfirstlane((int)(c0 >> 32)) << 32);
        t0 = (long long)(unsigned)__builtin_amdgcn_readfirstlane((int)c1) |
             ((long long)__builtin_amdgcn_readfirstlane((int)(c1 >> 32)) << 32);
        if (__builtin_amdgcn_readfirstlane((int)c2) & 1) {
            S_in = a.S_out;
            S_out = const_cast<uint8_t*>(a.S);
        }
        if (rec_V) rec_V += (size_t)t0 * n;
        if (rec_S) rec_S += (size_t)t0 * n;
    }
    for (int k = tid; k < PS; k += kTpThreads) coef_s[k] = k < L ? a.coef[k] : 0.0;
    if (HL) {   // history rows from the HBM ring (rows past nloc repeat the last neuron)
        for (int e = tid; e < NB * HR; e += kTpThreads) {
            const int rr = e / HR, q = e - rr * HR;
            const int sl = L - 1 - (q < L ? q : (q - L) % L);
            hist_s[e] = a.ring[(size_t)sl * n + i0 + (rr < nloc ? rr : nloc - 1)];
        }
    }
    const int head0 = (int)(st0 % L);
    {
        const int sp = head0 == 0 ? L - 1 : head0 - 1;   // V_{st0-1}
        const int slot = (int)((st0 - 1) & 1);
        for (int e = tid; e < NB; e += kTpThreads)
            stage[slot * NB + e] = e < nloc ? a.ring[(size_t)sp * n + i0 + e] : 0.0;
    }
    const int64_t E0 = a.rp[i0];
    const int ne = (int)(a.rp[i0 + nloc] - E0);
    const bool csr_lds = a.ne_lds > 0;
    const bool bitmap = a.rpm == 1;
    if (bitmap) {
        for (int e = tid; e < NB * nwp; e += kTpThreads) {
            const int rr = e / nwp;
            bm_s[e] = rr < nloc ? a.rbm[(size_t)(i0 + rr) * nwp + (e - rr * nwp)] : 0u;
        }
        for (int e = tid; e < ne + kBitmapBatch; e += kTpThreads) list_s[e] = 0;
        if (csr_lds)
            for (int e = tid; e < ne; e += kTpThreads) eval_s[e] = a.val[E0 + e];
        if (wave == kTpXchWave) {   // first edge of each (part, row) scan range
            const int rr = lane & (NB - 1), w_end = min(nwp, (lane / NB) * a.cw);
            int pb = 0;
            if (rr < nloc) {
                pb = (int)(a.rp[i0 + rr] - E0);
                for (int w = 0; w < w_end; ++w) pb += __builtin_popcount(a.rbm[(size_t)(i0 + rr) * nwp + w]);
            }
            pb_s[lane] = pb;
        }
    } else if (csr_lds) {
        for (int e = tid; e < ne; e += kTpThreads) {
            eidx_s[e] = encode_col(a.col[E0 + e], NB, wpg);
            eval_s[e] = a.val[E0 + e];
        }
    }
    const int r = lane % NB;                           // neuron row of this lane (chain, producers)
    const int ir = i0 + (r < nloc ? r : nloc - 1);     // clamped global index for reads
    const bool own = (wave == 0) && (lane < nloc);
    const int i = i0 + lane;
    double V = own ? a.V[i] : 0.0;
    double tp = own ? a.tprev[i] : 0.0;
    bool spk = own ? (S_in[i] != 0) : false;
    int rs = 0, re = 0;
    constexpr int kMaxDinReg = 4;
    double win[kMaxDinReg];
    const bool win_reg = a.din <= kMaxDinReg;
#pragma unroll
    for (int j = 0; j < kMaxDinReg; ++j) win[j] = 0.0;
    if (wave == kTpXchWave && lane < nloc) {
        rs = (int)(a.rp[i] - E0);
        re = (int)(a.rp[i + 1] - E0);
        if (win_reg)
            for (int j = 0; j < a.din; ++j) win[j] = a.W_in[(size_t)i * a.din + j];
    }
    double p[7];
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        p[j] = a.np[j];
        asm volatile("" : "+v"(p[j]));
    }
    const bool producer = wave != 0 && wave != kTpXchWave;
    const int pw = wave - 1 - (wave > kTpXchWave ? 1 : 0);   // product wave index 0..3
    const int part = pw * (64 / NB) + lane / NB;
    int head = head0;
    int lim = st0 < L ? (int)st0 : L - 1;
    // products of the launch's first step (its first W chunks when windowed), all from the HBM ring
    {
        const int nk0 = lim >= 2 ? lim - 1 : 0;
        const int pre0 = W ? min(W, (nk0 + C - 1) / C) : 0;
        if (W && tid < kTpProducers) prod_s[tid] = (unsigned)pre0;
    }
    if constexpr (!WIN) {   // lags past the step's last are read by the chain as +0.0
        for (int e = tid; e < NB * PSr; e += kTpThreads) P[e] = 0.0;
        __syncthreads();
    }
    if (producer) {
        const int nk = lim >= 2 ? lim - 1 : 0;
        const int klim = W ? min(nk, W * C) : nk;
        for (int k0 = 2 + part * PER; k0 < 2 + klim; k0 += C) {
#pragma unroll
            for (int j = 0; j < PER; j += 2) {
                const int k = k0 + j;
                const int ka = k < L ? k : L - 1, kb = k + 1 < L ? k + 1 : L - 1;
                int sa = head - ka; if (sa < 0) sa += L;
                int sbb = head - kb; if (sbb < 0) sbb += L;
                const double pa = coef_s[k] * a.ring[(size_t)sa * n + ir];
                const double pb = coef_s[k + 1] * a.ring[(size_t)sbb * n + ir];
                *reinterpret_cast<double2*>(P + (size_t)r * PSr + k) = make_double2(k <= lim ? pa : 0.0,
                                                                                    k + 1 <= lim ? pb : 0.0);
            }
        }
    }
    __syncthreads();

    const double c1 = coef_s[1];
    const double2* prow = reinterpret_cast<const double2*>(P + (size_t)r * PSr);   // pair j: lags 2j, 2j+1
    long long st = st0;
    unsigned m = 0;                                    // micro-step index within the launch
    unsigned gb = 0;                                   // chunks of the steps before this one (windowed)
    unsigned seen = 0;                                 // producers: last chain progress read
    // One loop per wave role: with the roles as branches of a single loop, copies
    // of the producers' in-flight load registers sit on every path and make the
    // chain wave wait for its own global stores (the store counter is shared).
    auto steps = [&](auto role) __attribute__((always_inline)) {
        constexpr int R = decltype(role)::value;
        // exchange wave, simple synapses, CSR slice in LDS, rows of at most kRegEdges
        // edges: each lane keeps its row's column codes in registers, so a step's
        // pull is one batch of spike-word and weight reads, then the in-order adds
        unsigned pk_r[kRegEdges];
        double wv_r[kRegEdges];
        // inputs of 64 macro steps in the lanes: lane l of uw_cur holds u[uw_base + l],
        // of uw_nxt u[uw_base + 64 + l], loaded a window ahead (far-memory loads beside
        // the exchange polls take more than a step)
        double uw_cur[kMaxDinReg], uw_nxt[kMaxDinReg];
        long long uw_base = t0;
        auto uw_load = [&](double* w, long long base) {
            const long long tl = base + lane;
#pragma unroll
            for (int j = 0; j < kMaxDinReg; ++j) w[j] = (j < a.din && tl < t0 + a.T) ? a.u[tl * a.din + j] : 0.0;
        };
        if constexpr (R == 5) {
            const int len_r = lane < nloc ? re - rs : 0;
#pragma unroll
            for (int q = 0; q < kRegEdges; ++q) {   // absent edges: word 0, weight +0.0
                pk_r[q] = q < len_r ? eidx_s[rs + q] : 0u;
                wv_r[q] = q < len_r ? eval_s[rs + q] : 0.0;
            }
            if (win_reg) {
                uw_load(uw_cur, t0);
                uw_load(uw_nxt, t0 + 64);
            }
        }
        for (long long t = 0; t < a.T; ++t) {
            const long long tt = t0 + t;
            const int tb = (int)(t & 1);
            for (int mi = 0; mi < a.micro; ++mi, ++st, ++m) {
                const int sb = (int)(st & 1);
                const int nk = lim >= 2 ? lim - 1 : 0;     // lags 2..lim of this step
                const int nch = (nk + C - 1) / C;
                if constexpr (R == 0) {
                    // ---------------- chain: H = c_1 V_{st-1} + sum_k P[k]
                    // V_{st-1}: this lane's own V after the first step of the launch (the tail
                    // stored the same value in the ring and the stage), the staged ring value before
                    const double Vprev = m == 0u ? stage[(int)((st - 1) & 1) * NB + r] : V;
                    double h = 0.0;
                    if (lim >= 1) h = h + c1 * Vprev;
                    double2 bA[C / 2], bB[C / 2];
                    auto add = [&](const double2* buf, int cnt) {
                        if (cnt == C) {
                            add16_f64_ordered(h, buf);
                            add16_f64_ordered(h, buf + 8);
                        } else {
#pragma unroll
                            for (int j = 0; j < C / 2; ++j) {
                                if (2 * j < cnt) add_f64_ordered(h, buf[j].x);
                                if (2 * j + 1 < cnt) add_f64_ordered(h, buf[j].y);
                            }
                        }
                    };
                    if (!WIN && lane < NB) {   // lanes past NB would only repeat the rows of lanes 0..NB-1
                        if (nch > 0) h = chain_chunks(h, lds_offset(prow + 1), nch, lds_offset(&cons_s), m * 64u + 1u);
                    } else if (WIN && lane < NB) {
                        // windowed: chunk c of this step is chunk gb+c of the stream, in slot
                        // (gb+c) mod W; it is read once every producer wave has written it, and
                        // its slot is released (cons_s) after its adds
                        bool stalled = false;
                        auto wait_prod = [&](unsigned need) {
                            long long spins = 0;
                            while (!stalled) {
                                asm volatile("" ::: "memory");
                                bool ok = true;
#pragma unroll
                                for (int q = 0; q < kTpProducers; ++q)
                                    ok = ok && (int)(lds_counter_load(&prod_s[q]) - need) >= 0;
                                if (ok) break;
                                if (++spins > kTpSpinLimit) {
                                    stalled = true;
                                    __hip_atomic_store(&abort_s, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                                    __hip_atomic_store(a.abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                                }
                                __builtin_amdgcn_s_sleep(1);
                            }
                            asm volatile("" ::: "memory");
                        };
                        auto ldw = [&](double2* buf, int c) {
                            const int sl = (int)((gb + (unsigned)c) % (unsigned)W);
#pragma unroll
                            for (int j = 0; j < C / 2; ++j) buf[j] = prow[1 + (C / 2) * sl + j];
                        };
                        auto release = [&](unsigned v) {
                            asm volatile("" ::: "memory");
                            lds_counter_store(&cons_s, v);
                        };
                        if (nch > 0) {
                            wait_prod(gb + 1u);
                            ldw(bA, 0);
                        }
                        int c = 0;
                        for (; c + 2 <= nch; c += 2) {
                            wait_prod(gb + (unsigned)c + 2u);
                            ldw(bB, c + 1);
                            add(bA, C);
                            release(gb + (unsigned)c + 1u);
                            if (c + 2 < nch) {
                                wait_prod(gb + (unsigned)c + 3u);
                                ldw(bA, c + 2);
                            }
                            add(bB, (c + 1 == nch - 1) ? nk - (c + 1) * C : C);
                            release(gb + (unsigned)c + 2u);
                        }
                        if (c < nch) {
                            add(bA, nk - c * C);
                            release(gb + (unsigned)c + 1u);
                        }
                    }
                    double leak = -(Vprev - p[2]) / p[3];   // before the barrier: the wait hides it
                    asm volatile("" : "+v"(leak));          // (kept here, not sunk past the barrier)
                    __syncthreads();   // rec / ext of this step and the products of the next are ready
                    // the input terms are read together with the abort flag (one LDS round
                    // trip: the empty statement keeps the compiler from sinking the two reads
                    // past the abort branch, which would cost a second one)
                    const int ll = lane < NB ? lane : 0;
                    double xe = ext_s[tb * NB + ll], xr = rec_s[sb * NB + ll];
                    const int ab = abort_s;
                    asm volatile("" : "+v"(xe), "+v"(xr));
                    if (__builtin_amdgcn_readfirstlane(ab)) break;
                    if (own) {                                  // flif_gl.c:50-100
                        const double x = xe + xr;
                        spk = false;
                        if (tp > p[6]) {
                            const double rhs = leak + x + p[5];
                            V = p[4] * rhs - h;
                            if (V >= p[0]) { V = p[1]; spk = true; tp = 0.0; }
                        } else {
                            V = p[1];
                        }
                    }
                    publish_spikes(a.xch, nw, g, wpg, st, lane, __ballot(own && spk));
                    // LDS first (stage, history row, the producers' signal), then the HBM
                    // stores: the next chain's first LDS wait then finds the writes done.
                    // Lane 0 (always an owner) signals after the writes of every owner lane:
                    // a wave's LDS instructions complete in issue order, and the empty
                    // statement keeps the compiler from moving the signal above them.
                    if (own) {
                        stage[sb * NB + lane] = V;
                        if (HL) {
                            const int e = L - 1 - head;
                            hist_s[lane * HR + e] = V;
                            for (int q = e; q < HR - L; q += L) hist_s[lane * HR + L + q] = V;
                        }
                        asm volatile("" ::: "memory");   // V_st staged before the signal
                        if (lane == 0) {
                            if constexpr (!WIN) lds_counter_store(&cons_s, (m + 1u) * 64u);   // tail done
                            else lds_counter_store(&tail_s, m + 1u);
                        }
                    }
                    asm volatile("" ::: "memory");
                    if (own) {
                        a.ring[(size_t)head * n + i] = V;
                        tp = tp + a.dt;
                        if (mi == a.micro - 1) {
                            if (rec_V) rec_V[(size_t)t * n + i] = V;
                            if (rec_S) rec_S[(size_t)t * n + i] = spk ? 1 : 0;
                        }
                    }
                } else if constexpr (R == 4) {
                    // ---------------- exchange, row pull, external input
                    double uin[kMaxDinReg];
                    if (mi == 0 && win_reg) {
#pragma unroll
                        for (int j = 0; j < kMaxDinReg; ++j) uin[j] = (j < a.din) ? a.u[tt * a.din + j] : 0.0;
                    }
                    const bool aborted = !gather_spikes(a.xch, nw, NB, wpg, n, st, st0, S_in, bits_s,
                                                        a.abort_flag, &abort_s, lane);
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    if (bitmap && !aborted) {
                        if (NB == 16) {   // two 16-neuron exchange words per column-order word
                            for (int w = lane; w < nwp; w += 64)
                                pk_s[w] = bits_s[2 * w] | (2 * w + 1 < nw ? bits_s[2 * w + 1] << 16 : 0u);
                            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                            __builtin_amdgcn_wave_barrier();
                            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                        }
                        bitmap_scan(bm_s, pk_s, list_s, cnt_s, pb_s, nwp, a.cw, NB, (lane & (NB - 1)) < nloc, lane);
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                        __builtin_amdgcn_wave_barrier();
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    }
                    if (!aborted && lane < nloc) {
                        double rec;
                        if (bitmap)
                            rec = csr_lds ? bitmap_sum<kBitmapBatch / 2>(list_s, cnt_s, pb_s, eval_s, NB, lane)
                                          : bitmap_sum<kBitmapBatch / 2>(list_s, cnt_s, pb_s, a.val + E0, NB, lane);
                        else
                            rec = row_pull(a.stype, csr_lds, rs, re, E0, eidx_s, eval_s, a.col, a.val, bits_s, NB, wpg,
                                           a.edge_decay, a.edge_tr);
                        rec_s[sb * NB + lane] = rec;
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
                    __syncthreads();
                    if (__builtin_amdgcn_readfirstlane(abort_s)) break;
                } else if constexpr (R == 5) {
                    // ---------------- exchange with the row pull from registers (simple
                    // synapses, CSR slice in LDS, rows of at most kRegEdges edges): one batch
                    // of spike-word and weight reads, then the in-order adds over all
                    // kRegEdges slots (sparse.c:68-76; a silent or absent edge adds +0.0, which
                    // leaves the sum unchanged: the sum starts at +0.0 and is never -0.0).
                    const bool aborted = !gather_spikes(a.xch, nw, NB, wpg, n, st, st0, S_in, bits_s,
                                                        a.abort_flag, &abort_s, lane);
                    __builtin_amdgcn_wave_barrier();   // bits_s written by this wave, read by its lanes
                    if (!aborted) {
                        unsigned bw[kRegEdges];
#pragma unroll
                        for (int q = 0; q < kRegEdges; ++q) bw[q] = bits_s[pk_r[q] >> 5];
                        double sum = 0.0;
#pragma unroll
                        for (int q = 0; q < kRegEdges; ++q)
                            sum = sum + (((bw[q] >> (pk_r[q] & 31)) & 1u) ? wv_r[q] : 0.0);
                        if (mi == 0 && win_reg && tt - uw_base == 64) {   // next input window
#pragma unroll
                            for (int j = 0; j < kMaxDinReg; ++j) uw_cur[j] = uw_nxt[j];
                            uw_base += 64;
                            uw_load(uw_nxt, uw_base + 64);
                        }
                        if (lane < nloc) {
                            rec_s[sb * NB + lane] = sum;
                            if (mi == 0) {
                                double su = 0.0;
                                if (win_reg) {
                                    const int k = (int)(tt - uw_base);   // uniform, 0..63
#pragma unroll
                                    for (int j = 0; j < kMaxDinReg; ++j)
                                        if (j < a.din) {
                                            const unsigned long long b = __double_as_longlong(uw_cur[j]);
                                            const unsigned lo = __builtin_amdgcn_readlane((unsigned)b, k);
                                            const unsigned hi = __builtin_amdgcn_readlane((unsigned)(b >> 32), k);
                                            su = su + win[j] * __longlong_as_double(((long long)hi << 32) | lo);
                                        }
                                } else {
                                    for (int j = 0; j < a.din; ++j)
                                        su = su + a.W_in[(size_t)i * a.din + j] * a.u[tt * a.din + j];
                                }
                                ext_s[tb * NB + lane] = a.gain * su;
                            }
                        }
                    }
                    __syncthreads();
                    if (__builtin_amdgcn_readfirstlane(abort_s)) break;
                } else if constexpr (!WIN) {
                    // ---------------- products of step st+1, behind the chain of step st
                    const int headn = head + 1 == L ? 0 : head + 1;
                    const int limn = lim < L - 1 ? lim + 1 : L - 1;
                    const int nkn = limn >= 2 ? limn - 1 : 0;
                    const int nchn = (nkn + C - 1) / C;
                    // V_{st+1-k}, k >= 3, is final since the last barrier: loads run D - 1
                    // chunks ahead of the stores.  Lag k of step st+1 sits in slot
                    // (headn - k) mod L; the lane's two lags of chunk c are
                    // k = 2 + c*C + part*PER + j.  Each lane keeps the byte offset of its next
                    // slot per j (32-bit: N*L*8 < 2^31 for this kernel) and steps it back C
                    // slots per chunk, in chunk order; chunks past the first exist only when
                    // L > C, so one wrap suffices.  Lag 2 (replaced by the staged value) and
                    // lags past lim are read but discarded.
                    const int rowb = n * 8;
                    unsigned offb[PER];
#pragma unroll
                    for (int j = 0; j < PER; ++j) {
                        int sl = (headn - (2 + part * PER + j)) % L;
                        sl += sl < 0 ? L : 0;
                        offb[j] = (unsigned)(sl * rowb + ir * 8);
                    }
                    const int stepb = C * rowb, wrapb = L * rowb;
                    const char* ring_b = reinterpret_cast<const char*>(a.ring);
                    // history rows in LDS: lag k0 = 2 + c*C + part*PER of chunk c is entry
                    // (L - 1 - headn + k0) mod L of row r, its next PER - 1 lags the entries after it
                    int hA = (L - 1 - headn + 2 + part * PER) % L;
                    const double* hrow = hist_s + (size_t)r * HR;
                    auto ldr = [&](double* v, int) {   // called for chunks 0, 1, 2, ... in order
                        if (HL) {
#pragma unroll
                            for (int j = 0; j < PER; ++j) v[j] = hrow[hA + j];
                            hA += C;
                            hA -= hA >= L ? L : 0;
                            return;
                        }
#pragma unroll
                        for (int j = 0; j < PER; ++j) {
                            v[j] = *reinterpret_cast<const double*>(ring_b + offb[j]);
                            int o = (int)offb[j] - stepb;
                            o += o < 0 ? wrapb : 0;
                            offb[j] = (unsigned)o;
                        }
                    };
                    bool stalled = false;
                    // chunk q is loaded into vb[q % D], D - 1 chunks ahead of its products
                    constexpr int D = kTpLoadDepth;
                    constexpr int U = kTpProdUnit;
                    static_assert(D % U == 0, "units of chunks tile the load ring");
                    double vb[D][PER];
                    // Units of U chunks (cnt <= U in the last one): the coefficient pairs of the
                    // unit and, when the last value read is not enough, the chain's progress are
                    // read together, so a unit costs one LDS round trip (a wave's LDS reads
                    // complete in order, so any wait also waits for the loads in flight).  The
                    // unit's products are written once the chain of step st is past its last
                    // chunk (or does not read it) and the tail of step st-1 has staged V_{st-1};
                    // then the history of the chunks D - 1 ahead is loaded into the freed slots.
                    auto unit = [&](int c0, int s0, int cnt) {
                        const int cl = c0 + cnt - 1;
                        const unsigned target = m * 64u + (unsigned)(cl + 1 < nch ? cl + 1 : nch);
                        const bool need = (int)(seen - target) < 0;
                        unsigned fresh = seen;
                        if (need) fresh = lds_counter_load(&cons_s);
                        double2 pr[U][PER / 2];
#pragma unroll
                        for (int u = 0; u < U; ++u)
                            if (u < cnt) {
#pragma unroll
                                for (int j = 0; j < PER / 2; ++j) {
                                    const int k = 2 + (c0 + u) * C + part * PER + 2 * j;
                                    pr[u][j] = make_double2(coef_s[k] * vb[s0 + u][2 * j],
                                                            coef_s[k + 1] * vb[s0 + u][2 * j + 1]);
                                }
                            }
                        if (need) {
                            seen = fresh;
                            long long spins = 0;
                            while ((int)(seen - target) < 0 && !stalled) {
                                if (++spins > kTpSpinLimit) {
                                    stalled = true;
                                    __hip_atomic_store(&abort_s, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                                    __hip_atomic_store(a.abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                                }
                                __builtin_amdgcn_s_sleep(1);
                                seen = lds_counter_load(&cons_s);
                            }
                        }
                        if (c0 == 0 && part == 0)
                            pr[0][0].x = coef_s[2] * stage[(int)((st - 1) & 1) * NB + r];   // lag 2: V_{st-1}
#pragma unroll
                        for (int u = 0; u < U; ++u)
                            if (u < cnt) {
                                double2* dst = reinterpret_cast<double2*>(P + (size_t)r * PS + 2 + (c0 + u) * C + part * PER);
#pragma unroll
                                for (int j = 0; j < PER / 2; ++j) dst[j] = pr[u][j];
                            }
#pragma unroll
                        for (int u = 0; u < U; ++u)
                            if (u < cnt && c0 + u + D - 1 < nchn) ldr(vb[(s0 + u + D - 1) % D], c0 + u + D - 1);
                    };
#pragma unroll
                    for (int j = 0; j < D - 1; ++j)
                        if (j < nchn) ldr(vb[j], j);
                    int c = 0;
                    for (; c + D <= nchn; c += D) {
#pragma unroll
                        for (int j = 0; j < D; j += U) unit(c + j, j, U);
                    }
#pragma unroll
                    for (int j = 0; j < D; j += U)
                        if (c + j < nchn) unit(c + j, j, min(U, nchn - c - j));
                    if (nchn > 0) {   // the chain reads the lags past limn (last chunk only) as +0.0
#pragma unroll
                        for (int j = 0; j < PER; ++j) {
                            const int k = 2 + (nchn - 1) * C + part * PER + j;
                            if (k > limn) P[(size_t)r * PS + k] = 0.0;
                        }
                    }
                    __syncthreads();
                    if (__builtin_amdgcn_readfirstlane(abort_s)) break;
                } else {
                    // ---------------- windowed products: the chunks of step st past the first W
                    // (into the slots the chain releases), then the first W of step st+1.  Chunk g
                    // of the stream goes to slot g mod W once the chain has released chunk g-W.
                    const int headn = head + 1 == L ? 0 : head + 1;
                    const int limn = lim < L - 1 ? lim + 1 : L - 1;
                    const int nkn = limn >= 2 ? limn - 1 : 0;
                    const int nchn = (nkn + C - 1) / C;
                    const int pre = min(W, nch), qa = nch - pre;
                    const int Q = qa + min(W, nchn);
                    // item q: chunk pre+q of step st (q < qa) or chunk q-qa of step st+1.  The
                    // ring holds V_{st-k} at (head - k) mod L for step st and at (headn - k)
                    // for step st+1; every value read is final since the last barrier except
                    // lag 2 of step st+1 (V_{st-1}), which comes from the stage.
                    auto ldq = [&](double* v, int q) {
                        const bool nx = q >= qa;
                        const int cc = nx ? q - qa : pre + q;
                        const int sb0 = (nx ? headn : head) - (2 + cc * C + part * PER);
#pragma unroll
                        for (int j = 0; j < PER; ++j) {
                            int sl = sb0 - j;
                            sl += sl < 0 ? L : 0;
                            sl += sl < 0 ? L : 0;
                            v[j] = a.ring[(size_t)sl * n + ir];
                        }
                    };
                    bool stalled = false;
                    auto spin_until = [&](const unsigned* ctr, unsigned need) {
                        long long spins = 0;
                        while (!stalled) {
                            asm volatile("" ::: "memory");
                            if ((int)(lds_counter_load(ctr) - need) >= 0) break;
                            if (++spins > kTpSpinLimit) {
                                stalled = true;
                                __hip_atomic_store(&abort_s, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                                __hip_atomic_store(a.abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            }
                            __builtin_amdgcn_s_sleep(1);
                        }
                        asm volatile("" ::: "memory");
                    };
                    auto putq = [&](const double* v, int q) {
                        const bool nx = q >= qa;
                        const int cc = nx ? q - qa : pre + q;
                        const unsigned gq = gb + (unsigned)(nx ? nch + cc : cc);
                        double2 pr[PER / 2];
#pragma unroll
                        for (int j = 0; j < PER / 2; ++j) {
                            const int k = 2 + cc * C + part * PER + 2 * j;
                            pr[j] = make_double2(coef_s[k] * v[2 * j], coef_s[k + 1] * v[2 * j + 1]);
                        }
                        spin_until(&cons_s, gq + 1u - (unsigned)W);   // slot released by chunk gq-W
                        if (nx && cc == 0 && part == 0) {
                            spin_until(&tail_s, m);                    // V_{st-1} staged
                            pr[0].x = coef_s[2] * stage[(int)((st - 1) & 1) * NB + r];
                        }
                        double2* dst = reinterpret_cast<double2*>(P + (size_t)r * PSr + 2 +
                                                                  (size_t)(gq % (unsigned)W) * C + part * PER);
#pragma unroll
                        for (int j = 0; j < PER / 2; ++j) dst[j] = pr[j];
                        asm volatile("" ::: "memory");
                        lds_counter_store(&prod_s[pw], gq + 1u);
                    };
                    double vA[PER], vB[PER], vC[PER], vD[PER];
                    if (Q > 0) ldq(vA, 0);
                    if (Q > 1) ldq(vB, 1);
                    if (Q > 2) ldq(vC, 2);
                    int q = 0;
                    for (; q + 4 <= Q; q += 4) {
                        ldq(vD, q + 3);
                        putq(vA, q);
                        if (q + 4 < Q) ldq(vA, q + 4);
                        putq(vB, q + 1);
                        if (q + 5 < Q) ldq(vB, q + 5);
                        putq(vC, q + 2);
                        if (q + 6 < Q) ldq(vC, q + 6);
                        putq(vD, q + 3);
                    }
                    if (q < Q) putq(vA, q);
                    if (q + 1 < Q) putq(vB, q + 1);
                    if (q + 2 < Q) putq(vC, q + 2);
                    __syncthreads();
                    if (__builtin_amdgcn_readfirstlane(abort_s)) break;
                }
                head = head + 1 == L ? 0 : head + 1;
                if (lim < L - 1) ++lim;
                gb += (unsigned)nch;
            }
            if (__builtin_amdgcn_readfirstlane(abort_s)) break;
        }
    };
    if (wave == 0) steps(RoleTag<0>{});
    else if (wave == kTpXchWave) {
        const int len = lane < nloc ? re - rs : 0;
        if (NB == 16 && !bitmap && csr_lds && a.stype == ST_SIMPLE && __all(len <= kRegEdges))
            steps(RoleTag<5>{});
        else
            steps(RoleTag<4>{});
    }
    else steps(RoleTag<1>{});
    if (own) {
        a.V[i] = V;
        S_out[i] = spk ? 1 : 0;
        a.tprev[i] = tp;
    }
}

template <int NB>
__global__ void __launch_bounds__(kTpThreads) k_persist_tp(PersistArgs a) {
    tp_body<NB, false>(a, blockIdx.x);
}

// The same with a window of product chunks (PersistArgs::win), for rows too
// long to keep whole in LDS; a separate kernel so that the whole-row kernel
// keeps its registers.
template <int NB>
__global__ void __launch_bounds__(kTpThreads) k_persist_tpw(PersistArgs a) {
    tp_body<NB, true>(a, blockIdx.x);
}

// Several reservoirs in one launch: workgroup b runs local workgroup
// wg_map[b].y of reservoir wg_map[b].x with that reservoir's arguments.  The
// reservoirs share nothing but the launch (each has its own exchange words).
template <int NB>
__global__ void __launch_bounds__(kTpThreads) k_persist_tp_batch(const PersistArgs* __restrict__ table,
                                                                 const int2* __restrict__ wg_map) {
    const int2 mw = wg_map[blockIdx.x];
    tp_body<NB, false>(table[mw.x], mw.y);
}

namespace {

struct Plan {
    int nb = 0, G = 0, ring_lds = 0, ne_lds = 0;
    int tp = 0;                 // FLIF-GL with the trailing product stage (k_persist_tp)
    int rpm = 0, cw = 0;        // row pull over spike-word bitmaps (k_persist_tp only)
    int win = 0;                // k_persist_tp product window in chunks (0: whole rows)
    int hl = 0;                 // k_persist_tp: history rows in LDS (products read no HBM)
    size_t lds = 0;
};

// batch: the plan for a batched launch, which prefers 32 neurons per
// workgroup (more reservoirs per launch); a reservoir launched alone prefers
// 16 when the device has enough CUs (less work per workgroup, a shorter step).
bool persist_plan(spires_dev* g, int ncu, Plan* pl, bool batch = false) {
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
        const size_t LP = (size_t)L + (L & 1) + 2;
        size_t base = (rl ? LP * nb * 8 : 0) + (gl ? (size_t)(L + 16) * 8 : 0) + 4 * (size_t)nb * 8 +
                      (size_t)G * wpg * 4 + 64;
        if (base > cap) return false;
        const int64_t me = max_edges(nb);
        pl->ne_lds = (base + (size_t)me * 12 <= cap) ? (int)std::max<int64_t>(me, 1) : 0;
        pl->nb = nb; pl->G = G; pl->ring_lds = rl; pl->tp = 0; pl->rpm = 0; pl->cw = 0; pl->win = 0;
        pl->lds = base + (size_t)pl->ne_lds * 12;
        return true;
    };
    // Product-wave kernel.  Row pull: spike-word bitmaps when a workgroup's rows
    // hold more edges than half the words a bitmap scan reads, else the edge
    // list; weights in LDS when they fit ("fast"), else read from HBM.
    // Whole product rows when they fit in LDS; otherwise a window of W 32-lag
    // chunks per row (the largest of 16, 8, 4 that fits), refilled by the
    // product waves while the chain runs.
    // whole-row product plans keep the workgroup's history rows in LDS when they fit
    auto add_history = [&](int nb) {
        const size_t hb = (size_t)nb * tp_hist_stride(L) * 8;
        if (pl->win == 0 && pl->lds + hb <= cap) { pl->hl = 1; pl->lds += hb; }
    };
    auto fill_tp = [&](int nb, bool fast_only) -> bool {
        const int G = (n + nb - 1) / nb;
        const int wpg = (nb + 31) / 32;
        if (L < 2 || G > ncu || G * wpg > 512) return false;
        const size_t PS = (size_t)tp_row_stride(L);
        const size_t rest = PS * 8 + 6 * (size_t)nb * 8 + (size_t)G * wpg * 4 + 64;
        size_t base = PS * nb * 8 + rest;
        int win = 0;
        if (base > cap) {
            if (batch) return false;   // batched launches take whole rows only
            const int nch_max = (L - 2 + kTpChunk - 1) / kTpChunk;
            for (int w : {16, 8, 4}) {
                const size_t b = (size_t)(32 * w + 2) * nb * 8 + rest;
                if (w < nch_max && b <= cap) { win = w; base = b; break; }
            }
            if (!win) return false;
        } else if (nb == 64) {
            return false;   // whole rows at 64 neurons per workgroup: not instantiated (32 fits first)
        } else if ((long long)n * L * 8 >= (1ll << 31)) {
            return false;   // whole-row product waves address the history ring with 32-bit offsets
        }
        const int64_t me = max_edges(nb);
        const int nwp = (n + 31) / 32;
        pl->nb = nb; pl->G = G; pl->ring_lds = 0; pl->tp = 1; pl->rpm = 0; pl->cw = 0; pl->win = win; pl->hl = 0;
        const bool dense = nb <= 32 && g->stype == ST_SIMPLE && me + kBitmapBatch < 65536 &&
                           2 * me > (int64_t)nb * nwp;
        if (dense) {
            const size_t bm = ((size_t)(nb + (nb == 32 ? 0 : 1)) * nwp * 4 + 512 + ((size_t)me + kBitmapBatch) * 2 + 7) /
                              8 * 8;
            if (base + bm <= cap) {
                const bool wl = base + bm + (size_t)me * 8 <= cap;
                if (fast_only && !wl) return false;
                const int parts = 64 / nb;
                pl->rpm = 1;
                pl->cw = ((nwp + parts - 1) / parts + 7) / 8 * 8;
                pl->ne_lds = wl ? (int)std::max<int64_t>(me, 1) : 0;
                pl->lds = base + bm + (size_t)pl->ne_lds * 8;
                add_history(nb);
                return true;
            }
        }
        const bool el = base + (size_t)me * 12 <= cap;
        if (fast_only && !el) return false;
        pl->ne_lds = el ? (int)std::max<int64_t>(me, 1) : 0;
        pl->lds = base + (size_t)pl->ne_lds * 12;
        add_history(nb);
        return true;
    };
    if (gl) {
        // path 2 / auto: trailing-product kernel (16 neurons per workgroup when only
        // that keeps the row pull in LDS), then the single-wave chain with the
        // history in LDS, then in HBM; path 3 / 4 force the single-wave HBM / LDS variants.
        if (g->path == 0 || g->path == 2) {
            // 64 neurons per workgroup only for a reservoir launched alone with a window
            const int order[3] = {batch ? 32 : 16, batch ? 16 : 32, batch ? 0 : 64};
            for (int nb : order)
                if (nb && fill_tp(nb, true)) return true;
            for (int nb : order)
                if (nb && fill_tp(nb, false)) return true;
        }
        if (g->path != 3)
            for (int nb : {32, 16})
                if (fill(nb, 1)) return true;
        if (g->path == 4) return false;
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

namespace {

// Exchange words (cleared when the workgroup count changes) and the abort flag
// of one reservoir's launch.
int prepare_launch(spires_dev* g, const Plan& pl) {
    const int wpg = (pl.nb + 31) / 32;
    const size_t nw = (size_t)pl.G * wpg;
    if (g->xch_words < 2 * nw * kXchStride) {
        dfree(g->xch);
        int rc = dalloc(&g->xch, 2 * nw * kXchStride);
        if (rc) return rc;
        g->xch_words = 2 * nw * kXchStride;
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
    if (pl.rpm == 1 && !g->rbm) {
        // row i: bit j % 32 of word j / 32 for each column j
        const size_t nwp = ((size_t)g->n + 31) / 32;
        std::vector<unsigned> bm((size_t)g->n * nwp, 0u);
        for (int i = 0; i < g->n; ++i)
            for (int64_t k = g->rp_host[i]; k < g->rp_host[i + 1]; ++k) {
                const int j = g->col_host[k];
                bm[(size_t)i * nwp + (j >> 5)] |= 1u << (j & 31);
            }
        int rc = dalloc(&g->rbm, bm.size());
        if (rc) return rc;
        SDEV_CK(hipMemcpyAsync(g->rbm, bm.data(), bm.size() * 4, hipMemcpyHostToDevice, g->stream));
        SDEV_CK(hipStreamSynchronize(g->stream));
    }
    if (g->abort_dirty) {
        SDEV_CK(hipMemsetAsync(g->abort_flag, 0, 4, g->stream));
        g->abort_dirty = false;
    }
    return 0;
}

PersistArgs make_args(spires_dev* g, const Plan& pl, const double* u_series, long long t0, long long T,
                      double* rec_V, uint8_t* rec_S) {
    PersistArgs a;
    memset(&a, 0, sizeof(a));
    a.n = g->n; a.din = g->din; a.L = g->L; a.G = pl.G; a.wpg = (pl.nb + 31) / 32; a.micro = g->micro;
    a.stype = g->stype; a.ne_lds = pl.ne_lds;
    a.rpm = pl.rpm; a.cw = pl.cw; a.rbm = g->rbm; a.win = pl.win; a.hl = pl.hl;
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
    return a;
}

// After a launch: advance the step counters and report an aborted exchange.
// The device flag stays raised until the error has been reported, so a launch
// queued behind an aborted one gives up at its first wait instead of hiding it.
int finish_launch(spires_dev* g, long long T) {
    g->step += (long long)T * g->micro;
    g->msteps += (size_t)T;
    g->cur ^= 1;
    unsigned flag = 0;
    SDEV_CK(hipMemcpyAsync(&flag, g->abort_flag, 4, hipMemcpyDeviceToHost, g->stream));
    SDEV_CK(hipStreamSynchronize(g->stream));
    if (flag) {
        g->abort_dirty = true;
        set_error("persistent step kernel timed out waiting for the spike exchange");
        return SPIRES_DEV_EINTERNAL;
    }
    return 0;
}

// As finish_launch without waiting: the flag is copied to pinned host memory
// in stream order and checked by the next settle(), so the host can queue the
// next launch while this one runs.
int finish_launch_async(spires_dev* g, long long T) {
    g->step += (long long)T * g->micro;
    g->msteps += (size_t)T;
    g->cur ^= 1;
    if (!g->abort_host) {
        SDEV_CK(hipHostMalloc((void**)&g->abort_host, sizeof(unsigned), hipHostMallocDefault));
        *g->abort_host = 0u;
    }
    SDEV_CK(hipMemcpyAsync(g->abort_host, g->abort_flag, 4, hipMemcpyDeviceToHost, g->stream));
    g->abort_pending = true;
    return 0;
}

}  // namespace

int launch_persist(spires_dev* g, const double* u_series, long long t0, long long T, double* rec_V,
                   uint8_t* rec_S) {
    Plan pl;
    if (!persist_plan(g, num_cus(), &pl)) { set_error("persistent path unsupported"); return SPIRES_DEV_EINVAL; }
    int rc = prepare_launch(g, pl);
    if (rc) return rc;
    PersistArgs a = make_args(g, pl, u_series, t0, T, rec_V, rec_S);
    void* args[] = {&a};
    const void* fn = nullptr;
    int threads = kPersistThreads;
    if (pl.tp) {
        if (pl.win)
            fn = pl.nb == 64 ? (const void*)k_persist_tpw<64>
                 : pl.nb == 32 ? (const void*)k_persist_tpw<32> : (const void*)k_persist_tpw<16>;
        else
            fn = pl.nb == 32 ? (const void*)k_persist_tp<32> : (const void*)k_persist_tp<16>;
        threads = kTpThreads;
    } else {
        switch (g->ntype) {
        case NT_LIF_DISCRETE: fn = pick_kernel<NT_LIF_DISCRETE>(pl.nb, pl.ring_lds); break;
        case NT_LIF_BIO: fn = pick_kernel<NT_LIF_BIO>(pl.nb, pl.ring_lds); break;
        default: fn = pick_kernel<NT_FLIF_GL>(pl.nb, pl.ring_lds); break;
        }
    }
    SDEV_CK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pl.lds));
    SDEV_CK(hipLaunchCooperativeKernel(fn, dim3(pl.G), dim3(threads), args, (unsigned)pl.lds, g->stream));
    return finish_launch_async(g, T);
}

int settle_ready(spires_dev* g) {
    if (!g->abort_pending || !g->abort_host) return 0;
    if (__atomic_load_n(g->abort_host, __ATOMIC_ACQUIRE) == 0u) return 0;
    return settle(g);
}

int settle(spires_dev* g) {
    if (!g || !g->abort_pending) return 0;
    SDEV_CK(hipStreamSynchronize(g->stream));
    g->abort_pending = false;
    if (*g->abort_host) {
        *g->abort_host = 0u;
        g->abort_dirty = true;
        set_error("persistent step kernel timed out waiting for the spike exchange");
        return SPIRES_DEV_EINTERNAL;
    }
    return 0;
}

int batch_plan(spires_dev* const* gs, int count, int* nb_out, int* wgs_out) {
    const int ncu = num_cus();
    int nb = 0, total = 0;
    for (int r = 0; r < count; ++r) {
        Plan pl;
        if (!persist_plan(gs[r], ncu, &pl, true) || !pl.tp) {
            set_error("batched runs need FLIF-GL reservoirs that fit the product-wave kernel");
            return SPIRES_DEV_EINVAL;
        }
        if (r == 0) nb = pl.nb;
        if (pl.nb != nb) { set_error("batched reservoirs need the same neurons-per-workgroup (history length class)"); return SPIRES_DEV_EINVAL; }
        total += pl.G;
    }
    if (total > ncu) { set_error("batched reservoirs need more workgroups than the device has compute units"); return SPIRES_DEV_EINVAL; }
    if (nb_out) *nb_out = nb;
    if (wgs_out) *wgs_out = total;
    return 0;
}

int launch_persist_batch(spires_dev* const* gs, int count, long long t0, long long T, double* const* rec_V,
                         uint8_t* const* rec_S) {
    int nb = 0, total = 0;
    int rc = batch_plan(gs, count, &nb, &total);
    if (rc) return rc;
    const int ncu = num_cus();
    std::vector<PersistArgs> table(count);
    std::vector<int2> map;
    map.reserve(total);
    size_t lds = 0;
    for (int r = 0; r < count; ++r) {
        spires_dev* g = gs[r];
        Plan pl;
        persist_plan(g, ncu, &pl, true);
        if ((rc = prepare_launch(g, pl))) return rc;
        SDEV_CK(hipStreamSynchronize(g->stream));
        table[r] = make_args(g, pl, g->din > 0 ? g->u_series : nullptr, t0, T, rec_V[r], rec_S[r]);
        for (int b = 0; b < pl.G; ++b) map.push_back(make_int2(r, b));
        lds = std::max(lds, pl.lds);
    }
    spires_dev* g0 = gs[0];
    PersistArgs* table_d = nullptr;
    int2* map_d = nullptr;
    SDEV_CK(hipMalloc(&table_d, table.size() * sizeof(PersistArgs)));
    if (hipMalloc(&map_d, map.size() * sizeof(int2)) != hipSuccess) {
        (void)hipFree(table_d);
        set_error("device allocation failed");
        return SPIRES_DEV_EALLOC;
    }
    hipError_t e = hipMemcpyAsync(table_d, table.data(), table.size() * sizeof(PersistArgs), hipMemcpyHostToDevice,
                                  g0->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(map_d, map.data(), map.size() * sizeof(int2), hipMemcpyHostToDevice, g0->stream);
    const void* fn = nb == 32 ? (const void*)k_persist_tp_batch<32> : (const void*)k_persist_tp_batch<16>;
    if (e == hipSuccess) e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    void* args[] = {&table_d, &map_d};
    if (e == hipSuccess)
        e = hipLaunchCooperativeKernel(fn, dim3(total), dim3(kTpThreads), args, (unsigned)lds, g0->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g0->stream);
    (void)hipFree(table_d);
    (void)hipFree(map_d);
    if (e != hipSuccess) { set_error(hipGetErrorString(e)); return SPIRES_DEV_EINTERNAL; }
    for (int r = 0; r < count; ++r)
        if ((rc = finish_launch(gs[r], T))) return rc;
    return 0;
}

}  // namespace sdev

namespace sdev {
// The spike exchange alone: G workgroups publish and gather their words every
// step with the product-wave kernel's own protocol (publish_spikes /
// gather_spikes), and do nothing else.  Its time per step is the floor that
// exchange puts under a persistent step of G workgroups.
__global__ void __launch_bounds__(64) k_exchange_floor(unsigned long long* xch, int G, long long T,
                                                       const uint8_t* S, unsigned* abort_flag) {
    __shared__ unsigned bits_s[kMaxWordsPerLane * 64];
    __shared__ int abort_s;
    if (threadIdx.x == 0) abort_s = 0;
    __syncthreads();
    const int lane = threadIdx.x, g = blockIdx.x;
    for (long long st = 0; st < T; ++st) {
        if (!gather_spikes(xch, G, 32, 1, G * 32, st, 0, S, bits_s, abort_flag, &abort_s, lane)) break;
        publish_spikes(xch, G, g, 1, st, lane, 0ull);
    }
}
}  // namespace sdev

extern "C" int spires_dev_exchange_floor(int G, long long T, double* us_per_step) {
    using namespace sdev;
    sdev::RandShield rand_shield_;
    if (G < 1 || G > 64 * kMaxWordsPerLane || T < 2 || !us_per_step) return SPIRES_DEV_EINVAL;
    if (G > num_cus()) { set_error("exchange floor: more workgroups than compute units"); return SPIRES_DEV_EINVAL; }
    unsigned long long* xch = nullptr;
    uint8_t* S = nullptr;
    unsigned* ab = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto done = [&](int rc) {
        if (st) (void)hipStreamSynchronize(st);
        dfree(xch); dfree(S); dfree(ab);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (st) (void)hipStreamDestroy(st);
        return rc;
    };
    int rc = 0;
    if ((rc = dalloc(&xch, 2 * (size_t)G * kXchStride)) || (rc = dalloc(&S, (size_t)G * 32)) || (rc = dalloc(&ab, 1)))
        return done(rc);
    hipError_t e = hipStreamCreate(&st);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipMemsetAsync(xch, 0, 2 * (size_t)G * kXchStride * 8, st);
    if (e == hipSuccess) e = hipMemsetAsync(S, 0, (size_t)G * 32, st);
    if (e == hipSuccess) e = hipMemsetAsync(ab, 0, 4, st);
    void* args[] = {&xch, &G, &T, &S, &ab};
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    if (e == hipSuccess)
        e = hipLaunchCooperativeKernel((const void*)k_exchange_floor, dim3(G), dim3(64), args, 0, st);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    unsigned flag = 0;
    if (e == hipSuccess) e = hipMemcpy(&flag, ab, 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return done(hip_fail(e, "exchange floor", __LINE__));
    if (flag) { set_error("exchange floor: a wait timed out"); return done(SPIRES_DEV_EINTERNAL); }
    *us_per_step = (double)ms * 1e3 / (double)T;
    return done(0);
}

// Raises g's exchange abort flag as a timed-out spike-exchange wait would, so
// that the next persistent launch gives up at its first unfinished wait and the
// error path can be exercised without a real timeout.
extern "C" int spires_dev_inject_exchange_timeout(spires_dev* g) {
    sdev::RandShield rand_shield_;
    using namespace sdev;
    if (!g) return SPIRES_DEV_EINVAL;
    if (const int rc = settle(g)) return rc;
    if (!g->abort_flag) {
        const int rc = dalloc(&g->abort_flag, 1);
        if (rc) return rc;
    }
    const unsigned one = 1u;
    SDEV_CK(hipMemcpyAsync(g->abort_flag, &one, 4, hipMemcpyHostToDevice, g->stream));
    SDEV_CK(hipStreamSynchronize(g->stream));
    g->abort_dirty = false;
    return 0;
}

extern "C" int spires_dev_plan_info(spires_dev* g, int* info) {
    sdev::RandShield rand_shield_;
    if (!g || !info) return SPIRES_DEV_EINVAL;
    for (int k = 0; k < 7; ++k) info[k] = 0;
    sdev::Plan pl, pb;
    const int ncu = sdev::num_cus();
    if (g->mode != SPIRES_DEV_MODE_EXACT || g->path == 1 || !sdev::persist_plan(g, ncu, &pl)) return 0;
    info[0] = pl.tp ? (pl.win ? 3 : 2) : 1;
    info[1] = pl.nb;
    info[2] = pl.G;
    info[3] = pl.rpm;
    info[4] = pl.ne_lds > 0 ? 1 : 0;
    info[5] = (int)pl.lds;
    info[6] = sdev::persist_plan(g, ncu, &pb, true) && pb.tp ? pb.nb : 0;
    return 0;
}

// ---------------------------------------------------------------------------
// Time blocks replayed from a HIP graph.  The launches of one block of K macro
// steps (product-wave batches, packed per neurons-per-workgroup class) are
// captured once; each kernel reads its reservoirs' step counters from a
// device clock, and a one-thread-per-reservoir kernel after it advances the
// clocks, so every replay runs the next K steps.  Abort flags are folded into
// a sticky count by that kernel and checked after the replays.

namespace sdev {

struct TbItem {
    long long* clk;       // [4]: micro-step, macro step, spike-buffer parity, unused
    unsigned* abort_flag;
    unsigned* sticky;
    long long steps;
    int micro;
};

__global__ void k_timeblock_advance(const TbItem* __restrict__ items, int count) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= count) return;
    const TbItem it = items[r];
    it.clk[0] += it.steps * it.micro;
    it.clk[1] += it.steps;
    it.clk[2] ^= 1;
    *it.sticky += *it.abort_flag;
    *it.abort_flag = 0u;
}

// ---------------------------------------------------------------------------
// All reservoirs of a time block in one persistent launch.  Workgroup b is one
// wave that owns the 64-neuron slice wg_map[b].y of reservoir wg_map[b].x for
// the whole block: lane = neuron.  Per micro-step the wave sums the GL history
// from the ring in HBM (the single-wave chain of k_persist, lags in ascending
// order), gathers the reservoir's spike words of the previous step, pulls its
// CSR rows from HBM (CSR order), forms the input drive and updates the neurons
// (flif_gl.c:50-100), then publishes its two spike words.  Reservoirs share
// nothing but the launch: a reservoir with a short history runs ahead of one
// with a long history, and the launch lasts as long as the slowest reservoir,
// not the sum over reservoirs.  Every slice of every reservoir is resident at
// once (cooperative launch; the host checks the occupancy).

// ---- one-launch kernel helpers: global-memory pointers ----------------------
// The reservoir arguments come from a table in memory, so the compiler cannot
// infer that their pointers address global memory and would use FLAT accesses
// (a 64-bit address per lane and per load, both wait counters).  The kernel
// converts them to address-space-1 pointers once and these helpers take them
// as such; each does what its counterpart in dev_exchange.hpp / gl_seg_hbm
// does, in the same order.
#define SPIRES_GLOBAL __attribute__((address_space(1)))
template <typename T>
using gptr = SPIRES_GLOBAL T*;

template <typename T>
__device__ __forceinline__ gptr<T> to_global(T* p) {
    return (gptr<T>)p;
}

// h += sum_{k=1..cnt} c[k] * ring[((head-k) mod L)*n + i], ascending k, one
// multiply and one add per lag (flif_gl.c:66-72).  The ring is read through a
// buffer resource: the lane's neuron is a fixed byte offset (voff), the slot a
// scalar offset computed per load, so a load costs no address registers;
// blocks of 16 lags, NBLK blocks in flight.  The coefficients come from LDS
// (cs, broadcast reads).
template <int NBLK>
__device__ __forceinline__ double gl_seg_mux(double h, __amdgpu_buffer_rsrc_t rs, unsigned voff, const double* cs,
                                             int head, int cnt, int L, unsigned rowb) {
    constexpr int B = 16;
    const int kb = 1, sb = head - 1;
    const int nblk = cnt / B;
    double v[NBLK][B];
    auto at = [&](int slot) {
        slot += slot < 0 ? L : 0;
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, voff, (unsigned)slot * rowb, 0));
    };
    auto load = [&](double* buf, int blk) {
#pragma unroll
        for (int j = 0; j < B; ++j) buf[j] = at(sb - blk * B - j);
    };
#pragma unroll
    for (int u = 0; u < NBLK; ++u)
        if (u < nblk) load(v[u], u);
    for (int b0 = 0; b0 < nblk; b0 += NBLK) {
#pragma unroll
        for (int u = 0; u < NBLK; ++u) {
            const int blk = b0 + u;
            if (blk < nblk) {
#pragma unroll
                for (int j = 0; j < B; ++j) h = h + cs[kb + blk * B + j] * v[u][j];
                if (blk + NBLK < nblk) load(v[u], blk + NBLK);
            }
        }
    }
    const int q = nblk * B, rem = cnt - q;
    for (int j = 0; j < rem; ++j) h = h + cs[kb + q + j] * at(sb - q - j);
    return h;
}

// gather_spikes for 64-neuron slices (two words each): spike bits of micro-step
// st-1 of one reservoir into bits_s[nw]; at the launch's first step from S.
__device__ __forceinline__ bool mux_gather(gptr<unsigned long long> xch, int nw, int n, long long st, long long st0,
                                           gptr<const uint8_t> S, unsigned* bits_s, gptr<unsigned> abort_flag,
                                           int* abort_s, int lane) {
    if (st == st0) {
        for (int w = lane; w < nw; w += 64) {
            unsigned b = 0;
            for (int j = 0; j < 32; ++j) {
                const int nj = w * 32 + j;   // slice w/2, lanes 32 (w%2) + j
                if (nj < n && S[nj]) b |= 1u << j;
            }
            bits_s[w] = b;
        }
        return true;
    }
    gptr<const unsigned long long> slot = xch + (size_t)((st - 1) & 1) * nw;
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
        __builtin_amdgcn_s_sleep(2);
    }
#pragma unroll
    for (int q = 0; q < kMaxWordsPerLane; ++q)
        if (lane + 64 * q < nw) bits_s[lane + 64 * q] = (unsigned)word[q];
    return ok;
}

// row_pull from the CSR in HBM (sparse.c:68-76: the row's edges in CSR order;
// simple synapses add the weights of spiking columns, PSC synapses advance
// each edge's trace in place and add w * trace).  Columns are natural neuron
// ids: bit j % 32 of word j / 32.
__device__ __forceinline__ double mux_row_pull(int stype, int rs, int re, gptr<const int32_t> col,
                                               gptr<const double> val, const unsigned* bits,
                                               gptr<const double> edge_decay, gptr<double> edge_tr) {
    double s = 0.0;
    if (stype == ST_SIMPLE) {
        int e = rs;
        for (; e + 8 <= re; e += 8) {   // loads of a batch issued before the in-order sum
            int32_t cj[8];
            double wv[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) { cj[q] = col[e + q]; wv[q] = val[e + q]; }
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if ((bits[cj[q] >> 5] >> (cj[q] & 31)) & 1u) s = s + wv[q];
        }
        for (; e < re; ++e) {
            const int32_t j = col[e];
            if ((bits[j >> 5] >> (j & 31)) & 1u) s = s + val[e];
        }
        return s;
    }
    for (int e = rs; e < re; ++e) {
        const int32_t j = col[e];
        const unsigned sbit = (bits[j >> 5] >> (j & 31)) & 1u;
        const double d = edge_decay[e];
        const double tr = edge_tr[e] * d + (1.0 - d) * (double)sbit;
        edge_tr[e] = tr;
        s = s + val[e] * tr;
    }
    return s;
}

constexpr int kMuxNb = 64;
constexpr int kMuxMaxWords = kMaxWordsPerLane * 64;   // exchange words of one reservoir (n <= 16384)
constexpr int kMuxBlocks = 2;                          // 16-lag history blocks in flight per lane

__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) k_mux_gl(
    const PersistArgs* __restrict__ table, const int2* __restrict__ wg_map) {
    __shared__ unsigned bits_s[kMuxMaxWords];
    __shared__ int abort_s;
    extern __shared__ __align__(16) double coef_s[];   // [L] GL coefficients of the slice's reservoir
    const int2 mw = wg_map[blockIdx.x];
    const PersistArgs& a = table[mw.x];
    const int lane = threadIdx.x, g = mw.y;
    const int n = a.n, L = a.L, wpg = 2, nw = a.G * 2, din = a.din, micro = a.micro, stype = a.stype;
    const int i0 = g * kMuxNb, nloc = min(kMuxNb, n - i0);
    const bool own = lane < nloc;
    const int i = i0 + (own ? lane : nloc - 1);      // lanes past the slice mirror its last neuron
    gptr<const double> W_in = to_global(a.W_in), u = to_global(a.u), ring_c = to_global((const double*)a.ring);
    gptr<double> ring = to_global(a.ring);
    gptr<const double> coef = to_global(a.coef), val = to_global(a.val), decay = to_global(a.edge_decay);
    gptr<double> edge_tr = to_global(a.edge_tr);
    gptr<const int32_t> col = to_global(a.col);
    gptr<const int64_t> rp = to_global(a.rp);
    gptr<unsigned long long> xch = to_global(a.xch);
    gptr<unsigned> abort_flag = to_global(a.abort_flag);
    gptr<double> Vg = to_global(a.V), tpg = to_global(a.tprev);
    gptr<const uint8_t> S_in = to_global((const uint8_t*)a.S);
    gptr<uint8_t> S_out = to_global(a.S_out);
    gptr<double> rec_V = to_global(a.rec_V);
    gptr<uint8_t> rec_S = to_global(a.rec_S);
    if (lane == 0) abort_s = 0;
    long long st0 = a.st0, t0 = a.t0;
    const long long T = a.T;
    const double gain = a.gain, dt = a.dt;
    if (a.clk) {
        const long long c0 = a.clk[0], c1 = a.clk[1], c2 = a.clk[2];
        st0 = (long long)(unsigned)__builtin_amdgcn_readfirstlane((int)c0) |
              ((long long)__builtin_amdgcn_readfirstlane((int)(c0 >> 32)) << 32);
        t0 = (long long)(unsigned)__builtin_amdgcn_readfirstlane((int)c1) |
             ((long long)__builtin_amdgcn_readfirstlane((int)(c1 >> 32)) << 32);
        if (__builtin_amdgcn_readfirstlane((int)c2) & 1) {
            S_in = to_global((const uint8_t*)a.S_out);
            S_out = to_global(const_cast<uint8_t*>(a.S));
        }
        if (rec_V) rec_V += (size_t)t0 * n;
        if (rec_S) rec_S += (size_t)t0 * n;
    }
    double V = Vg[i], tp = tpg[i];
    bool spk = S_in[i] != 0;
    // lanes past the slice pull no row (PSC traces advance once per edge)
    const int rs = own ? (int)rp[i] : 0, re = own ? (int)rp[i + 1] : 0;
    double p[7];
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        p[j] = a.np[j];
        asm volatile("" : "+v"(p[j]));
    }
    const unsigned ring_off = (unsigned)i * 8u;   // byte offset of the lane's neuron in a ring slot
    const unsigned rowb = (unsigned)n * 8u;        // bytes per ring slot (N*L*8 < 2^31: checked on the host)
    const __amdgpu_buffer_rsrc_t ring_rs =
        __builtin_amdgcn_make_buffer_rsrc((void*)a.ring, (short)0, (int)((unsigned)L * rowb), 0x00020000);
    for (int k = lane; k < L; k += 64) coef_s[k] = coef[k];
    __syncthreads();
    long long st = st0;
    int head = (int)(st % L);
    int lim = st < L ? (int)st : L - 1;
    for (long long t = 0; t < T; ++t) {
        const long long tt = t0 + t;
        double ext = 0.0;
        for (int mi = 0; mi < micro; ++mi, ++st) {
            // GL history of this micro-step (lags 1..lim in ascending order)
            const double H = gl_seg_mux<kMuxBlocks>(0.0, ring_rs, ring_off, coef_s, head, lim, L, rowb);
            // spikes of micro-step st-1, CSR rows, input drive
            const bool aborted = !mux_gather(xch, nw, n, st, st0, S_in, bits_s, abort_flag, &abort_s, lane);
            __syncthreads();
            if (__builtin_amdgcn_readfirstlane(abort_s) || aborted) break;
            const double rec = mux_row_pull(stype, rs, re, col, val, bits_s, decay, edge_tr);
            if (mi == 0) {
                double su = 0.0;
                for (int j = 0; j < din; ++j) su = su + W_in[(size_t)i * din + j] * u[tt * din + j];
                ext = gain * su;
            }
            // neuron update (flif_gl.c:50-100)
            const double x = ext + rec;
            spk = false;
            if (tp > p[6]) {
                // V_{st-1}: the lane's own V after the launch's first step (the same value
                // went to the ring), the ring value before
                const int sp = head == 0 ? L - 1 : head - 1;
                const double Vp = st == st0 ? ring[(size_t)sp * n + i] : V;
                const double rhs = (-(Vp - p[2]) / p[3]) + x + p[5];
                V = p[4] * rhs - H;
                if (V >= p[0]) { V = p[1]; spk = true; tp = 0.0; }
            } else {
                V = p[1];
            }
            // publish: words 2g and 2g+1 of the parity slot (tag st + 1, 32 spike bits each)
            const unsigned long long bal = __ballot(own && spk);
            if (lane < wpg) {
                const unsigned bits = (unsigned)(bal >> (32 * lane));
                const unsigned long long word = ((unsigned long long)(unsigned)(st + 1) << 32) | bits;
                __hip_atomic_store(&xch[(size_t)(st & 1) * nw + (size_t)g * wpg + lane], word, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
            if (own) {
                ring[(size_t)head * n + i] = V;
                tp = tp + dt;
                if (mi == micro - 1) {
                    if (rec_V) rec_V[(size_t)t * n + i] = V;
                    if (rec_S) rec_S[(size_t)t * n + i] = spk ? 1 : 0;
                }
            }
            head = head + 1 == L ? 0 : head + 1;
            if (lim < L - 1) ++lim;
            __syncthreads();   // bits_s is rewritten by the next gather
        }
        if (__builtin_amdgcn_readfirstlane(abort_s)) break;
    }
    if (own) {
        Vg[i] = V;
        S_out[i] = spk ? 1 : 0;
        tpg[i] = tp;
    }
}

}  // namespace sdev

struct spires_dev_timeblock {
    std::vector<spires_dev*> gs;     // in launch order
    long long K = 0;
    unsigned flags = 0;
    int launches = 0;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    sdev::PersistArgs* table = nullptr;
    int2* map = nullptr;
    sdev::TbItem* items = nullptr;
    long long* clk = nullptr;
    unsigned* sticky = nullptr;
    std::vector<const double*> u_at;
    std::vector<double*> recv_at;
    std::vector<uint8_t*> recs_at;
    std::vector<sdev::Plan> plans;          // per reservoir, launch order
    std::vector<sdev::PersistArgs> args;    // host copy of the table
};

extern "C" void spires_dev_timeblock_destroy(spires_dev_timeblock* tb) {
    sdev::RandShield rand_shield_;
    if (!tb) return;
    if (tb->stream) (void)hipStreamSynchronize(tb->stream);
    if (tb->exec) (void)hipGraphExecDestroy(tb->exec);
    if (tb->graph) (void)hipGraphDestroy(tb->graph);
    sdev::dfree(tb->table); sdev::dfree(tb->map); sdev::dfree(tb->items); sdev::dfree(tb->clk); sdev::dfree(tb->sticky);
    if (tb->stream) (void)hipStreamDestroy(tb->stream);
    delete tb;
}

extern "C" int spires_dev_timeblock_create(spires_dev* const* gs, int count, long long K, unsigned flags,
                                           spires_dev_timeblock** out) {
    sdev::RandShield rand_shield_;
    using namespace sdev;
    if (!gs || count < 1 || K < 1 || !out) return SPIRES_DEV_EINVAL;
    *out = nullptr;
    if (flags & ~(SPIRES_DEV_REC_STATE | SPIRES_DEV_REC_RASTER | SPIRES_DEV_REC_APPEND | SPIRES_DEV_TB_ONE_LAUNCH)) {
        set_error("time blocks record states and rasters only");
        return SPIRES_DEV_EINVAL;
    }
    const bool one = (flags & SPIRES_DEV_TB_ONE_LAUNCH) != 0;
    size_t mux_lds = 0;   // one-launch kernel: the GL coefficients of the longest history
    const int ncu = num_cus();
    std::vector<Plan> plans(count);
    for (int r = 0; r < count; ++r) {
        spires_dev* g = gs[r];
        if (!g) return SPIRES_DEV_EINVAL;
        if (const int rc = settle(g)) return rc;
        if (one) {
            // the one-launch kernel: FLIF-GL, CSR, simple or PSC synapses, any history length
            if (g->mode != SPIRES_DEV_MODE_EXACT || g->path == 1 || g->backend != BK_SPARSE || g->ntype != NT_FLIF_GL ||
                !(g->stype == ST_SIMPLE || g->stype == ST_PSC_HET) || g->L < 2) {
                set_error("one-launch time blocks need exact-mode FLIF-GL reservoirs with CSR weights");
                return SPIRES_DEV_EINVAL;
            }
        } else if (g->mode != SPIRES_DEV_MODE_EXACT || g->path == 1 || !persist_plan(g, ncu, &plans[r], true) ||
                   !plans[r].tp) {
            set_error("time blocks need exact-mode FLIF-GL reservoirs that fit the product-wave kernel");
            return SPIRES_DEV_EINVAL;
        }
        if (g->din > 0 && !g->u_series) { set_error("time blocks need uploaded inputs"); return SPIRES_DEV_EINVAL; }
        if ((flags & SPIRES_DEV_REC_STATE) && !g->rec_state) { set_error("reserve state records before creating a time block"); return SPIRES_DEV_EINVAL; }
        if ((flags & SPIRES_DEV_REC_RASTER) && !g->rec_raster) { set_error("reserve raster records before creating a time block"); return SPIRES_DEV_EINVAL; }
        for (int q = 0; q < r; ++q)
            if (gs[q] == g) { set_error("a reservoir appears twice in a time block"); return SPIRES_DEV_EINVAL; }
    }
    if (one) {
        int total = 0;
        for (int r = 0; r < count; ++r) {
            if (gs[r]->n > kMuxMaxWords * 32) { set_error("one-launch time blocks take reservoirs of at most 16384 neurons"); return SPIRES_DEV_EINVAL; }
            total += (gs[r]->n + kMuxNb - 1) / kMuxNb;
        }
        int per_cu = 0;
        int Lmax = 0;
        for (int r = 0; r < count; ++r) {
            Lmax = std::max(Lmax, gs[r]->L);
            if ((long long)gs[r]->n * gs[r]->L * 8 >= (1ll << 31)) {
                set_error("one-launch time blocks address a history ring of at most 2^31 bytes");
                return SPIRES_DEV_EINVAL;
            }
        }
        mux_lds = (size_t)Lmax * 8;
        SDEV_CK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_mux_gl, 64, mux_lds));
        if ((long long)per_cu * ncu < total) {
            set_error("one-launch time block: " + std::to_string(total) + " slices do not fit the " +
                      std::to_string((long long)per_cu * ncu) + " resident workgroups");
            return SPIRES_DEV_EINVAL;
        }
    }
    // pack launches per neurons-per-workgroup class, at most one workgroup per CU.
    // A launch lasts as long as its slowest member, whose step time grows with its
    // history length: reservoirs are placed longest history first (then most
    // workgroups first) into the first launch with room, so a shorter history only
    // fills room next to longer ones and never lengthens a launch.
    std::vector<std::vector<int>> launches;
    if (one) {   // every reservoir in one launch of 64-neuron slices
        launches.emplace_back();
        for (int r = 0; r < count; ++r) {
            launches[0].push_back(r);
            plans[r] = Plan();
            plans[r].nb = kMuxNb;
            plans[r].G = (gs[r]->n + kMuxNb - 1) / kMuxNb;
        }
    }
    for (int nb : {32, 16}) {
        if (one) break;
        std::vector<int> idx;
        for (int r = 0; r < count; ++r)
            if (plans[r].nb == nb) idx.push_back(r);
        std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) {
            if (gs[x]->L != gs[y]->L) return gs[x]->L > gs[y]->L;
            return plans[x].G > plans[y].G;
        });
        std::vector<std::vector<int>> bins;
        std::vector<int> used;
        for (int r : idx) {
            size_t b = 0;
            while (b < bins.size() && used[b] + plans[r].G > ncu) ++b;
            if (b == bins.size()) { bins.emplace_back(); used.push_back(0); }
            bins[b].push_back(r);
            used[b] += plans[r].G;
        }
        for (auto& v : bins) launches.push_back(v);
    }
    spires_dev_timeblock* tb = new spires_dev_timeblock();
    tb->K = K;
    tb->flags = flags;
    tb->launches = (int)launches.size();
    auto fail = [&](int rc) { spires_dev_timeblock_destroy(tb); return rc; };
    std::vector<PersistArgs> table;
    std::vector<int2> map;
    std::vector<TbItem> items;
    std::vector<size_t> map_off, item_off, lds;
    std::vector<int> nbs;
    size_t lds32 = 0, lds16 = 0;
    int total = 0;
    for (auto& L : launches) total += (int)L.size();
    int rc = 0;
    if ((rc = dalloc(&tb->clk, (size_t)total * 4)) || (rc = dalloc(&tb->sticky, (size_t)total))) return fail(rc);
    for (auto& L : launches) {
        map_off.push_back(map.size());
        item_off.push_back(items.size());
        size_t l = 0;
        for (int r : L) {
            spires_dev* g = gs[r];
            const Plan& pl = plans[r];
            if (!g->ring64_valid) {
                if ((rc = fast_sync_rings(g, false))) return fail(rc);
                g->ring64_valid = 1;
            }
            g->ring32_valid = 0;
            if ((rc = prepare_launch(g, pl))) return fail(rc);
            const int idx = (int)table.size();
            PersistArgs a = make_args(g, pl, g->din > 0 ? g->u_series : nullptr, 0, K,
                                      (flags & SPIRES_DEV_REC_STATE) ? g->rec_state : nullptr,
                                      (flags & SPIRES_DEV_REC_RASTER) ? g->rec_raster : nullptr);
            a.S = g->S[0];
            a.S_out = g->S[1];
            a.clk = tb->clk + 4 * idx;
            table.push_back(a);
            for (int b = 0; b < pl.G; ++b) map.push_back(make_int2(idx, b));
            items.push_back(TbItem{tb->clk + 4 * idx, g->abort_flag, tb->sticky + idx, K, g->micro});
            tb->gs.push_back(g);
            tb->u_at.push_back(g->u_series);
            tb->recv_at.push_back(g->rec_state);
            tb->recs_at.push_back(g->rec_raster);
            tb->plans.push_back(pl);
            l = std::max(l, pl.lds);
        }
        lds.push_back(l);
        const int nb = plans[L[0]].nb;
        nbs.push_back(nb);
        if (nb != kMuxNb) (nb == 32 ? lds32 : lds16) = std::max(nb == 32 ? lds32 : lds16, l);
    }
    // one launch: the slices of the shortest histories first (then the smallest reservoirs).
    // Workgroups are dispatched in order, so every CU gets a like mix either way; with the
    // short histories first (the older waves) config 2's block is 6 % shorter than with the
    // longest first or in the caller's order (187-190 against 200 us per step).
    if (one)
        std::stable_sort(map.begin(), map.end(), [&](const int2& x, const int2& y) {
            if (tb->gs[x.x]->L != tb->gs[y.x]->L) return tb->gs[x.x]->L < tb->gs[y.x]->L;
            return tb->gs[x.x]->n < tb->gs[y.x]->n;
        });
    if ((rc = dalloc(&tb->table, table.size())) || (rc = dalloc(&tb->map, map.size())) ||
        (rc = dalloc(&tb->items, items.size())))
        return fail(rc);
    hipError_t e = hipStreamCreate(&tb->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(tb->table, table.data(), table.size() * sizeof(PersistArgs), hipMemcpyHostToDevice, tb->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(tb->map, map.data(), map.size() * sizeof(int2), hipMemcpyHostToDevice, tb->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(tb->items, items.data(), items.size() * sizeof(TbItem), hipMemcpyHostToDevice, tb->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(tb->stream);
    if (e == hipSuccess && lds32) e = hipFuncSetAttribute((const void*)k_persist_tp_batch<32>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds32);
    if (e == hipSuccess && lds16) e = hipFuncSetAttribute((const void*)k_persist_tp_batch<16>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds16);
    if (e != hipSuccess) { set_error(hipGetErrorString(e)); return fail(SPIRES_DEV_EINTERNAL); }
    tb->args = table;
    // capture one block: each launch (cooperative, so that the runtime checks that
    // all its workgroups can be resident at once, as the spike exchange needs),
    // then the clock advance of its reservoirs
    e = hipStreamBeginCapture(tb->stream, hipStreamCaptureModeThreadLocal);
    for (size_t q = 0; q < launches.size() && e == hipSuccess; ++q) {
        const int G = (int)((q + 1 < launches.size() ? map_off[q + 1] : map.size()) - map_off[q]);
        const int cnt = (int)launches[q].size();
        const PersistArgs* tab = tb->table;
        const int2* mp = tb->map + map_off[q];
        void* kargs[] = {(void*)&tab, (void*)&mp};
        if (nbs[q] == kMuxNb)
            e = hipLaunchCooperativeKernel((const void*)k_mux_gl, dim3(G), dim3(64), kargs, mux_lds, tb->stream);
        else
            e = hipLaunchCooperativeKernel(nbs[q] == 32 ? (const void*)k_persist_tp_batch<32>
                                                        : (const void*)k_persist_tp_batch<16>,
                                           dim3(G), dim3(kTpThreads), kargs, (unsigned)lds[q], tb->stream);
        if (e != hipSuccess) break;
        hipLaunchKernelGGL(k_timeblock_advance, dim3(1), dim3(64 * ((cnt + 63) / 64)), 0, tb->stream,
                           (const TbItem*)(tb->items + item_off[q]), cnt);
        e = hipGetLastError();
    }
    hipError_t e2 = hipStreamEndCapture(tb->stream, &tb->graph);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&tb->exec, tb->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) { set_error(std::string("time block capture: ") + hipGetErrorString(e)); return fail(SPIRES_DEV_EINTERNAL); }
    *out = tb;
    return 0;
}

extern "C" int spires_dev_timeblock_launches(const spires_dev_timeblock* tb) { return tb ? tb->launches : 0; }

extern "C" int spires_dev_timeblock_run(spires_dev_timeblock* tb, long long blocks) {
    sdev::RandShield rand_shield_;
    using namespace sdev;
    if (!tb || blocks < 0) return SPIRES_DEV_EINVAL;
    if (blocks == 0) return 0;
    const size_t cnt = tb->gs.size();
    const long long steps = blocks * tb->K;
    std::vector<long long> clk(cnt * 4, 0);
    for (spires_dev* g : tb->gs) {
        const int rc = settle(g);
        if (rc) return rc;
    }
    for (size_t r = 0; r < cnt; ++r) {
        spires_dev* g = tb->gs[r];
        if (g->u_series != tb->u_at[r] || g->rec_state != tb->recv_at[r] || g->rec_raster != tb->recs_at[r]) {
            set_error("inputs or records were reallocated after the time block was made");
            return SPIRES_DEV_EINVAL;
        }
        if (g->mode != SPIRES_DEV_MODE_EXACT) { set_error("time blocks run in exact mode"); return SPIRES_DEV_EINVAL; }
        const size_t end = g->msteps + (size_t)steps;
        if (g->din > 0 && end > g->u_T) { set_error("run beyond uploaded inputs"); return SPIRES_DEV_EINVAL; }
        if (((tb->flags & SPIRES_DEV_REC_STATE) && end * g->n > g->rec_cap_s) ||
            ((tb->flags & SPIRES_DEV_REC_RASTER) && end * g->n > g->rec_cap_r)) {
            set_error("records too small for the time blocks (rows = macro step index)");
            return SPIRES_DEV_EINVAL;
        }
        if (!g->ring64_valid) { set_error("the FP64 history is stale (fast-mode run after the time block was made)"); return SPIRES_DEV_EINVAL; }
        // a run outside the block may have grown (reallocated) the exchange words:
        // size them for this block's plan again and point the table at them
        if (const int rc = prepare_launch(g, tb->plans[r])) return rc;
        if (tb->args[r].xch != g->xch) {
            tb->args[r].xch = g->xch;
            SDEV_CK(hipMemcpyAsync(&tb->table[r].xch, &tb->args[r].xch, sizeof(g->xch), hipMemcpyHostToDevice, tb->stream));
        }
        SDEV_CK(hipStreamSynchronize(g->stream));
        clk[4 * r + 0] = g->step;
        clk[4 * r + 1] = (long long)g->msteps;
        clk[4 * r + 2] = g->cur;
    }
    SDEV_CK(hipMemcpyAsync(tb->clk, clk.data(), clk.size() * 8, hipMemcpyHostToDevice, tb->stream));
    SDEV_CK(hipMemsetAsync(tb->sticky, 0, cnt * 4, tb->stream));
    for (spires_dev* g : tb->gs) SDEV_CK(hipMemsetAsync(g->abort_flag, 0, 4, tb->stream));
    for (long long b = 0; b < blocks; ++b) SDEV_CK(hipGraphLaunch(tb->exec, tb->stream));
    std::vector<unsigned> sticky(cnt, 0);
    SDEV_CK(hipMemcpyAsync(sticky.data(), tb->sticky, cnt * 4, hipMemcpyDeviceToHost, tb->stream));
    SDEV_CK(hipStreamSynchronize(tb->stream));
    for (size_t r = 0; r < cnt; ++r) {
        spires_dev* g = tb->gs[r];
        g->step += steps * g->micro;
        g->msteps += (size_t)steps;
        g->cur ^= (int)(blocks & 1);
        g->ring32_valid = 0;
        if (tb->flags & (SPIRES_DEV_REC_STATE | SPIRES_DEV_REC_RASTER)) g->rec_rows = std::max(g->rec_rows, g->msteps);
    }
    for (size_t r = 0; r < cnt; ++r)
        if (sticky[r]) { set_error("a time-block kernel timed out waiting for the spike exchange"); return SPIRES_DEV_EINTERNAL; }
    return 0;
}
