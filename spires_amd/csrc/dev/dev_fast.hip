// FP32 fast mode for FLIF-GL: time-blocked history contraction on MFMA.
//
// For a block of K micro-steps starting at n0, the GL history of step n0+j
// (flif_gl.c:66-72) splits into
//   H[n0+j] = sum_{m=1}^{min(n0, L-1-j)} c_{m+j} V[n0-m]      (samples before the block)
//           + sum_{k=1}^{j} c_k V[n0+j-k]                      (in-block triangle)
// The first part, for all j at once, is the [nb x L] . [L x K] product of the
// workgroup's FP32 voltage history (ring in HBM, slot = time mod L)
// with a Toeplitz coefficient matrix B[s][j] = c_{m(s)+j}, m(s) = (n0-s) mod L,
// generated on the fly from the coefficients in LDS.  Four waves compute it
// with v_mfma_f32_16x16x4f32 at each block start (below); then the K steps run like
// the exact persistent kernel (spike exchange on wave 1) with only the j-term
// triangle summed step-wise from LDS.  The membrane update, threshold
// comparison and reset stay FP64 (V = pow(dt,alpha) rhs - H), with rhs grouped
// as ((leak + ext) + bias) + rec, so that only the last addition waits for the
// exchange; the results are deterministic but the additions are not in the
// exact kernel's order.
//
// Neurons are interleaved over the G workgroups: neuron i belongs to workgroup
// i mod G as its lane i div G.  The hubs of a scale-free graph (the lowest
// neuron ids) are thereby spread over the workgroups instead of loading one
// with most of the edges.  The FP32 ring is kept in workgroup order ([L][64 G],
// column 64 g + r for neuron g + G r), so that a workgroup's history columns
// are contiguous; the conversions from and to the FP64 ring apply the order.
//
// History contraction: each wave sums a quarter of the ring's slots for all 64
// neurons (float4 loads, eight in flight per lane), and the four partial sums
// are added into O in wave order, so the result does not depend on timing.
//
// In-block history: H[n0 + j] = O[j] + c_1 Vb[j-1] + T_j with
// T_j = sum_{k=2..j} c_k Vb[j-k].  T_{j+1} needs only voltages up to step j - 1,
// which wave 0 wrote itself, so wave 0 sums it (in a register) while it waits
// for step j's exchange, and adds the three terms at the next update.
//
// Row pull: each row's CSR range is cut into segments of kSegEdges edges; the
// lanes of waves 1-3 sum the segments (each in CSR order, from +0.0) after the
// exchange, and the row's lane adds its segments' partial sums in segment
// order.  A row of at most kSegEdges edges is one segment, summed exactly as
// the exact kernel sums it; a hub row of a scale-free graph (hundreds of
// in-edges) no longer serialises the pull on one lane.  In LDS the edges are
// stored segment-interleaved (edge q of segment sg at q * seg_max + sg), so the
// lanes of a wave, which take consecutive segments, read consecutive words.
#include "dev_internal.hpp"
#include "dev_exchange.hpp"

#include <algorithm>
#include <cstring>

namespace sdev {

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kFastThreads = 256;     // 4 waves: contraction on all; per step wave 0 sums the
                                      // triangle and updates, wave 1 exchanges, waves 1-3 pull
constexpr int kFastNb = 64;           // neurons per workgroup (one wave of lanes)
constexpr int kSegEdges = 8;          // edges per row-pull segment (the pull loads 8 slots)
constexpr int kOStride = kFastNb + 1; // row stride of O in LDS (padded: conflict-free column writes)

struct FastArgs {
    int n, din, L, nb, G, wpg, micro, K, stype;
    int ne_lds;                   // CSR edge slots in LDS, kSegEdges * seg_max (0: CSR from HBM)
    long long st0, t0, T;
    double gain, dt;
    double np[16];
    const double* W_in;
    const double* u;
    const int64_t* rp; const int32_t* col; const double* val;
    const double* edge_decay; double* edge_tr;
    double* V; const uint8_t* S; uint8_t* S_out; double* tprev;
    float* ring;                  // [L][ldr] FP32, workgroup order
    int ldr;                      // ring row length, 64 G
    const double* coef;           // FP64 coefficients (converted to FP32 in LDS)
    unsigned long long* xch;
    unsigned* abort_flag;
    double* rec_V; uint8_t* rec_S;
    // segments: [G + 1] first segment of each workgroup, [G][nb + 1] first segment
    // of each row (relative to its workgroup's first), [segments][2] CSR edge range
    const int* seg;
    int seg_max;
};

// Exchange word and bit of column j when neuron j is lane j / G of workgroup j mod G.
__device__ __forceinline__ unsigned encode_col_ileave(int j, int G, int wpg) {
    const int gg = j % G, r = j / G;
    return ((unsigned)(gg * wpg + (r >> 5)) << 5) | (unsigned)(r & 31);
}

// The same column as (byte offset of its exchange word << 5) | bit: the LDS
// address of the word (the spike words start the dynamic LDS) is code >> 5,
// and v_bfe takes the bit from the low five bits of the code itself.
__device__ __forceinline__ unsigned encode_col_lds(int j, int G, int wpg) {
    const int gg = j % G, r = j / G;
    return ((unsigned)(gg * wpg + (r >> 5)) << 7) | (unsigned)(r & 31);
}

// Row pull from the CSR in HBM (segment tables too large for the LDS): edges
// [b, e) in CSR order, as row_pull does for the natural neuron order.
__device__ __forceinline__ double row_pull_ileave(int stype, int b, int e, const int32_t* __restrict__ col,
                                                  const double* __restrict__ val, const unsigned* __restrict__ bits,
                                                  int G, int wpg, const double* __restrict__ edge_decay,
                                                  double* __restrict__ edge_tr) {
    double s = 0.0;
    for (int q = b; q < e; ++q) {
        const unsigned pk = encode_col_ileave(col[q], G, wpg);
        const unsigned sbit = (bits[pk >> 5] >> (pk & 31)) & 1u;
        if (stype == ST_SIMPLE) {
            s = s + (sbit ? val[q] : 0.0);
        } else {
            const double d = edge_decay[q];
            const double tr = edge_tr[q] * d + (1.0 - d) * (double)sbit;
            edge_tr[q] = tr;
            s = s + val[q] * tr;
        }
    }
    return s;
}

static_assert(kSegEdges == 8, "the row pull loads a segment as 8 slots");

__global__ void __launch_bounds__(kFastThreads) k_fast(FastArgs a) {
    extern __shared__ __align__(16) float flds[];   // 16-byte aligned: b128 LDS reads
    // wave index and abort flag are read through readfirstlane so the compiler
    // treats them as wave-uniform (scalar branches, scalar loop counters).
    const int g = blockIdx.x, tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nb = a.nb, L = a.L, K = a.K;
    const int nloc = (a.n - g + a.G - 1) / a.G;   // lanes r with g + G r < n
    const int wpg = a.wpg, nw = a.G * wpg;
    // LDS (each part 16-byte aligned): spike bits [nw] (first, at a fixed address),
    // coefficients [L] (FP32), pre-block contraction O [K][nb + 1], in-block voltages
    // Vb [K][nb], segment partial sums
    // [seg_max + 8] (FP64), ext [2][nb] (FP64), CSR values [seg_max][kSegEdges]
    // (FP64), CSR columns [seg_max][kSegEdges] (encode_col_lds), segment ranges
    // [seg_max][2], row segment offsets [nb + 1]
    const int nwp = (nw + 3) & ~3, Lp = (L + 3) & ~3;
    unsigned* bits_s = (unsigned*)flds;
    float* c_s = flds + nwp;
    float* O_s = c_s + Lp;
    float* Vb_s = O_s + ((K * kOStride + 3) & ~3);
    double* part_s = (double*)(Vb_s + K * nb);
    double* ext_s = part_s + ((a.seg_max + 8 + 1) & ~1);
    double* eval_s = ext_s + 2 * nb;
    unsigned* eidx_s = (unsigned*)(eval_s + a.ne_lds);
    int* segr_s = (int*)(eidx_s + a.ne_lds);
    int* rowf_s = segr_s + 2 * a.seg_max;
    __shared__ int abort_s;
    if (tid == 0) abort_s = 0;
    for (int k = tid; k < L; k += kFastThreads) c_s[k] = (float)a.coef[k];
    const bool csr_lds = a.ne_lds > 0;
    const int sg0 = a.seg[g], nseg = a.seg[g + 1] - sg0;
    const int* segr = a.seg + (a.G + 1) + (size_t)a.G * (nb + 1) + 2 * (size_t)sg0;
    for (int e = tid; e < 2 * nseg; e += kFastThreads) segr_s[e] = segr[e];
    // every segment owns kSegEdges slots; the slots past its length hold column 0
    // with weight 0 so that the pull loads them unconditionally
    if (csr_lds)
        for (int sg = tid; sg < nseg; sg += kFastThreads) {
            const int b = segr[2 * sg], len = segr[2 * sg + 1] - b;
            for (int q = 0; q < kSegEdges; ++q) {
                eidx_s[sg * kSegEdges + q] = q < len ? encode_col_lds(a.col[b + q], a.G, wpg) : 0u;
                eval_s[sg * kSegEdges + q] = q < len ? a.val[b + q] : 0.0;
            }
        }
    for (int e = tid; e < a.seg_max + 8; e += kFastThreads) part_s[e] = 0.0;
    for (int e = tid; e <= nb; e += kFastThreads) rowf_s[e] = a.seg[(a.G + 1) + (size_t)g * (nb + 1) + e];
    const bool own = wave == 0 && lane < nloc;
    const int i = g + a.G * lane;
    double V = own ? a.V[i] : 0.0;
    double tp = own ? a.tprev[i] : 0.0;
    bool spk = own ? (a.S[i] != 0) : false;
    constexpr int kMaxDinReg = 16;
    double win[kMaxDinReg];
    const bool win_reg = a.din <= kMaxDinReg;
#pragma unroll
    for (int j = 0; j < kMaxDinReg; ++j) win[j] = 0.0;
    if (wave == 1 && lane < nloc) {
        if (win_reg)
            for (int j = 0; j < a.din; ++j) win[j] = a.W_in[(size_t)i * a.din + j];
    }
    // neuron parameters (V_th, V_reset, V_rest, tau_m, pow(dt, alpha), bias, t_ref)
    // and dt in LDS: read at their point of use, they hold no scalar registers
    // across the step loop (a kernel-argument array there is spilled and reloaded
    // as a whole)
    __shared__ double np_s[8];
    if (tid < 7) np_s[tid] = a.np[tid];
    if (tid == 7) np_s[7] = a.dt;
    const double* p = np_s;
    float vb_last = 0.f;   // wave 0: FP32 voltage of the previous step of the block
    float tri = 0.f;       // wave 0: T_j of the coming step
    __syncthreads();
    // wave 0: the row's segment range in part_s (fixed for the launch)
    const int f0 = own ? rowf_s[lane] : 0, f1 = own ? rowf_s[lane + 1] : 0;

    long long st = a.st0;
    long long t = 0;
    int mi = 0;
    int head = (int)(st % L);   // ring slot of micro-step st, advanced without division
    while (t < a.T) {
        // ---------------- block start: O = Hist . B on MFMA (all 4 waves) ----------------
        const long long n0 = st;
        {
            // Wave w sums the history slots of its quarter of the ring, all nb = 64
            // neurons and all K columns.  Lane l reads slot s = 4 g + l / 16 of slot
            // group g and the four neurons 4 (l % 16) + c (one float4); MFMA c takes
            // component c, so its row i is neuron 4 i + c and its k is the slot.
            // The four partial sums are added into O in wave order.
            const int kq = lane >> 4, li = lane & 15;
            floatx4 acc[4][4];   // [c][q]
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[c][q] = floatx4{0.f, 0.f, 0.f, 0.f};
            const int ntile = K / 16;
            // ring slot s holds V[n0 - mm], mm = (n0 - s) mod L steps before the block
            const int nm = head;
            const int ng = (L + 3) / 4, gpw = (ng + 3) / 4;
            const int g_lo = wave * gpw, g_hi = min(ng, g_lo + gpw);
            // loads are unconditional, the slot clamped in bounds; what a clamped load
            // brings is multiplied by a zero coefficient.  Columns of lanes without a
            // neuron hold zeros and land in rows of O that are never read.
            const float* ring_g = a.ring + (size_t)g * kFastNb + 4 * li;   // this lane's 4 columns
            constexpr int kBatch = 8;   // float4 ring loads in flight before their MFMAs
            for (int g0 = g_lo; g0 < g_hi; g0 += kBatch) {
                floatx4 av[kBatch];
                int mmv[kBatch];
#pragma unroll
                for (int b = 0; b < kBatch; ++b) {
                    const int s = 4 * (g0 + b) + kq;
                    const bool ok = g0 + b < g_hi && s < L;
                    av[b] = *reinterpret_cast<const floatx4*>(ring_g + (size_t)min(s, L - 1) * a.ldr);
                    int mm = nm - s;
                    if (mm < 0) mm += L;
                    mmv[b] = (ok && mm >= 1 && (long long)mm <= n0) ? mm : -(1 << 30);   // no coefficient
                }
#pragma unroll
                for (int b = 0; b < kBatch; ++b) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        if (q < ntile) {
                            const int idx = mmv[b] + 16 * q + li;
                            const float cv = c_s[min(max(idx, 0), L - 1)];
                            const float bv = (idx >= 1 && idx <= L - 1) ? cv : 0.f;
#pragma unroll
                            for (int c = 0; c < 4; ++c)
                                acc[c][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[b][c], bv, acc[c][q], 0, 0, 0);
                        }
                    }
                }
            }
            // D layout: lane l, reg r -> row 4 (l / 16) + r (neuron 4 row + c), column l % 16
            for (int w2 = 0; w2 < 4; ++w2) {
                if (wave == w2) {
#pragma unroll
                    for (int c = 0; c < 4; ++c)
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            if (q < ntile) {
#pragma unroll
                                for (int r = 0; r < 4; ++r) {
                                    const int nl = 4 * (4 * kq + r) + c;
                                    float* o = O_s + (16 * q + li) * kOStride + nl;
                                    *o = (w2 == 0) ? acc[c][q][r] : *o + acc[c][q][r];
                                }
                            }
                        }
                }
                __syncthreads();
            }
        }
        // ---------------- K steps (one workgroup barrier each) ----------------
        for (int j = 0; j < K && t < a.T; ++j) {
            const long long tt = a.t0 + t;
            const int tb = (int)(t & 1);
            // wave 0: the history term H and the leak of the update, which need only
            // the previous steps, then T_{j+1} for the next step.  Wave 1: spikes of
            // the previous step and the input drive.
            double H = 0.0, leak = 0.0;
            if (wave == 0) {
                if (own && tp > p[6]) {
                    float h = O_s[j * kOStride + lane];
                    if (j >= 1) h = h + c_s[1] * vb_last;
                    if (j >= 2) h = h + tri;
                    H = (double)h;
                    leak = -(V - p[2]) / p[3];
                }
                // T_{j+1} = sum_{k=2..j+1} c_k Vb[j+1-k] in batches of 8: loads from
                // clamped indices first, a term past k = j + 1 gets coefficient 0 (adds
                // an exact zero)
                float h = 0.f;
                if (j + 1 < K)
                    for (int k = 2; k <= j + 1; k += 8) {
                        float cv[8], vv[8];
#pragma unroll
                        for (int q = 0; q < 8; ++q) {
                            cv[q] = c_s[min(k + q, L - 1)];
                            vv[q] = Vb_s[max(j + 1 - k - q, 0) * nb + lane];
                        }
#pragma unroll
                        for (int q = 0; q < 8; ++q) h = h + (k + q <= j + 1 ? cv[q] : 0.f) * vv[q];
                    }
                tri = h;
            }
            if (wave == 1) {
                double uin[kMaxDinReg];
                if (mi == 0 && win_reg) {
#pragma unroll
                    for (int q = 0; q < kMaxDinReg; ++q) uin[q] = (q < a.din) ? a.u[tt * a.din + q] : 0.0;
                }
                const bool aborted = !gather_spikes(a.xch, nw, nb, wpg, a.n, st, a.st0, a.S, bits_s,
                                                    a.abort_flag, &abort_s, lane, a.G);
                if (!aborted && lane < nloc && mi == 0) {
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
            __syncthreads();   // spike bits and ext in LDS (abort_s is only set before this barrier)
            if (__builtin_amdgcn_readfirstlane(abort_s)) break;
            // wave 0, while the others pull: the update's terms that do not depend on
            // the recurrent input, a0 = (leak + ext) + bias
            double a0 = 0.0;
            if (wave == 0 && own && tp > p[6]) a0 = (leak + ext_s[tb * nb + lane]) + p[5];
            // waves 1-3: segment partial sums
            if (wave >= 1)
                for (int sg = tid - 64; sg < nseg; sg += 3 * 64) {
                const int b = segr_s[2 * sg], e = segr_s[2 * sg + 1];
                double su = 0.0;
                if (!csr_lds) {
                    su = row_pull_ileave(a.stype, b, e, a.col, a.val, bits_s, a.G, wpg, a.edge_decay, a.edge_tr);
                } else if (a.stype == ST_SIMPLE) {
                    // the segment's kSegEdges slots (padded: column code 0, weight 0) as
                    // b128 loads, then the spike words, then the in-order sum; a silent
                    // edge's weight is masked to +0.0, which leaves su unchanged (su
                    // starts at +0.0, so it is never -0.0)
                    const uint4* ep = reinterpret_cast<const uint4*>(eidx_s + sg * kSegEdges);
                    const double2* wp = reinterpret_cast<const double2*>(eval_s + sg * kSegEdges);
                    const uint4 e0 = ep[0], e1 = ep[1];
                    const double2 w0 = wp[0], w1 = wp[1], w2 = wp[2], w3 = wp[3];
                    const unsigned pk[8] = {e0.x, e0.y, e0.z, e0.w, e1.x, e1.y, e1.z, e1.w};
                    const double wv[8] = {w0.x, w0.y, w1.x, w1.y, w2.x, w2.y, w3.x, w3.y};
                    unsigned bw[8];
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                        bw[q] = *reinterpret_cast<const unsigned*>(reinterpret_cast<const char*>(bits_s) + (pk[q] >> 5));
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        const unsigned m = (unsigned)__builtin_amdgcn_sbfe((int)bw[q], pk[q], 1);   // 0 or ~0
                        const unsigned long long wb = __builtin_bit_cast(unsigned long long, wv[q]) &
                                                      (((unsigned long long)m << 32) | m);
                        su = su + __builtin_bit_cast(double, wb);
                    }
                } else {   // per-edge synaptic traces in HBM (edge index b + q)
                    for (int q = 0; q < e - b; ++q) {
                        const unsigned pk = eidx_s[sg * kSegEdges + q];
                        const unsigned sbit = (bits_s[pk >> 7] >> (pk & 31)) & 1u;
                        const int64_t k = b + q;
                        const double d = a.edge_decay[k];
                        const double tr = a.edge_tr[k] * d + (1.0 - d) * (double)sbit;
                        a.edge_tr[k] = tr;
                        su = su + eval_s[sg * kSegEdges + q] * tr;
                    }
                }
                part_s[sg] = su;
            }
            __syncthreads();   // segment partial sums ready
            if (wave == 0) {
                if (own) {
                    // recurrent input: the row's segment partial sums, 8 at a time added
                    // as a tree (depth 3; slots past the row hold +0.0 after masking)
                    double rec = 0.0;
                    for (int f = f0; f < f1; f += 8) {
                        double pv[8];   // part_s has 8 spare entries past the last segment
#pragma unroll
                        for (int q = 0; q < 8; ++q) pv[q] = part_s[f + q];
#pragma unroll
                        for (int q = 0; q < 8; ++q) pv[q] = f + q < f1 ? pv[q] : 0.0;
                        const double s4a = (pv[0] + pv[1]) + (pv[2] + pv[3]);
                        const double s4b = (pv[4] + pv[5]) + (pv[6] + pv[7]);
                        rec = rec + (s4a + s4b);
                    }
                    spk = false;
                    if (tp > p[6]) {
                        // V = alpha_term * ((leak + ext + bias) + rec) - H
                        V = p[4] * (a0 + rec) - H;
                        if (V >= p[0]) { V = p[1]; spk = true; tp = 0.0; }
                    } else {
                        V = p[1];
                    }
                }
                publish_spikes(a.xch, nw, g, wpg, st, lane, __ballot(own && spk));
                if (own) {
                    tp = tp + p[7];
                    const float vf = (float)V;
                    Vb_s[j * nb + lane] = vf;
                    vb_last = vf;
                    a.ring[(size_t)head * a.ldr + (size_t)g * kFastNb + lane] = vf;
                    if (mi == a.micro - 1) {
                        if (a.rec_V) a.rec_V[(size_t)t * a.n + i] = V;
                        if (a.rec_S) a.rec_S[(size_t)t * a.n + i] = spk ? 1 : 0;
                    }
                }
            }
            ++st;
            if (++head == L) head = 0;
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

// FP64 ring [L][n] (natural order) -> FP32 ring [L][64 G] (workgroup order; the
// columns of lanes without a neuron are zero), and back.
__global__ void k_ring_f64_to_f32(int L, int n, int G, const double* src, float* dst) {
    const size_t ldr = (size_t)G * kFastNb;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)L * ldr) return;
    const size_t s = q / ldr;
    const int p = (int)(q - s * ldr);
    const int i = (p / kFastNb) + G * (p % kFastNb);
    dst[q] = i < n ? (float)src[s * n + i] : 0.f;
}
__global__ void k_ring_f32_to_f64(int L, int n, int G, const float* src, double* dst) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)L * n) return;
    const size_t s = q / n;
    const int i = (int)(q - s * n);
    dst[q] = (double)src[s * (size_t)G * kFastNb + (size_t)(i % G) * kFastNb + i / G];
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
    const int G = (g->n + kFastNb - 1) / kFastNb;
    const size_t cnt = (size_t)g->L * g->n, cnt32 = (size_t)g->L * G * kFastNb;
    if (cnt == 0) return 0;
    if (to_fast) {
        if (g->ring32_cap < cnt32) {
            dfree(g->ring32);
            int rc = dalloc(&g->ring32, cnt32);
            if (rc) return rc;
            g->ring32_cap = cnt32;
        }
        hipLaunchKernelGGL(k_ring_f64_to_f32, dim3((unsigned)((cnt32 + 255) / 256)), dim3(256), 0, g->stream, g->L,
                           g->n, G, g->ring, g->ring32);
    } else if (g->ring32) {
        hipLaunchKernelGGL(k_ring_f32_to_f64, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, g->stream, g->L,
                           g->n, G, g->ring32, g->ring);
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
    if (g->xch_words < 2 * nw * kXchStride || g->xch_plan_G != -G) {
        if (g->xch_words < 2 * nw * kXchStride) {
            dfree(g->xch);
            int rc = dalloc(&g->xch, 2 * nw * kXchStride);
            if (rc) return rc;
            g->xch_words = 2 * nw * kXchStride;
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
    if (!g->fseg) {   // row-pull segments, built once per reservoir (the CSR does not change)
        if (g->rp_host[g->n] > (int64_t)INT32_MAX) { set_error("fast step kernel: more than 2^31 - 1 synapses"); return SPIRES_DEV_EINVAL; }
        std::vector<int> off(G + 1, 0), rowf((size_t)G * (nb + 1), 0), rng;
        int smax = 0;
        for (int w = 0; w < G; ++w) {
            int cnt = 0;
            for (int r = 0; r < nb; ++r) {
                rowf[(size_t)w * (nb + 1) + r] = cnt;
                const int64_t row = w + (int64_t)G * r;   // neuron of lane r
                if (row >= g->n) continue;
                const int64_t b = g->rp_host[row], e = g->rp_host[row + 1];
                for (int64_t q = b; q < e; q += kSegEdges) {
                    rng.push_back((int)q);
                    rng.push_back((int)std::min<int64_t>(e, q + kSegEdges));
                    ++cnt;
                }
            }
            rowf[(size_t)w * (nb + 1) + nb] = cnt;
            off[w + 1] = off[w] + cnt;
            smax = std::max(smax, cnt);
        }
        std::vector<int> tab;
        tab.reserve(off.size() + rowf.size() + rng.size());
        tab.insert(tab.end(), off.begin(), off.end());
        tab.insert(tab.end(), rowf.begin(), rowf.end());
        tab.insert(tab.end(), rng.begin(), rng.end());
        int rc = dalloc(&g->fseg, tab.size());
        if (rc) return rc;
        SDEV_CK(hipMemcpyAsync(g->fseg, tab.data(), tab.size() * 4, hipMemcpyHostToDevice, g->stream));
        SDEV_CK(hipStreamSynchronize(g->stream));
        g->fseg_max = std::max(1, smax);
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
    a.ring = g->ring32; a.ldr = G * kFastNb; a.coef = g->coef;
    a.xch = g->xch; a.abort_flag = g->abort_flag;
    a.rec_V = rec_V; a.rec_S = rec_S;
    a.seg = g->fseg; a.seg_max = g->fseg_max;
    // dynamic LDS of k_fast (the layout at the top of the kernel) for ne CSR slots
    auto lds_bytes = [&](size_t ne) {
        const size_t f32 = (size_t)((nw + 3) & ~3) + (size_t)((g->L + 3) & ~3) + (size_t)((K * kOStride + 3) & ~3) +
                           (size_t)K * nb;
        const size_t f64 = (size_t)((g->fseg_max + 8 + 1) & ~1) + 2 * (size_t)nb + ne;
        const size_t i32 = ne + 2 * (size_t)g->fseg_max + nb + 1;
        return f32 * 4 + f64 * 8 + i32 * 4;
    };
    const size_t cap = 160 * 1024 - 2048;
    const int me = kSegEdges * g->fseg_max;   // CSR slots, kSegEdges per segment
    const int ne_lds = lds_bytes(me) <= cap ? me : 0;
    const size_t lds = lds_bytes(ne_lds);
    if (lds > 160 * 1024) { set_error("fast step kernel: history length and segment tables exceed the LDS"); return SPIRES_DEV_EINVAL; }
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
