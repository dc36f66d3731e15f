// Reservoir construction on the device.
//
// Erdos-Renyi recurrent weights with the reference's distribution
// (init_weights + generate_weight, reservoir.c:345-545): edge i->j, i != j,
// when the uniform u1(i, j) in [0, 1) is below p; weight magnitude u2 in
// (0, 1], positive when u3 < ei_ratio.  u1..u3 come from SplitMix64 of
// (seed, i, j, stream) instead of the reference's rand() sequence, so any row
// can be generated independently and in ascending column order.  One wave
// per row counts, the host turns the counts into row pointers, one wave per
// row fills columns and values.
//
// Spectral radius: the reference's power iteration (sparse.c:93-127) -- x = 1,
// y = W x with each row summed in CSR order, lambda = max |y|, x = y / lambda,
// stop when |lambda - lambda_prev| < 1e-6 or after 1000 iterations -- one
// launch per iteration (rows are owned by threads, the max is a reduction over
// workgroup maxima), so every operation and its order match the CPU loop and
// lambda is bit-identical.
// The rescale W *= rho / lambda (when lambda > 1e-9) follows reservoir.c:548-557
// and runs on the device too, so only the finished CSR crosses PCIe.
//
// Small-world and scale-free topologies (reservoir.c:399-443, 445-536) follow
// the reference's construction in law, from hashed draws instead of rand():
// - small world: the directed ring of K/2 forward neighbours, then each ring
//   edge (i, i+s), s = 1..K/2 in order, is rewired with probability 0.1 to the
//   first of a sequence of uniform targets that is neither i nor already an
//   edge of row i (10 n attempts, as the reference).  Rows are independent:
//   one wave per row keeps the row's edge set as a bitmap in LDS and tests 64
//   candidate targets at once (the first valid one in attempt order wins).
// - scale free: the seed clique on m0 nodes, then node v = m0..n-1 takes m
//   distinct partners among 0..v-1 with probability proportional to deg + 1
//   (attempts past 50 m draw uniformly, as the reference's fallback), and each
//   undirected edge is oriented by a fair coin.  A draw proportional to
//   deg + 1 is a uniform node with probability v / total, otherwise a uniform
//   endpoint of the edges among 0..v-1 (and of v's own edges so far, which are
//   rejected as duplicates, as the reference's roulette picks and rejects
//   them).  The endpoint of a seed edge or the source of a grown edge is known
//   at once; the target of a grown edge is another node's draw, so the nodes
//   run in rounds: a node that meets an undecided target stops and resumes
//   with the same draw in the next round.  Each round at least completes the
//   lowest unfinished node; in practice a few tens of rounds suffice.
// Edge weights keep generate_weight's law, hashed from (seed, row, column).
#include "dev_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace sdev {

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// 53 random bits as a double: unit0 in [0, 1), unit1 in (0, 1]
__device__ __forceinline__ double unit0(unsigned long long h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }
__device__ __forceinline__ double unit1(unsigned long long h) {
    return ((double)(h >> 11) + 1.0) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ unsigned long long pair_key(unsigned long long seed, long long n, int i, int j, int stream) {
    return mix64(seed ^ mix64(((unsigned long long)((long long)i * n + j) << 2) | (unsigned)stream));
}

// One wave per row: lane l tests columns l, l+64, ...; a ballot gives the
// chunk's edges, so the fill writes them in ascending column order.
constexpr int kErRows = 4;   // rows (waves) per workgroup

__device__ __forceinline__ bool er_edge(unsigned long long seed, int n, double p, int i, int j) {
    return j < n && j != i && unit0(pair_key(seed, n, i, j, 0)) < p;
}

__global__ void __launch_bounds__(64 * kErRows) k_er_count(int n, double p, unsigned long long seed,
                                                           int* __restrict__ cnt) {
    const int i = blockIdx.x * kErRows + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= n) return;
    int c = 0;
    for (int j0 = 0; j0 < n; j0 += 64) c += __popcll(__ballot(er_edge(seed, n, p, i, j0 + lane)));
    if (lane == 0) cnt[i] = c;
}

__global__ void __launch_bounds__(64 * kErRows) k_er_fill(int n, double p, double ei, unsigned long long seed,
                                                          const int64_t* __restrict__ rp, int32_t* __restrict__ col,
                                                          double* __restrict__ val) {
    const int i = blockIdx.x * kErRows + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= n) return;
    int64_t k = rp[i];
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int j0 = 0; j0 < n; j0 += 64) {
        const int j = j0 + lane;
        const bool e = er_edge(seed, n, p, i, j);
        const unsigned long long m = __ballot(e);
        if (e) {
            const double mag = unit1(pair_key(seed, n, i, j, 1)), u3 = unit0(pair_key(seed, n, i, j, 2));
            const int64_t at = k + __popcll(m & below);
            col[at] = j;
            val[at] = u3 < ei ? mag : -mag;
        }
        k += __popcll(m);
    }
}

// Power iteration, one launch per iteration `it` (0..1000) on one stream.
// Launch `it` takes lambda_{it-1} = max of the previous launch's workgroup
// maxima, applies the reference's stop test to it, and otherwise builds
// x_it = y_{it-1} / lambda_{it-1} (x_0 = 1) in LDS -- every workgroup the whole
// vector, so no grid-wide barrier is needed -- and sums its rows of y_it = W x_it
// (one thread per row, CSR order).  y and the maxima alternate between two
// buffers by parity; lam_hist[it] keeps lambda_it for the next launch's test.
// Once the test has passed, `state[0]` is set and later launches return at once.
constexpr int kPowerRows = 256;          // rows (threads) per workgroup
constexpr int kPowerMaxN = 19968;        // x_it in LDS (156 KB) beside the 2 KB reduction array

struct PowerWs {
    double* y;         // [2][n]
    double* part;      // [2][groups]
    double* lam_hist;  // [1001]
    int* state;        // [0] done, [1] iterations run
    double* lam;       // result
};

__global__ void __launch_bounds__(kPowerRows) k_power_step(int it, int n, const int64_t* __restrict__ rp,
                                                           const int32_t* __restrict__ col,
                                                           const double* __restrict__ val, PowerWs w) {
    extern __shared__ double x_s[];
    __shared__ double red[kPowerRows];
    const int tid = threadIdx.x, groups = gridDim.x;
    if (w.state[0]) return;
    if (it > 0) {
        const double* pp = w.part + (size_t)((it - 1) & 1) * groups;
        double m = 0.0;
        for (int g = tid; g < groups; g += kPowerRows)
            if (pp[g] > m) m = pp[g];
        red[tid] = m;
        __syncthreads();
        for (int s = kPowerRows / 2; s > 0; s >>= 1) {
            if (tid < s && red[tid + s] > red[tid]) red[tid] = red[tid + s];
            __syncthreads();
        }
        const double ln = red[0], lo = it > 1 ? w.lam_hist[it - 2] : 0.0;
        const bool stop = fabs(ln - lo) < 1e-6 || it == 1000;
        if (blockIdx.x == 0 && tid == 0) {
            w.lam_hist[it - 1] = ln;
            if (stop) { w.state[0] = 1; w.state[1] = it; *w.lam = ln; }
        }
        if (stop) return;
        const double* yp = w.y + (size_t)((it - 1) & 1) * n;
        for (int j = tid; j < n; j += kPowerRows) x_s[j] = yp[j] / ln;
    } else {
        for (int j = tid; j < n; j += kPowerRows) x_s[j] = 1.0;
    }
    __syncthreads();
    const int i = blockIdx.x * kPowerRows + tid;
    double a = 0.0;
    if (i < n) {
        double s = 0.0;
        for (int64_t k = rp[i]; k < rp[i + 1]; ++k) s = s + val[k] * x_s[col[k]];
        w.y[(size_t)(it & 1) * n + i] = s;
        a = fabs(s);
    }
    red[tid] = a;
    __syncthreads();
    for (int s = kPowerRows / 2; s > 0; s >>= 1) {
        if (tid < s && red[tid + s] > red[tid]) red[tid] = red[tid + s];
        __syncthreads();
    }
    if (tid == 0) w.part[(size_t)(it & 1) * groups + blockIdx.x] = red[0];
}

// Doubles of power-iteration work buffers for n rows: y, part, lam_hist, lam, state.
constexpr size_t power_ws_doubles(int n) {
    return 2 * (size_t)n + 2 * (size_t)((n + kPowerRows - 1) / kPowerRows) + 1001 + 1 + 1;
}

// Carves the work buffers out of ws (power_ws_doubles(n) doubles).
void power_ws_layout(double* ws, int n, PowerWs* w) {
    const int groups = (n + kPowerRows - 1) / kPowerRows;
    w->y = ws;
    w->part = w->y + 2 * (size_t)n;
    w->lam_hist = w->part + 2 * (size_t)groups;
    w->lam = w->lam_hist + 1001;
    w->state = (int*)(w->lam + 1);
}

// The 1001 iteration launches (w.state must be zero when the first one runs).
int power_iter_launch(hipStream_t st, int n, const int64_t* rp, const int32_t* col, const double* val,
                      const PowerWs& w) {
    const size_t lds = (size_t)n * sizeof(double);
    const int groups = (n + kPowerRows - 1) / kPowerRows;
    SDEV_CK(hipFuncSetAttribute((const void*)k_power_step, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    for (int it = 0; it <= 1000; ++it)
        hipLaunchKernelGGL(k_power_step, dim3(groups), dim3(kPowerRows), lds, st, it, n, rp, col, val, w);
    SDEV_CK(hipGetLastError());
    return 0;
}

// Enqueue the power iteration of the device CSR (rp, col, val) on st; *w.lam
// and w.state[1] hold lambda and the iteration count once st drains.  The
// workspace comes from one allocation released by the caller after the sync.
int power_iter_enqueue(hipStream_t st, int n, const int64_t* rp, const int32_t* col, const double* val,
                       double** ws_out, PowerWs* w) {
    if (n > kPowerMaxN) {
        set_error("power iteration: n above 19968 does not fit x in LDS");
        return SPIRES_DEV_EINVAL;
    }
    double* ws = nullptr;
    int rc = dalloc(&ws, power_ws_doubles(n));
    if (rc) return rc;
    *ws_out = ws;
    power_ws_layout(ws, n, w);
    SDEV_CK(hipMemsetAsync(w->state, 0, 2 * sizeof(int), st));
    return power_iter_launch(st, n, rp, col, val, *w);
}

// ---- Power iteration without allocation ------------------------------------
// The work buffers of spires_dev_spectral_radius_noalloc live in one fixed
// array per device, sized for the largest n the kernel takes (kPowerMaxN:
// 2 x 19968 + 2 x 78 + 1003 doubles = 328,856 bytes); each device holds its
// copy until the process ends.  The host never passes `scratch` itself to a
// kernel: its device address comes from hipGetSymbolAddress, once per device.
constexpr size_t kScratchDoubles = power_ws_doubles(kPowerMaxN);
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of the current device's scratch array (looked up on the
// first call per device, kept afterwards).
static int scratch_address(double** out) {
    int dev = 0;
    SDEV_CK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxScratchDevices) {
        set_error("power iteration scratch: device index out of range");
        return SPIRES_DEV_EINVAL;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            set_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
            return SPIRES_DEV_EINTERNAL;
        }
        g_scratch_ptr[dev] = (double*)p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

// Clears the two state words (a kernel, not a memset, so the call can be captured in a graph).
__global__ void k_power_clear(int* __restrict__ state) {
    if (threadIdx.x < 2) state[threadIdx.x] = 0;
}

// Copies the iteration count out of the state words.
__global__ void k_power_iters_out(const int* __restrict__ state, int* __restrict__ iters) {
    if (threadIdx.x == 0) *iters = state[1];
}

__global__ void k_rescale(size_t nnz, double rho, const double* __restrict__ lam, double* __restrict__ val) {
    const double l = *lam;
    if (!(l > 1e-9)) return;
    const double f = rho / l;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += (size_t)gridDim.x * blockDim.x)
        val[k] = val[k] * f;
}

// Draw key for sequential draws: tag (4 bits), row or node (15 bits), ring
// offset or edge number (15 bits), attempt (30 bits).
__device__ __forceinline__ unsigned long long draw_key(unsigned long long seed, unsigned tag, int i, int s,
                                                       long long a) {
    return mix64(seed ^ mix64(((unsigned long long)tag << 60) | ((unsigned long long)(unsigned)i << 45) |
                              ((unsigned long long)(unsigned)s << 30) | (unsigned long long)a));
}

__device__ __forceinline__ double edge_weight(unsigned long long seed, int n, double ei, int i, int j) {
    const double mag = unit1(pair_key(seed, n, i, j, 1)), u3 = unit0(pair_key(seed, n, i, j, 2));
    return u3 < ei ? mag : -mag;
}

// ---- small world --------------------------------------------------------------
constexpr int kSwRows = 4;   // rows (waves) per workgroup; the bitmap words of each in LDS

// Row i's final edge set into the wave's bitmap bm (nw32 words).  LDS accesses
// go through a volatile pointer (program order), with a fence after each
// phase so that every lane sees the bits the others set.
__device__ void sw_row(volatile unsigned* bm, int nw32, int n, int half, unsigned long long seed, int i, int lane) {
    for (int w = lane; w < nw32; w += 64) bm[w] = 0u;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    for (int s = 1 + lane; s <= half; s += 64) {   // ring: i -> i + s (LDS atomics: lanes share words)
        const int j = (int)(((long long)i + s) % n);
        if (j != i) atomicOr((unsigned*)(bm + (j >> 5)), 1u << (j & 31));
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    const long long limit = 10ll * n;
    for (int s = 1; s <= half; ++s) {
        if (!(unit0(draw_key(seed, 1, i, s, 0)) < 0.1)) continue;
        const int jo = (int)(((long long)i + s) % n);
        if (lane == 0) bm[jo >> 5] = bm[jo >> 5] & ~(1u << (jo & 31));
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        int jn = -1;
        for (long long a0 = 0; a0 <= limit; a0 += 64) {
            const long long a = a0 + lane;   // attempt a + 1 of the reference's loop
            int c = (int)(unit0(draw_key(seed, 2, i, s, a)) * (double)n);
            if (c >= n) c = n - 1;
            const bool ok = a < limit && c != i && !((bm[c >> 5] >> (c & 31)) & 1u);
            const unsigned long long bal = __ballot(ok);
            if (bal) { jn = __shfl(c, __builtin_ctzll(bal)); break; }
            if (a0 + 64 > limit) {   // no valid target in 10 n attempts: the reference keeps draw 10 n + 1
                jn = __shfl(c, (int)(limit - a0));
                break;
            }
        }
        if (lane == 0 && jn >= 0 && jn != i) bm[jn >> 5] = bm[jn >> 5] | (1u << (jn & 31));
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    }
}

// Count (col == nullptr) or fill row i's edges in ascending column order.
__global__ void __launch_bounds__(64 * kSwRows) k_sw_rows(int n, int half, double ei, unsigned long long seed,
                                                          int* __restrict__ cnt, const int64_t* __restrict__ rp,
                                                          int32_t* __restrict__ col, double* __restrict__ val) {
    extern __shared__ unsigned sw_bm[];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.x * kSwRows + wv;
    if (i >= n) return;
    const int nw32 = (n + 31) >> 5;
    volatile unsigned* bm = sw_bm + (size_t)wv * nw32;
    sw_row(bm, nw32, n, half, seed, i, lane);
    // lane l owns words [l*per, (l+1)*per): ascending columns in lane order
    const int per = (nw32 + 63) / 64;
    const int w0 = min(nw32, lane * per), w1 = min(nw32, w0 + per);
    int c = 0;
    for (int w = w0; w < w1; ++w) c += __popc(bm[w]);
    int incl = c;   // inclusive prefix over lanes
    for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_up(incl, d);
        if (lane >= d) incl += o;
    }
    if (!col) {
        if (lane == 63) cnt[i] = incl;
        return;
    }
    int64_t k = rp[i] + incl - c;
    for (int w = w0; w < w1; ++w) {
        unsigned b = bm[w];
        while (b) {
            const int j = (w << 5) + __builtin_ctz(b);
            b &= b - 1;
            col[k] = j;
            val[k] = edge_weight(seed, n, ei, i, j);
            ++k;
        }
    }
}

// ---- scale free ----------------------------------------------------------------
struct BaArgs {
    int n, m, m0;
    long long Eseed, E;
    unsigned long long seed;
    int* src;                  // [E] first endpoint: the seed pair's lower node, or the growing node
    int* tgt;                  // [E] second endpoint (-1 until decided)
    int* k_done;               // [n] edges node v has decided
    long long* tries;          // [n] attempts node v has made (the reference's guard, then the fallback's)
    int* unfinished;           // nodes that stopped in this round
};

__global__ void k_ba_init(BaArgs b) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < b.E) {
        if (e < b.Eseed) {   // seed clique pairs (u, v), u < v, in lexicographic order
            int u = 0;
            long long rem = e;
            while (rem >= b.m0 - 1 - u) { rem -= b.m0 - 1 - u; ++u; }
            b.src[e] = u;
            b.tgt[e] = u + 1 + (int)rem;
        } else {
            b.src[e] = b.m0 + (int)((e - b.Eseed) / b.m);
            b.tgt[e] = -1;
        }
    }
    if (e < b.n) { b.k_done[e] = 0; b.tries[e] = 0; }
}

__global__ void k_ba_round(BaArgs b) {
    const int v = b.m0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (v >= b.n) return;
    int k = b.k_done[v];
    if (k >= b.m) return;
    long long t = b.tries[v];
    const long long Eb = b.Eseed + (long long)(v - b.m0) * b.m;   // edges among 0..v-1
    const long long guard = 50ll * b.m;
    while (k < b.m) {
        int u;
        if (t < guard) {   // proportional to deg + 1 over 0..v-1
            const double total = 2.0 * (double)Eb + (double)k + (double)v;
            const double r = unit0(draw_key(b.seed, 3, v, 0, t)) * total;
            if (r < (double)v) {
                u = min((int)r, v - 1);
            } else {
                long long q = (long long)(r - (double)v);
                if (q > 2 * Eb + k - 1) q = 2 * Eb + k - 1;
                if (q < 2 * Eb) {
                    const long long e = q >> 1;
                    if (q & 1) {
                        u = __hip_atomic_load(&b.tgt[e], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        if (u < 0) {   // another node's draw not decided yet: resume here next round
                            b.k_done[v] = k;
                            b.tries[v] = t;
                            atomicAdd(b.unfinished, 1);
                            return;
                        }
                    } else {
                        u = b.src[e];
                    }
                } else {
                    u = b.tgt[Eb + (q - 2 * Eb)];   // one of v's own partners: rejected below
                }
            }
        } else {   // fallback: uniform partners
            const double r = unit0(draw_key(b.seed, 4, v, 0, t - guard)) * (double)v;
            u = min((int)r, v - 1);
        }
        ++t;
        bool dup = false;
        for (int q = 0; q < k; ++q) dup = dup || b.tgt[Eb + q] == u;
        if (!dup) {
            __hip_atomic_store(&b.tgt[Eb + k], u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ++k;
        }
    }
    b.k_done[v] = k;
    b.tries[v] = t;
}

// Orientation by a fair coin per undirected edge (lower -> higher when the coin
// is below 0.5, as the reference's i < j loop), row counts.
__global__ void k_ba_orient(BaArgs b, int* __restrict__ row_of, int* __restrict__ col_of, int* __restrict__ cnt) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= b.E) return;
    const int x = b.src[e], y = b.tgt[e];
    const int lo = min(x, y), hi = max(x, y);
    const bool fwd = unit0(pair_key(b.seed, b.n, lo, hi, 3)) < 0.5;
    const int r = fwd ? lo : hi, c = fwd ? hi : lo;
    row_of[e] = r;
    col_of[e] = c;
    atomicAdd(&cnt[r], 1);
}

__global__ void k_ba_scatter(long long E, const int* __restrict__ row_of, const int* __restrict__ col_of,
                             const int64_t* __restrict__ rp, int* __restrict__ fill, int32_t* __restrict__ col) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const int r = row_of[e];
    col[rp[r] + atomicAdd(&fill[r], 1)] = col_of[e];
}

// Each row's columns in ascending order (insertion sort; the order of the
// scatter is not deterministic, the sorted row is), then the weights.
__global__ void k_rows_sort_weigh(int n, double ei, unsigned long long seed, const int64_t* __restrict__ rp,
                                  int32_t* __restrict__ col, double* __restrict__ val) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t b = rp[i], e = rp[i + 1];
    for (int64_t x = b + 1; x < e; ++x) {
        const int32_t c = col[x];
        int64_t y = x - 1;
        while (y >= b && col[y] > c) { col[y + 1] = col[y]; --y; }
        col[y + 1] = c;
    }
    for (int64_t x = b; x < e; ++x) val[x] = edge_weight(seed, n, ei, i, col[x]);
}

// Device buffers of one construction and its host outputs; the common tail runs
// the power iteration on the device CSR, rescales it and copies it out.
struct Build {
    hipStream_t st = nullptr;
    int n = 0;
    int64_t* rp = nullptr;
    int32_t* col = nullptr;
    double* val = nullptr;
    double* ws = nullptr;
    std::vector<void*> extra;   // construction-specific device buffers
    int64_t* orp = nullptr;     // host row pointers (malloc'd, handed to the caller)
    int32_t* ocol = nullptr;
    double* oval = nullptr;
    ~Build() {
        if (st) (void)hipStreamSynchronize(st);
        dfree(rp); dfree(col); dfree(val); dfree(ws);
        for (void* q : extra) if (q) (void)hipFree(q);
        if (st) (void)hipStreamDestroy(st);
        free(orp); free(ocol); free(oval);
    }
};

// Row pointers from per-row counts (device), then host and device copies of them
// and room for the columns and values.
static int build_rows(Build& b, const int* cnt) {
    std::vector<int> hc(b.n);
    hipError_t e = hipMemcpyAsync(hc.data(), cnt, (size_t)b.n * sizeof(int), hipMemcpyDeviceToHost, b.st);
    if (e == hipSuccess) e = hipStreamSynchronize(b.st);
    if (e != hipSuccess) return hip_fail(e, "construction row counts", __LINE__);
    b.orp = (int64_t*)malloc(((size_t)b.n + 1) * 8);
    if (!b.orp) return SPIRES_DEV_EALLOC;
    b.orp[0] = 0;
    for (int i = 0; i < b.n; ++i) b.orp[i + 1] = b.orp[i] + hc[i];
    const size_t nnz = (size_t)b.orp[b.n];
    b.ocol = (int32_t*)malloc((nnz ? nnz : 1) * 4);
    b.oval = (double*)malloc((nnz ? nnz : 1) * 8);
    if (!b.ocol || !b.oval) return SPIRES_DEV_EALLOC;
    int rc = 0;
    if ((rc = dalloc(&b.rp, (size_t)b.n + 1)) || (rc = dalloc(&b.col, nnz)) || (rc = dalloc(&b.val, nnz))) return rc;
    SDEV_CK(hipMemcpyAsync(b.rp, b.orp, ((size_t)b.n + 1) * 8, hipMemcpyHostToDevice, b.st));
    return 0;
}

// Power iteration and rescale of the filled device CSR, then the copies out.
static int build_finish(Build& b, double rho, int64_t** rp_out, int32_t** col_out, double** val_out,
                        size_t* nnz_out, double* lam_out) {
    const size_t nnz = (size_t)b.orp[b.n];
    PowerWs w;
    int rc = power_iter_enqueue(b.st, b.n, b.rp, b.col, b.val, &b.ws, &w);
    if (rc) return rc;
    if (nnz) {
        const unsigned rb = (unsigned)std::min<size_t>((nnz + 255) / 256, 4096);
        hipLaunchKernelGGL(k_rescale, dim3(rb), dim3(256), 0, b.st, nnz, rho, w.lam, b.val);
    }
    hipError_t e = hipGetLastError();
    double lam = 0.0;
    if (e == hipSuccess && nnz) e = hipMemcpyAsync(b.ocol, b.col, nnz * 4, hipMemcpyDeviceToHost, b.st);
    if (e == hipSuccess && nnz) e = hipMemcpyAsync(b.oval, b.val, nnz * 8, hipMemcpyDeviceToHost, b.st);
    if (e == hipSuccess) e = hipMemcpyAsync(&lam, w.lam, 8, hipMemcpyDeviceToHost, b.st);
    if (e == hipSuccess) e = hipStreamSynchronize(b.st);
    if (e != hipSuccess) return hip_fail(e, "construction", __LINE__);
    *rp_out = b.orp; *col_out = b.ocol; *val_out = b.oval; *nnz_out = nnz;
    if (lam_out) *lam_out = lam;
    b.orp = nullptr; b.ocol = nullptr; b.oval = nullptr;
    return 0;
}

template <typename T>
static int build_alloc(Build& b, T** p, size_t count) {
    int rc = dalloc(p, count);
    if (!rc) b.extra.push_back((void*)*p);
    return rc;
}

static int build_er(Build& b, double p, double ei, unsigned long long seed) {
    int* cnt = nullptr;
    int rc = build_alloc(b, &cnt, (size_t)b.n);
    if (rc) return rc;
    const unsigned blocks = (unsigned)((b.n + kErRows - 1) / kErRows);
    hipLaunchKernelGGL(k_er_count, dim3(blocks), dim3(64 * kErRows), 0, b.st, b.n, p, seed, cnt);
    if ((rc = build_rows(b, cnt))) return rc;
    hipLaunchKernelGGL(k_er_fill, dim3(blocks), dim3(64 * kErRows), 0, b.st, b.n, p, ei, seed, b.rp, b.col, b.val);
    SDEV_CK(hipGetLastError());
    return 0;
}

// Watts-Strogatz sizes as reservoir.c:404-413: K = round(conn (n - 1)), at
// least 2, even, below n; K / 2 forward ring neighbours.  n < 3: no edges.
static int build_sw(Build& b, double conn, double ei, unsigned long long seed) {
    int* cnt = nullptr;
    int rc = build_alloc(b, &cnt, (size_t)b.n);
    if (rc) return rc;
    const int n = b.n;
    int half = 0;
    if (n >= 3) {
        int K = (int)(conn * (double)(n - 1) + 0.5);
        if (K < 2) K = 2;
        if (K % 2) K++;
        if (K >= n) K = n - 1;
        half = K / 2;
    }
    const unsigned blocks = (unsigned)((n + kSwRows - 1) / kSwRows);
    const size_t lds = (size_t)kSwRows * ((n + 31) / 32) * 4;
    SDEV_CK(hipFuncSetAttribute((const void*)k_sw_rows, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_sw_rows, dim3(blocks), dim3(64 * kSwRows), lds, b.st, n, half, ei, seed, cnt,
                       (const int64_t*)nullptr, (int32_t*)nullptr, (double*)nullptr);
    SDEV_CK(hipGetLastError());
    if ((rc = build_rows(b, cnt))) return rc;
    hipLaunchKernelGGL(k_sw_rows, dim3(blocks), dim3(64 * kSwRows), lds, b.st, n, half, ei, seed, cnt,
                       (const int64_t*)b.rp, b.col, b.val);
    SDEV_CK(hipGetLastError());
    return 0;
}

// Barabasi-Albert sizes as reservoir.c:451-463: m = round(conn (n - 1)) in
// [1, n - 1], seed clique m0 = max(m, 2) (below n).  n < 2: no edges.
static int build_ba(Build& b, double conn, double ei, unsigned long long seed) {
    const int n = b.n;
    int* cnt = nullptr;
    int rc = build_alloc(b, &cnt, (size_t)n);
    if (rc) return rc;
    SDEV_CK(hipMemsetAsync(cnt, 0, (size_t)n * 4, b.st));
    if (n < 2) return build_rows(b, cnt);
    int m = (int)(conn * (double)(n - 1) + 0.5);
    if (m < 1) m = 1;
    if (m >= n) m = n - 1;
    int m0 = m > 2 ? m : 2;
    if (m0 >= n) m0 = n - 1;
    BaArgs a;
    a.n = n; a.m = m; a.m0 = m0; a.seed = seed;
    a.Eseed = (long long)m0 * (m0 - 1) / 2;
    a.E = a.Eseed + (long long)(n - m0) * m;
    int *row_of = nullptr, *col_of = nullptr, *fill = nullptr;
    if ((rc = build_alloc(b, &a.src, (size_t)a.E)) || (rc = build_alloc(b, &a.tgt, (size_t)a.E)) ||
        (rc = build_alloc(b, &a.k_done, (size_t)n)) || (rc = build_alloc(b, &a.tries, (size_t)n)) ||
        (rc = build_alloc(b, &a.unfinished, 1)) || (rc = build_alloc(b, &row_of, (size_t)a.E)) ||
        (rc = build_alloc(b, &col_of, (size_t)a.E)) || (rc = build_alloc(b, &fill, (size_t)n)))
        return rc;
    const long long span = std::max<long long>(a.E, n);
    hipLaunchKernelGGL(k_ba_init, dim3((unsigned)((span + 255) / 256)), dim3(256), 0, b.st, a);
    const unsigned gblocks = (unsigned)((n - m0 + 255) / 256);
    // every round completes at least the lowest unfinished node, so n - m0 rounds bound the loop
    int left = 1;
    for (int round = 0; round <= n - m0 && left; ++round) {
        SDEV_CK(hipMemsetAsync(a.unfinished, 0, 4, b.st));
        if (gblocks) hipLaunchKernelGGL(k_ba_round, dim3(gblocks), dim3(256), 0, b.st, a);
        SDEV_CK(hipMemcpyAsync(&left, a.unfinished, 4, hipMemcpyDeviceToHost, b.st));
        SDEV_CK(hipStreamSynchronize(b.st));
    }
    if (left) { set_error("scale-free construction did not converge"); return SPIRES_DEV_EINTERNAL; }
    const unsigned eblocks = (unsigned)((a.E + 255) / 256);
    hipLaunchKernelGGL(k_ba_orient, dim3(eblocks), dim3(256), 0, b.st, a, row_of, col_of, cnt);
    SDEV_CK(hipGetLastError());
    if ((rc = build_rows(b, cnt))) return rc;
    SDEV_CK(hipMemsetAsync(fill, 0, (size_t)n * 4, b.st));
    hipLaunchKernelGGL(k_ba_scatter, dim3(eblocks), dim3(256), 0, b.st, a.E, (const int*)row_of, (const int*)col_of,
                       (const int64_t*)b.rp, fill, b.col);
    hipLaunchKernelGGL(k_rows_sort_weigh, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, b.st, n, ei, seed,
                       (const int64_t*)b.rp, b.col, b.val);
    SDEV_CK(hipGetLastError());
    return 0;
}

}  // namespace sdev

using namespace sdev;

extern "C" {

int spires_dev_build_topology(int type, int n, double conn, double ei, unsigned long long seed, double rho,
                              int64_t** rp_out, int32_t** col_out, double** val_out, size_t* nnz_out,
                              double* lam_out) {
    sdev::RandShield rand_shield_;
    if (n < 1 || type < 0 || type > 2 || !(conn >= 0.0 && conn <= 1.0) || !(ei >= 0.0 && ei <= 1.0) || !rp_out ||
        !col_out || !val_out || !nnz_out)
        return SPIRES_DEV_EINVAL;
    *rp_out = nullptr; *col_out = nullptr; *val_out = nullptr; *nnz_out = 0;
    if (n > kPowerMaxN) {
        set_error("device construction: n above 19968 is not supported");
        return SPIRES_DEV_EINVAL;
    }
    Build b;
    b.n = n;
    SDEV_CK(hipStreamCreate(&b.st));
    int rc = type == 0 ? build_er(b, conn, ei, seed) : type == 1 ? build_sw(b, conn, ei, seed) : build_ba(b, conn, ei, seed);
    if (rc) return rc;
    return build_finish(b, rho, rp_out, col_out, val_out, nnz_out, lam_out);
}

int spires_dev_random_er(int n, double p, double ei, unsigned long long seed, double rho, int64_t** rp_out, int32_t** col_out,
                         double** val_out, size_t* nnz_out, double* lam_out) {
    return spires_dev_build_topology(0, n, p, ei, seed, rho, rp_out, col_out, val_out, nnz_out, lam_out);
}

int spires_dev_spectral_radius(int n, const int64_t* rp_h, const int32_t* col_h, const double* val_h, double* lam,
                               int* iters) {
    sdev::RandShield rand_shield_;
    if (n < 1 || !rp_h || !lam) return SPIRES_DEV_EINVAL;
    const size_t nnz = (size_t)rp_h[n];
    if (nnz && (!col_h || !val_h)) return SPIRES_DEV_EINVAL;
    if (rp_h[0] != 0) { set_error("spectral_radius: row_ptr[0] must be 0"); return SPIRES_DEV_EINVAL; }
    for (int i = 0; i < n; ++i)
        if (rp_h[i + 1] < rp_h[i]) { set_error("spectral_radius: row pointers not monotone"); return SPIRES_DEV_EINVAL; }
    for (size_t k = 0; k < nnz; ++k)
        if (col_h[k] < 0 || col_h[k] >= n) { set_error("spectral_radius: column out of range"); return SPIRES_DEV_EINVAL; }
    hipStream_t st = nullptr;
    SDEV_CK(hipStreamCreate(&st));
    int64_t* rp = nullptr;
    int32_t* col = nullptr;
    double *val = nullptr, *ws = nullptr;
    auto done = [&](int rc) {
        if (st) (void)hipStreamSynchronize(st);
        dfree(rp); dfree(col); dfree(val); dfree(ws);
        if (st) (void)hipStreamDestroy(st);
        return rc;
    };
    int rc = 0;
    if ((rc = dalloc(&rp, (size_t)n + 1)) || (rc = dalloc(&col, nnz)) || (rc = dalloc(&val, nnz))) return done(rc);
    hipError_t e = hipMemcpyAsync(rp, rp_h, ((size_t)n + 1) * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && nnz) e = hipMemcpyAsync(col, col_h, nnz * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && nnz) e = hipMemcpyAsync(val, val_h, nnz * 8, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return done(hip_fail(e, "power iteration upload", __LINE__));
    PowerWs w;
    if ((rc = power_iter_enqueue(st, n, rp, col, val, &ws, &w))) return done(rc);
    int hs[2] = {0, 0};
    e = hipMemcpyAsync(lam, w.lam, 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(hs, w.state, 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return done(hip_fail(e, "power iteration", __LINE__));
    if (iters) *iters = hs[1];
    return done(0);
}

int spires_dev_spectral_radius_noalloc(int n, const int64_t* rp, const int32_t* col, const double* val,
                                       double* lam, int* iters, void* stream) {
    sdev::RandShield rand_shield_;
    if (n < 1 || !rp || !col || !val || !lam || !iters) return SPIRES_DEV_EINVAL;
    if (n > kPowerMaxN) {
        set_error("spectral_radius_noalloc: n above 19968 exceeds the fixed work buffers");
        return SPIRES_DEV_EINVAL;
    }
    double* ws = nullptr;
    int rc = scratch_address(&ws);
    if (rc) return rc;
    PowerWs w;
    power_ws_layout(ws, n, &w);
    w.lam = lam;                         // the result goes straight to the caller's buffer
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_power_clear, dim3(1), dim3(64), 0, st, w.state);
    SDEV_CK(hipGetLastError());
    if ((rc = power_iter_launch(st, n, rp, col, val, w))) return rc;
    hipLaunchKernelGGL(k_power_iters_out, dim3(1), dim3(64), 0, st, (const int*)w.state, iters);
    SDEV_CK(hipGetLastError());
    return 0;
}

}  // extern "C"
