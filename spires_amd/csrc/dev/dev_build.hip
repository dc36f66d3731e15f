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

}  // namespace sdev

using namespace sdev;

extern "C" {

int spires_dev_random_er(int n, double p, double ei, unsigned long long seed, double rho, int64_t** rp_out, int32_t** col_out,
                         double** val_out, size_t* nnz_out, double* lam_out) {
    sdev::RandShield rand_shield_;
    if (n < 1 || !(p >= 0.0 && p <= 1.0) || !(ei >= 0.0 && ei <= 1.0) || !rp_out || !col_out || !val_out || !nnz_out) return SPIRES_DEV_EINVAL;
    *rp_out = nullptr; *col_out = nullptr; *val_out = nullptr; *nnz_out = 0;
    hipStream_t st = nullptr;
    SDEV_CK(hipStreamCreate(&st));
    int* cnt = nullptr;
    int64_t* rp = nullptr;
    int32_t* col = nullptr;
    double *val = nullptr, *ws = nullptr;
    int64_t* orp = nullptr;
    int32_t* ocol = nullptr;
    double* oval = nullptr;
    auto done = [&](int rc) {
        if (st) (void)hipStreamSynchronize(st);
        dfree(cnt); dfree(rp); dfree(col); dfree(val); dfree(ws);
        if (st) (void)hipStreamDestroy(st);
        if (rc) { free(orp); free(ocol); free(oval); }
        return rc;
    };
    int rc = 0;
    if (n > kPowerMaxN) {
        set_error("random_er: n above 19968 is not supported");
        return done(SPIRES_DEV_EINVAL);
    }
    if ((rc = dalloc(&cnt, (size_t)n))) return done(rc);
    const unsigned blocks = (unsigned)((n + kErRows - 1) / kErRows);
    hipLaunchKernelGGL(k_er_count, dim3(blocks), dim3(64 * kErRows), 0, st, n, p, seed, cnt);
    std::vector<int> hc(n);
    hipError_t e = hipMemcpyAsync(hc.data(), cnt, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return done(hip_fail(e, "er count", __LINE__));
    orp = (int64_t*)malloc(((size_t)n + 1) * 8);
    if (!orp) return done(SPIRES_DEV_EALLOC);
    orp[0] = 0;
    for (int i = 0; i < n; ++i) orp[i + 1] = orp[i] + hc[i];
    const size_t nnz = (size_t)orp[n];
    ocol = (int32_t*)malloc((nnz ? nnz : 1) * 4);
    oval = (double*)malloc((nnz ? nnz : 1) * 8);
    if (!ocol || !oval) return done(SPIRES_DEV_EALLOC);
    if ((rc = dalloc(&rp, (size_t)n + 1)) || (rc = dalloc(&col, nnz)) || (rc = dalloc(&val, nnz))) return done(rc);
    e = hipMemcpyAsync(rp, orp, ((size_t)n + 1) * 8, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return done(hip_fail(e, "er row pointers", __LINE__));
    hipLaunchKernelGGL(k_er_fill, dim3(blocks), dim3(64 * kErRows), 0, st, n, p, ei, seed, rp, col, val);
    PowerWs w;
    if ((rc = power_iter_enqueue(st, n, rp, col, val, &ws, &w))) return done(rc);
    if (nnz) {
        const unsigned rb = (unsigned)std::min<size_t>((nnz + 255) / 256, 4096);
        hipLaunchKernelGGL(k_rescale, dim3(rb), dim3(256), 0, st, nnz, rho, w.lam, val);
    }
    e = hipGetLastError();
    double lam = 0.0;
    if (e == hipSuccess && nnz) e = hipMemcpyAsync(ocol, col, nnz * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && nnz) e = hipMemcpyAsync(oval, val, nnz * 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(&lam, w.lam, 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return done(hip_fail(e, "er fill", __LINE__));
    *rp_out = orp; *col_out = ocol; *val_out = oval; *nnz_out = nnz;
    if (lam_out) *lam_out = lam;
    orp = nullptr; ocol = nullptr; oval = nullptr;
    return done(0);
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
