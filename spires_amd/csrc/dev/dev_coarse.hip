// Coarse-graining on the device (coarse_grain_reservoir, reservoir.c:782-967).
//
// The reference merges greedily: for i = 0..N-1 and j = i+1..N-1 in that order,
// a pair with W[i][j] > thr or W[j][i] > thr merges j into i (row and column
// i become the pairwise combination of rows i and j, row and column j are
// zeroed, V, W_in and W_out are combined), and the scan repeats until a whole
// pass merges nothing.  Each merge changes the rows that later checks read,
// so the order is kept: one workgroup runs a pass, row by row; what runs in
// parallel is the search for the next partner j of row i (1024 candidates per
// step, first hit by ballot and a workgroup minimum) and the merge itself,
// whose updates over k are independent of each other (no k reads what
// another k writes).  Every value is the same IEEE operation on the same
// operands as in the reference, so the result is bit-identical.
//
// Layout: the dense N x N matrix W and its transpose WT, row-major in HBM, so
// the partner search reads row i of both (coalesced); merges pay with strided
// column writes, and there are at most N-1 merges in all passes together.
// Rows of dead neurons are zero and are skipped (for thr >= 0 they can never
// merge, as in the reference).  A negative or NaN threshold never terminates
// in the reference (zero weights exceed it, so dead neurons keep merging) and
// is rejected.
#include "dev_internal.hpp"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace sdev {

constexpr int kCgThreads = 1024;

__global__ void k_cg_scatter(int n, const int64_t* __restrict__ rp, const int32_t* __restrict__ ci,
                             const double* __restrict__ val, double* __restrict__ W, double* __restrict__ WT) {
    const int i = blockIdx.x;
    for (int64_t k = rp[i] + threadIdx.x; k < rp[i + 1]; k += blockDim.x) {
        const int c = ci[k];
        W[(size_t)i * n + c] = val[k];
        WT[(size_t)c * n + i] = val[k];
    }
}

__device__ __forceinline__ double cg_combine(double a, double b) { return (a * b == 0.0) ? a + b : (a + b) / 2.0; }

__global__ void __launch_bounds__(kCgThreads) k_cg_pass(int n, double thr, double* __restrict__ W,
                                                         double* __restrict__ WT, int din, double* __restrict__ Wi,
                                                         int dout, double* __restrict__ Wo, double* __restrict__ V,
                                                         int* __restrict__ alive, int* __restrict__ merged_out) {
    __shared__ int wmin_s[kCgThreads / 64];
    __shared__ int found_s;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    int merged = 0;
    for (int i = 0; i < n; ++i) {
        if (!alive[i]) continue;   // a dead row and column are zero: no pair (i, j) can pass thr >= 0
        const double* wi = W + (size_t)i * n;    // W[i][j]
        const double* ti = WT + (size_t)i * n;   // W[j][i]
        int cur = i + 1;
        while (cur < n) {
            int found = n;
            for (int base = cur; base < n; base += kCgThreads) {
                const int j = base + tid;
                const bool hit = j < n && (wi[j] > thr || ti[j] > thr);
                const unsigned long long b = __ballot(hit);
                if (lane == 0) wmin_s[wv] = b ? base + wv * 64 + __ffsll((long long)b) - 1 : n;
                __syncthreads();
                if (tid == 0) {
                    int m = n;
                    for (int w = 0; w < kCgThreads / 64; ++w) m = min(m, wmin_s[w]);
                    found_s = m;
                }
                __syncthreads();
                found = found_s;
                __syncthreads();   // wmin_s and found_s are rewritten by the next step
                if (found < n) break;
            }
            if (found >= n) break;
            const int j = found;
            merged = 1;
            double* wj = W + (size_t)j * n;
            double* tj = WT + (size_t)j * n;
            double* wiw = W + (size_t)i * n;
            double* tiw = WT + (size_t)i * n;
            for (int k = tid; k < n; k += kCgThreads) {
                if (k == i || k == j) continue;
                const double nw = cg_combine(wiw[k], wj[k]);
                wiw[k] = nw;                      // W[i][k]
                W[(size_t)k * n + i] = nw;        // W[k][i]
                WT[(size_t)k * n + i] = nw;       // (W[i][k])^T
                tiw[k] = nw;                      // (W[k][i])^T
            }
            for (int k = tid; k < din; k += kCgThreads) {
                Wi[(size_t)i * din + k] = cg_combine(Wi[(size_t)i * din + k], Wi[(size_t)j * din + k]);
                Wi[(size_t)j * din + k] = 0.0;
            }
            for (int o = tid; o < dout; o += kCgThreads) {
                Wo[(size_t)o * n + i] += Wo[(size_t)o * n + j];
                Wo[(size_t)o * n + j] = 0.0;
            }
            if (tid == 0) {
                V[i] = (V[i] + V[j]) / 2.0;
                V[j] = 0.0;
                alive[j] = 0;
            }
            __syncthreads();   // row j read by the merge before it is cleared
            for (int k = tid; k < n; k += kCgThreads) {
                wj[k] = 0.0;
                W[(size_t)k * n + j] = 0.0;
                tj[k] = 0.0;
                WT[(size_t)k * n + j] = 0.0;
            }
            __syncthreads();   // the cleared row and column, and row i, before the next search
            cur = j + 1;
        }
    }
    if (tid == 0) *merged_out = merged;
}

// Nonzeros of each surviving row over surviving columns (one wave per row),
// then the compacted CSR in ascending column order.
constexpr int kCgRows = 4;
__global__ void __launch_bounds__(64 * kCgRows) k_cg_count(int n, const double* __restrict__ W,
                                                           const int* __restrict__ rows, int ns,
                                                           const int* __restrict__ map, int* __restrict__ cnt) {
    const int a = blockIdx.x * kCgRows + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (a >= ns) return;
    const double* w = W + (size_t)rows[a] * n;
    int c = 0;
    for (int j0 = 0; j0 < n; j0 += 64) {
        const int j = j0 + lane;
        c += __popcll(__ballot(j < n && map[j] >= 0 && w[j] != 0.0));
    }
    if (lane == 0) cnt[a] = c;
}

__global__ void __launch_bounds__(64 * kCgRows) k_cg_fill(int n, const double* __restrict__ W,
                                                          const int* __restrict__ rows, int ns,
                                                          const int* __restrict__ map, const int64_t* __restrict__ rp,
                                                          int32_t* __restrict__ ci, double* __restrict__ val) {
    const int a = blockIdx.x * kCgRows + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (a >= ns) return;
    const double* w = W + (size_t)rows[a] * n;
    int64_t k = rp[a];
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int j0 = 0; j0 < n; j0 += 64) {
        const int j = j0 + lane;
        const bool e = j < n && map[j] >= 0 && w[j] != 0.0;
        const unsigned long long m = __ballot(e);
        if (e) {
            const int64_t at = k + __popcll(m & below);
            ci[at] = map[j];
            val[at] = w[j];
        }
        k += __popcll(m);
    }
}

}  // namespace sdev

using namespace sdev;

extern "C" int spires_dev_coarse_grain(int n, const int64_t* rp_h, const int32_t* ci_h, const double* val_h,
                                       double thr, int din, double* W_in, int dout, double* W_out, double* V,
                                       int* alive_h, int* n_out, int64_t** rp_out, int32_t** ci_out,
                                       double** val_out, size_t* nnz_out, int* passes_out) {
    sdev::RandShield rand_shield_;
    if (n < 1 || !rp_h || !V || !alive_h || !n_out || !rp_out || !ci_out || !val_out || !nnz_out || din < 0 ||
        dout < 0 || (din && !W_in) || (dout && !W_out))
        return SPIRES_DEV_EINVAL;
    if (!(thr >= 0.0)) {
        set_error("coarse_grain: the threshold must be >= 0 (the reference's merge loop does not end otherwise)");
        return SPIRES_DEV_EINVAL;
    }
    const size_t nnz = (size_t)rp_h[n];
    if (rp_h[0] != 0 || (nnz && (!ci_h || !val_h))) return SPIRES_DEV_EINVAL;
    for (int i = 0; i < n; ++i)
        if (rp_h[i + 1] < rp_h[i]) { set_error("coarse_grain: row pointers not monotone"); return SPIRES_DEV_EINVAL; }
    for (size_t k = 0; k < nnz; ++k)
        if (ci_h[k] < 0 || ci_h[k] >= n) { set_error("coarse_grain: column out of range"); return SPIRES_DEV_EINVAL; }
    *rp_out = nullptr; *ci_out = nullptr; *val_out = nullptr; *nnz_out = 0; *n_out = 0;
    hipStream_t st = nullptr;
    SDEV_CK(hipStreamCreate(&st));
    const size_t nn = (size_t)n * n;
    double *W = nullptr, *WT = nullptr, *dWi = nullptr, *dWo = nullptr, *dV = nullptr, *dval = nullptr, *oval = nullptr;
    int *dalive = nullptr, *dmerged = nullptr, *drows = nullptr, *dmap = nullptr, *dcnt = nullptr;
    int64_t *drp = nullptr, *orp_d = nullptr;
    int32_t *dci = nullptr, *oci = nullptr;
    int64_t* hrp = nullptr;
    int32_t* hci = nullptr;
    double* hval = nullptr;
    auto done = [&](int rc) {
        if (st) (void)hipStreamSynchronize(st);
        dfree(W); dfree(WT); dfree(dWi); dfree(dWo); dfree(dV); dfree(dval); dfree(oval);
        dfree(dalive); dfree(dmerged); dfree(drows); dfree(dmap); dfree(dcnt); dfree(drp); dfree(orp_d);
        dfree(dci); dfree(oci);
        if (st) (void)hipStreamDestroy(st);
        if (rc) { std::free(hrp); std::free(hci); std::free(hval); }
        return rc;
    };
    int rc = 0;
    if ((rc = dalloc(&W, nn)) || (rc = dalloc(&WT, nn)) || (rc = dalloc(&dWi, (size_t)n * din)) ||
        (rc = dalloc(&dWo, (size_t)dout * n)) || (rc = dalloc(&dV, (size_t)n)) || (rc = dalloc(&dalive, (size_t)n)) ||
        (rc = dalloc(&dmerged, 1)) || (rc = dalloc(&drp, (size_t)n + 1)) || (rc = dalloc(&dci, nnz)) ||
        (rc = dalloc(&dval, nnz)))
        return done(rc);
    std::vector<int> ones(n, 1);
    hipError_t e = hipMemsetAsync(W, 0, nn * 8, st);
    if (e == hipSuccess) e = hipMemsetAsync(WT, 0, nn * 8, st);
    if (e == hipSuccess) e = hipMemcpyAsync(drp, rp_h, ((size_t)n + 1) * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && nnz) e = hipMemcpyAsync(dci, ci_h, nnz * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && nnz) e = hipMemcpyAsync(dval, val_h, nnz * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && din) e = hipMemcpyAsync(dWi, W_in, (size_t)n * din * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && dout) e = hipMemcpyAsync(dWo, W_out, (size_t)dout * n * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dV, V, (size_t)n * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dalive, ones.data(), (size_t)n * 4, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return done(hip_fail(e, "coarse_grain upload", __LINE__));
    hipLaunchKernelGGL(k_cg_scatter, dim3(n), dim3(64), 0, st, n, drp, dci, dval, W, WT);
    // at most n passes merge something (each merge removes a neuron for good), then one that does not
    int passes = 0, merged = 1;
    while (merged) {
        if (passes > n) { set_error("coarse_grain: pass limit exceeded"); return done(SPIRES_DEV_EINTERNAL); }
        hipLaunchKernelGGL(k_cg_pass, dim3(1), dim3(kCgThreads), 0, st, n, thr, W, WT, din, dWi, dout, dWo, dV,
                           dalive, dmerged);
        ++passes;
        e = hipMemcpyAsync(&merged, dmerged, 4, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return done(hip_fail(e, "coarse_grain pass", __LINE__));
    }
    e = hipMemcpyAsync(alive_h, dalive, (size_t)n * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(V, dV, (size_t)n * 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && din) e = hipMemcpyAsync(W_in, dWi, (size_t)n * din * 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && dout) e = hipMemcpyAsync(W_out, dWo, (size_t)dout * n * 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return done(hip_fail(e, "coarse_grain download", __LINE__));
    std::vector<int> map(n, -1), rows;
    for (int i = 0; i < n; ++i)
        if (alive_h[i]) { map[i] = (int)rows.size(); rows.push_back(i); }
    const int ns = (int)rows.size();
    if ((rc = dalloc(&drows, (size_t)ns)) || (rc = dalloc(&dmap, (size_t)n)) || (rc = dalloc(&dcnt, (size_t)ns)) ||
        (rc = dalloc(&orp_d, (size_t)ns + 1)))
        return done(rc);
    e = hipMemcpyAsync(drows, rows.data(), (size_t)ns * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dmap, map.data(), (size_t)n * 4, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return done(hip_fail(e, "coarse_grain map", __LINE__));
    const unsigned blocks = (unsigned)((ns + kCgRows - 1) / kCgRows);
    hipLaunchKernelGGL(k_cg_count, dim3(blocks), dim3(64 * kCgRows), 0, st, n, W, drows, ns, dmap, dcnt);
    std::vector<int> cnt(ns);
    e = hipMemcpyAsync(cnt.data(), dcnt, (size_t)ns * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return done(hip_fail(e, "coarse_grain count", __LINE__));
    hrp = (int64_t*)std::malloc(((size_t)ns + 1) * 8);
    if (!hrp) return done(SPIRES_DEV_EALLOC);
    hrp[0] = 0;
    for (int a = 0; a < ns; ++a) hrp[a + 1] = hrp[a] + cnt[a];
    const size_t onnz = (size_t)hrp[ns];
    hci = (int32_t*)std::malloc((onnz ? onnz : 1) * 4);
    hval = (double*)std::malloc((onnz ? onnz : 1) * 8);
    if (!hci || !hval) return done(SPIRES_DEV_EALLOC);
    if ((rc = dalloc(&oci, onnz)) || (rc = dalloc(&oval, onnz))) return done(rc);
    e = hipMemcpyAsync(orp_d, hrp, ((size_t)ns + 1) * 8, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return done(hip_fail(e, "coarse_grain rows", __LINE__));
    hipLaunchKernelGGL(k_cg_fill, dim3(blocks), dim3(64 * kCgRows), 0, st, n, W, drows, ns, dmap, orp_d, oci, oval);
    if (onnz) e = hipMemcpyAsync(hci, oci, onnz * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && onnz) e = hipMemcpyAsync(hval, oval, onnz * 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return done(hip_fail(e, "coarse_grain fill", __LINE__));
    *n_out = ns; *rp_out = hrp; *ci_out = hci; *val_out = hval; *nnz_out = onnz;
    if (passes_out) *passes_out = passes;
    hrp = nullptr; hci = nullptr; hval = nullptr;
    return done(0);
}
