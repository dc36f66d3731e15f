// Readout training on the device.
//
// Ridge (reservoir.c:573-679): the recorded state matrix X [T x N] stays in
// HBM.  A = X'X (+ lambda I) and B = X'Y are formed on the GPU (in fast mode
// the Gram comes from split-BF16 products hi'hi + hi'lo + lo'hi accumulated in
// FP32 on the matrix cores over chunks of rows and summed in FP64), then one
// FP64 Cholesky factorisation solves all Dout right-hand sides at once (the
// reference refactorises A with LU once per output), followed by FP64
// iterative refinement against the normal equations (in fast mode their
// residual comes from X itself in FP64).  A matrix that is not positive
// definite falls back to LU with partial pivoting.
//
// RLS (reservoir.c:681-767) and the delta rule (reservoir.c:244-280) run as
// fused FP64 kernels on the resident P / W_out.
#include "dev_internal.hpp"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

namespace sdev {

// One rocBLAS handle and one grow-only device workspace per device, shared by
// every reservoir (creating a handle and allocating the n x n systems per fit
// would cost more than the fit itself for the sizes a search evaluates).
// Users hold the device's lock for the whole call; the handle is pointed at
// the caller's stream each time.
struct ReadoutCtx {
    std::mutex mu;
    rocblas_handle h = nullptr;
    double* ws = nullptr; size_t ws_cap = 0;           // doubles
    rocblas_int* iw = nullptr; size_t iw_cap = 0;
};

constexpr int kMaxDevices = 64;
static ReadoutCtx g_readout[kMaxDevices];

static ReadoutCtx* readout_ctx() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    return &g_readout[dev];
}

// Destroys every device's handle and frees its workspace, each with its device
// current; the process-exit path then finds no rocBLAS or HIP object of ours alive.
void readout_shutdown() {
    RandShield rand_shield;
    int cur = 0;
    const bool have_cur = hipGetDevice(&cur) == hipSuccess;
    for (int d = 0; d < kMaxDevices; ++d) {
        ReadoutCtx& c = g_readout[d];
        std::lock_guard<std::mutex> lk(c.mu);
        if (!c.h && !c.ws && !c.iw) continue;
        if (hipSetDevice(d) != hipSuccess) continue;
        (void)hipDeviceSynchronize();
        if (c.h) rocblas_destroy_handle(c.h);
        c.h = nullptr;
        dfree(c.ws);
        dfree(c.iw);
        c.ws = nullptr; c.ws_cap = 0;
        c.iw = nullptr; c.iw_cap = 0;
    }
    if (have_cur) (void)hipSetDevice(cur);
}

// Locks the current device's context, creates its handle on first use and
// binds it to g's stream for the duration of the call.  On release the
// stream is drained and the handle goes back to the null stream: reservoirs
// (and their streams) come and go, and the shared handle must never keep a
// stream that may be destroyed, nor let its workspace be reused on another
// stream while work from this call is still in flight.
struct ReadoutLock {
    ReadoutCtx* c = nullptr;
    std::unique_lock<std::mutex> lk;
    rocblas_handle h = nullptr;
    hipStream_t bound = nullptr;
    explicit ReadoutLock(spires_dev* g) : c(readout_ctx()) {
        if (!c) return;
        lk = std::unique_lock<std::mutex>(c->mu);
        if (!c->h && rocblas_create_handle(&c->h) != rocblas_status_success) c->h = nullptr;
        if (!c->h) return;
        rocblas_set_stream(c->h, g->stream);
        rocblas_set_pointer_mode(c->h, rocblas_pointer_mode_host);
        h = c->h;
        bound = g->stream;
    }
    ~ReadoutLock() {
        if (h) {
            (void)hipStreamSynchronize(bound);
            rocblas_set_stream(h, nullptr);
        }
    }
    // Device scratch of at least `doubles` / `ints` entries (contents undefined).
    int reserve(spires_dev* g, size_t doubles, size_t ints) {
        if (doubles > c->ws_cap) {
            (void)hipStreamSynchronize(g->stream);
            dfree(c->ws);
            c->ws = nullptr;
            c->ws_cap = 0;
            int rc = dalloc(&c->ws, doubles);
            if (rc) return rc;
            c->ws_cap = doubles;
        }
        if (ints > c->iw_cap) {
            (void)hipStreamSynchronize(g->stream);
            dfree(c->iw);
            c->iw = nullptr;
            c->iw_cap = 0;
            int rc = dalloc(&c->iw, ints);
            if (rc) return rc;
            c->iw_cap = ints;
        }
        return 0;
    }
};

void readout_release(spires_dev*) {}

#define SDEV_RB(x)                                                                  \
    do {                                                                            \
        rocblas_status s_ = (x);                                                    \
        if (s_ != rocblas_status_success) {                                         \
            set_error(std::string(#x) + " failed: " + rocblas_status_to_string(s_)); \
            return SPIRES_DEV_EINTERNAL;                                            \
        }                                                                           \
    } while (0)

// A[i,i] += lambda; mirror the lower triangle into the upper one.
__global__ void k_ridge_diag_sym(int n, double* A, double lambda) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)n * n) return;
    const int c = (int)(idx / n), r = (int)(idx % n);   // column-major (r, c)
    if (r == c) A[idx] = A[idx] + lambda;
    else if (r < c) A[idx] = A[(size_t)r * n + c];      // upper <- lower (c, r)
}

// Split of an FP64 matrix into bf16 hi + lo parts: hi = bf16(x), lo = bf16(x - hi),
// round-to-nearest-even on the FP32 value (|x - hi - lo| <~ 2^-16 |x|).
__device__ __forceinline__ unsigned short f32_to_bf16(float f) {
    const unsigned u = __float_as_uint(f);
    const unsigned r = ((u >> 16) & 1u) + 0x7fffu;
    return (unsigned short)((u + r) >> 16);
}
__device__ __forceinline__ float bf16_to_f32(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

__global__ void k_split_bf16(size_t cnt, const double* x, unsigned short* hi, unsigned short* lo) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= cnt) return;
    const double v = x[q];
    const unsigned short h = f32_to_bf16((float)v);
    hi[q] = h;
    lo[q] = f32_to_bf16((float)(v - (double)bf16_to_f32(h)));
}

// A (+)= G1 + C + C' for one chunk of rows: FP32 products combined and summed
// over chunks in FP64.
__global__ void k_accum_gram(int n, const float* G1, const float* Cx, double* A, int first) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)n * n) return;
    const int c = (int)(idx / n), r = (int)(idx % n);
    const double v = (double)G1[idx] + (double)Cx[idx] + (double)Cx[(size_t)r * n + c];
    A[idx] = first ? v : A[idx] + v;
}

__global__ void k_add_diag(int n, double* A, double lambda) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) A[(size_t)i * n + i] += lambda;
}

__global__ void k_axpby(size_t n, double a, const double* x, double* y) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = y[i] + a * x[i];
}

constexpr int kGramChunk = 2048;   // rows per split-BF16 product (FP32 accumulation length)

__global__ void k_axpy(size_t n, const double* x, double* y) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = y[i] + x[i];
}

// ---- RLS, one step: P x, gain, error, rank-1 updates (reservoir.c:721-760)
__global__ void __launch_bounds__(256) k_rls_px(int n, const double* __restrict__ P,
                                                const double* __restrict__ x, double* px) {
    __shared__ double red[256];
    const int row = blockIdx.x;
    double s = 0.0;
    for (int j = threadIdx.x; j < n; j += 256) s += P[(size_t)row * n + j] * x[j];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) px[row] = red[0];
}

// single block: den = lambda + x.Px; k = Px/den; e = y - W x
__global__ void __launch_bounds__(256) k_rls_gain(int n, int dout, double lambda, const double* x,
                                                  double* px, const double* W, const double* y,
                                                  double* e, double* den_out) {
    __shared__ double red[256];
    __shared__ double den_s;
    double s = 0.0;
    for (int j = threadIdx.x; j < n; j += 256) s += x[j] * px[j];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) { den_s = lambda + red[0]; *den_out = den_s; }
    __syncthreads();
    const double sc = 1.0 / den_s;
    for (int o = 0; o < dout; ++o) {
        double t = 0.0;
        for (int j = threadIdx.x; j < n; j += 256) t += W[(size_t)o * n + j] * x[j];
        red[threadIdx.x] = t;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) e[o] = y[o] - red[0];
        __syncthreads();
    }
    for (int j = threadIdx.x; j < n; j += 256) px[j] = sc * px[j];
}

__global__ void k_rls_update(int n, int dout, double lambda, const double* k, const double* e,
                             const double* den, double* P, double* W) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nn = (size_t)n * n;
    if (idx < nn) {
        const int i = (int)(idx / n), j = (int)(idx % n);
        double p = P[idx] + (-(*den) * k[i]) * k[j];
        if (lambda != 1.0) p = (1.0 / lambda) * p;
        P[idx] = p;
    } else if (idx < nn + (size_t)n * dout) {
        const size_t w = idx - nn;
        const int o = (int)(w / n), j = (int)(w % n);
        W[w] = W[w] + (1.0 * e[o]) * k[j];
    }
}

__global__ void k_eye(int n, double* P, double v) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < (size_t)n * n) P[idx] = (idx / n == idx % n) ? v : 0.0;
}

// delta rule: W[o,:] += (lr * (t_o - y_o)) * V
__global__ void k_online(int n, int dout, double lr, const double* t, const double* y,
                         const double* V, double* W) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)n * dout) return;
    const int o = (int)(idx / n), j = (int)(idx % n);
    W[idx] = W[idx] + (lr * (t[o] - y[o])) * V[j];
}

static inline unsigned nb(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

// y[:, d] += alpha[d] x[:, d] over n x D column blocks (alpha on the device).
__global__ void k_col_axpy(int n, int D, const double* __restrict__ alpha, const double* __restrict__ x,
                           double* __restrict__ y) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)n * D) return;
    y[q] = y[q] + alpha[q / n] * x[q];
}

// p[:, d] = z[:, d] + beta[d] p[:, d].
__global__ void k_col_xpby(int n, int D, const double* __restrict__ beta, const double* __restrict__ z,
                           double* __restrict__ p) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)n * D) return;
    p[q] = z[q] + beta[q / n] * p[q];
}

constexpr int kPcgMaxIters = 240;     // over all restarts
constexpr int kPcgRestarts = 6;
constexpr double kPcgTol = 1e-12;      // relative residual of the FP64 normal equations
constexpr int kShiftTries = 10;
constexpr int kPcgStall = 6;           // iterations without a new lowest residual before CG stops

// Fast-mode ridge when n is a multiple of 128.  The Gram comes from the
// split-BF16 matrix-core kernel (dev_gram.hip); its Cholesky factor, of
// G_bf16 + lambda I, or of G_bf16 + (lambda + sigma) I with the smallest shift
// sigma in a doubling sequence that makes it positive definite (the split
// Gram is about 2^-16 off per entry, which can exceed a small lambda), serves
// as the preconditioner of conjugate gradients on the FP64 normal equations
// (X'X + lambda I) W = X'Y, applied matrix-free from the state record (two
// FP64 products with X per iteration), one CG per output.  info[0] gets
// 4 (split-BF16 Gram + CG) + 8 when a shift was needed; info[1] the final
// relative residual of the FP64 normal equations.  Returns kFastFallback
// when no shift in the sequence gives a positive definite factor.
constexpr int kFastFallback = -100;
static int ridge_fast_split(spires_dev* g, size_t row0, size_t row1, const double* Y, double lambda,
                            double* info_out) {
    hipError_t herr = hipSuccess;
    auto H = [&](hipError_t e) { if (e != hipSuccess && herr == hipSuccess) herr = e; };
    const int n = g->n, D = g->dout, T = (int)(row1 - row0);
    ReadoutLock lock(g);
    rocblas_handle h = lock.h;
    if (!h) { set_error("rocblas_create_handle failed"); return SPIRES_DEV_EINTERNAL; }
    const size_t nn = (size_t)n * n, nd = (size_t)n * D;
    const size_t bf_doubles = (gram_split_ws_elems(T, n) * 2 + 7) / 8 + 32;
    int rc = lock.reserve(g, 2 * nn + 7 * nd + 2 * (size_t)T * D + bf_doubles + 4 * (size_t)D, (size_t)1);
    if (rc) return rc;
    double* A = lock.c->ws;            // factor of G_bf16 + (lambda + sigma) I
    double* A0 = A + nn;               // G_bf16 + lambda I (lower)
    double* B = A0 + nn;               // X'Y
    double* W = B + nd;                // solution
    double* R = W + nd;                // residual
    double* Z = R + nd;                // preconditioned residual
    double* P = Z + nd;                // search direction
    double* Q = P + nd;                // A_exact P
    double* Tmp = Q + nd;
    double* XW = Tmp + nd;             // X v [T x D]
    double* Yd = XW + (size_t)T * D;   // targets [T x D]
    double* coef = Yd + (size_t)T * D; // per-output alpha / beta [2 D]
    unsigned short* bf = reinterpret_cast<unsigned short*>(coef + 4 * (size_t)D);
    rocblas_int* info = lock.c->iw;
    const double* Xs = g->rec_state + row0 * (size_t)n;   // column-major n x T (ld n) == X'
    const double one = 1.0, zero = 0.0;
    H(hipMemcpyAsync(Yd, Y, (size_t)T * D * 8, hipMemcpyHostToDevice, g->stream));
    SDEV_RB(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, n, D, T, &one, Xs, n, Yd, D, &zero,
                          B, n));
    if ((rc = gram_split_bf16(g->stream, Xs, T, n, lambda, A0, bf))) return rc;
    // diagonal scale for the shift sequence
    std::vector<double> dg(n);
    H(hipMemcpy2DAsync(dg.data(), 8, A0, (size_t)(n + 1) * 8, 8, (size_t)n, hipMemcpyDeviceToHost, g->stream));
    H(hipStreamSynchronize(g->stream));
    double dmax = 0.0;
    for (double v : dg) dmax = fmax(dmax, fabs(v));
    // The split Gram is off by about 2^-16 of the largest diagonal entry per entry;
    // when lambda does not cover that, start the shift there.
    const double sig0 = ldexp(dmax, -14);
    double sigma = lambda >= sig0 ? 0.0 : sig0 - lambda;
    bool pd = false;
    for (int k = 0; k < kShiftTries && herr == hipSuccess; ++k) {
        H(hipMemcpyAsync(A, A0, nn * 8, hipMemcpyDeviceToDevice, g->stream));
        if (sigma > 0.0) hipLaunchKernelGGL(k_add_diag, dim3(nb((size_t)n, 256)), dim3(256), 0, g->stream, n, A, sigma);
        SDEV_RB(rocsolver_dpotrf(h, rocblas_fill_lower, n, A, n, info));
        rocblas_int hinfo = 0;
        H(hipMemcpyAsync(&hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost, g->stream));
        H(hipStreamSynchronize(g->stream));
        if (hinfo == 0) { pd = true; break; }
        sigma = sigma > 0.0 ? 4.0 * sigma : sig0;
    }
    if (herr != hipSuccess) return hip_fail(herr, "split-BF16 ridge", __LINE__);
    if (!pd) return kFastFallback;
    auto apply = [&](const double* v, double* out) {   // out = (X'X + lambda I) v, FP64, from the record
        if (D == 1) {   // one output: two bandwidth-bound matrix-vector products
            rocblas_dgemv(h, rocblas_operation_transpose, n, T, &one, Xs, n, v, 1, &zero, XW, 1);
            rocblas_dgemv(h, rocblas_operation_none, n, T, &one, Xs, n, XW, 1, &zero, out, 1);
        } else {
            rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, T, D, n, &one, Xs, n, v, n, &zero,
                          XW, T);
            rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, n, D, T, &one, Xs, n, XW, T, &zero, out,
                          n);
        }
        hipLaunchKernelGGL(k_axpby, dim3(nb(nd, 256)), dim3(256), 0, g->stream, nd, lambda, v, out);
    };
    // M^-1 = L^-T L^-1 from the explicit inverse of the factor: two triangular
    // matrix products per application instead of two latency-bound solves
    SDEV_RB(rocsolver_dtrtri(h, rocblas_fill_lower, rocblas_diagonal_non_unit, n, A, n, info));
    auto precond = [&](const double* r, double* z) {
        H(hipMemcpyAsync(z, r, nd * 8, hipMemcpyDeviceToDevice, g->stream));
        if (D == 1) {
            rocblas_dtrmv(h, rocblas_fill_lower, rocblas_operation_none, rocblas_diagonal_non_unit, n, A, n, z, 1);
            rocblas_dtrmv(h, rocblas_fill_lower, rocblas_operation_transpose, rocblas_diagonal_non_unit, n, A, n, z, 1);
        } else {
            rocblas_dtrmm(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, rocblas_diagonal_non_unit,
                          n, D, &one, A, n, z, n, z, n);
            rocblas_dtrmm(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_transpose,
                          rocblas_diagonal_non_unit, n, D, &one, A, n, z, n, z, n);
        }
    };
    std::vector<double> hb(nd), hv(nd), hw(nd), dots(D), rho(D), bn(D);
    auto col_dots = [&](const double* a, const double* b, std::vector<double>& out) {   // host, fixed order
        H(hipMemcpyAsync(hv.data(), a, nd * 8, hipMemcpyDeviceToHost, g->stream));
        H(hipMemcpyAsync(hw.data(), b, nd * 8, hipMemcpyDeviceToHost, g->stream));
        H(hipStreamSynchronize(g->stream));
        for (int d = 0; d < D; ++d) {
            double s0 = 0.0;
            for (int i = 0; i < n; ++i) s0 += hv[(size_t)d * n + i] * hw[(size_t)d * n + i];
            out[d] = s0;
        }
    };
    col_dots(B, B, bn);
    std::vector<double> rr(D), al(D), be(D);
    auto residual_rel = [&]() {
        col_dots(R, R, rr);
        double m = 0.0;
        for (int d = 0; d < D; ++d) if (bn[d] > 0) m = fmax(m, sqrt(rr[d] / bn[d]));
        return m;
    };
    // (Re)start from W: R = B - A W exactly, Z = M^-1 R, P = Z.
    auto restart = [&]() {
        apply(W, Q);
        H(hipMemcpyAsync(R, B, nd * 8, hipMemcpyDeviceToDevice, g->stream));
        hipLaunchKernelGGL(k_axpby, dim3(nb(nd, 256)), dim3(256), 0, g->stream, nd, -1.0, Q, R);
        precond(R, Z);
        H(hipMemcpyAsync(P, Z, nd * 8, hipMemcpyDeviceToDevice, g->stream));
        col_dots(R, Z, rho);
        return residual_rel();
    };
    precond(B, W);                                      // W0 = M^-1 B
    double rel = restart();
    double best = rel;
    H(hipMemcpyAsync(Tmp, W, nd * 8, hipMemcpyDeviceToDevice, g->stream));   // best iterate so far
    // CG loses orthogonality in floating point once the residual is small next to
    // the conditioning of the preconditioned system; when it stops improving it is
    // restarted from its best iterate (recomputing the true residual)
    int iters = 0;
    for (int cycle = 0; cycle <= kPcgRestarts && best > kPcgTol && herr == hipSuccess; ++cycle) {
        if (cycle > 0) {
            H(hipMemcpyAsync(W, Tmp, nd * 8, hipMemcpyDeviceToDevice, g->stream));
            rel = restart();
        }
        int best_it = -1;
        for (int it = 0; iters < kPcgMaxIters && rel > kPcgTol && herr == hipSuccess; ++it, ++iters) {
            apply(P, Q);
            col_dots(P, Q, dots);
            for (int d = 0; d < D; ++d) al[d] = dots[d] != 0.0 ? rho[d] / dots[d] : 0.0;
            H(hipMemcpyAsync(coef, al.data(), D * 8, hipMemcpyHostToDevice, g->stream));
            hipLaunchKernelGGL(k_col_axpy, dim3(nb(nd, 256)), dim3(256), 0, g->stream, n, D, coef, P, W);
            for (int d = 0; d < D; ++d) al[d] = -al[d];
            H(hipMemcpyAsync(coef + D, al.data(), D * 8, hipMemcpyHostToDevice, g->stream));
            hipLaunchKernelGGL(k_col_axpy, dim3(nb(nd, 256)), dim3(256), 0, g->stream, n, D, coef + D, Q, R);
            rel = residual_rel();
            if (rel < best) {
                best = rel;
                best_it = it;
                H(hipMemcpyAsync(Tmp, W, nd * 8, hipMemcpyDeviceToDevice, g->stream));
            }
            if (rel <= kPcgTol || it - best_it >= kPcgStall) break;   // converged, or no progress lately
            precond(R, Z);
            col_dots(R, Z, dots);
            for (int d = 0; d < D; ++d) { be[d] = rho[d] != 0.0 ? dots[d] / rho[d] : 0.0; rho[d] = dots[d]; }
            H(hipMemcpyAsync(coef + 2 * D, be.data(), D * 8, hipMemcpyHostToDevice, g->stream));
            hipLaunchKernelGGL(k_col_xpby, dim3(nb(nd, 256)), dim3(256), 0, g->stream, n, D, coef + 2 * D, Z, P);
        }
        if (best_it < 0) break;                          // this cycle found nothing better
    }
    if (info_out) {
        info_out[0] = 4.0 + (sigma > 0.0 ? 8.0 : 0.0);
        info_out[1] = best;
    }
    H(hipMemcpyAsync(g->W_out, Tmp, nd * 8, hipMemcpyDeviceToDevice, g->stream));
    H(hipStreamSynchronize(g->stream));
    if (herr != hipSuccess) return hip_fail(herr, "split-BF16 ridge", __LINE__);
    return 0;
}

// Y [T x Dout] = X [T x N] W_out' for a recorded state matrix (row-major).
int outputs_from_states(spires_dev* g, const double* states, size_t T, double* out) {
    if (g->dout == 0 || T == 0) return 0;
    ReadoutLock lock(g);
    rocblas_handle h = lock.h;
    if (!h) { set_error("rocblas_create_handle failed"); return SPIRES_DEV_EINTERNAL; }
    const double one = 1.0, zero = 0.0;
    // column-major view: Y' (Dout x T) = W_out (Dout x N, ld N as row-major -> use transpose) * X' (N x T)
    SDEV_RB(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, g->dout, (int)T, g->n, &one,
                          g->W_out, g->n, states, g->n, &zero, out, g->dout));
    return 0;
}

}  // namespace sdev

using namespace sdev;

extern "C" {

void spires_dev_shutdown(void) { readout_shutdown(); }

static int ridge_batch_impl(spires_dev* const* gs, int count, size_t row0, size_t row1, const double* const* Y,
                            const double* lambda, double* info_out);

int spires_dev_ridge(spires_dev* g, size_t washout, const double* Y, double lambda, double* info_out) {
    sdev::RandShield rand_shield_;
    if (!g) return SPIRES_DEV_EINVAL;
    return spires_dev_ridge_rows(g, washout, g->rec_rows, Y, lambda, info_out);
}

int spires_dev_ridge_rows(spires_dev* g, size_t row0, size_t row1, const double* Y, double lambda,
                          double* info_out) {
    sdev::RandShield rand_shield_;
    hipError_t herr = hipSuccess;
    auto H = [&](hipError_t e) { if (e != hipSuccess && herr == hipSuccess) herr = e; };
    if (!g || !Y) return SPIRES_DEV_EINVAL;
    if (const int rc0 = settle(g)) return rc0;
    if (!g->rec_state || row0 >= row1 || row1 > g->rec_rows) { set_error("no recorded states for ridge"); return SPIRES_DEV_EINVAL; }
    const size_t washout = row0;
    const int n = g->n, D = g->dout;
    const int T = (int)(row1 - row0);
    if (D == 0) return 0;
    if (g->mode == SPIRES_DEV_MODE_EXACT && chol_batch_ok(n)) {   // Gram + Cholesky kernels of dev_chol.hip
        double inf[2] = {0.0, 0.0};
        const int rc = ridge_batch_impl(&g, 1, row0, row1, &Y, &lambda, inf);
        if (rc) return rc;
        if (info_out) {
            info_out[0] = inf[0];
            info_out[1] = inf[1];
        }
        if (inf[0] < 0) {   // W_out left unchanged, as the reference does
            set_error("ridge system is singular");
            return SPIRES_DEV_ESINGULAR;
        }
        return 0;
    }
    if (g->mode == SPIRES_DEV_MODE_FAST && n % 128 == 0) {   // split-BF16 Gram kernel + preconditioned CG
        const int rcf = ridge_fast_split(g, row0, row1, Y, lambda, info_out);
        if (rcf != kFastFallback) return rcf;
    }
    ReadoutLock lock(g);
    rocblas_handle h = lock.h;
    if (!h) { set_error("rocblas_create_handle failed"); return SPIRES_DEV_EINTERNAL; }
    const size_t nn = (size_t)n * n, nd = (size_t)n * D;
    // workspace: A, A0 [n x n]; B, X, R [n x D]; Y [T x D]; fast mode also
    // hi, lo (bf16 [n x T]) and G1, C (f32 [n x n]); info + pivots
    bool fast = g->mode == SPIRES_DEV_MODE_FAST;
    const int TC = std::min(T, kGramChunk);
    const size_t cnt = (size_t)n * TC;
    auto al = [](size_t bytes) { return (bytes + 255) / 256 * 256; };   // 256-byte aligned pieces
    const size_t fast_extra = fast ? (2 * al(nn * 4) + 2 * al(cnt * 2) + al((size_t)T * D * 8) + al(nd * 8)) / 8 + 48 : 0;
    int rc = lock.reserve(g, 2 * nn + 3 * nd + (size_t)T * D + fast_extra, (size_t)n + 1);
    if (rc) return rc;
    double* A = lock.c->ws;
    double* A0 = A + nn;
    double* B = A0 + nn;
    double* X = B + nd;
    double* R = X + nd;
    double* Yd = R + nd;
    char* fast_ws = reinterpret_cast<char*>(Yd + (size_t)T * D);
    fast_ws += (256 - reinterpret_cast<uintptr_t>(fast_ws) % 256) % 256;
    rocblas_int* info = lock.c->iw;
    rocblas_int* ipiv = info + 1;
    auto cleanup = [&]() {};
    const double* Xs = g->rec_state + washout * (size_t)n;   // column-major n x T (ld n) == X'
    const double one = 1.0, zero = 0.0, mone = -1.0;
    H(hipMemcpyAsync(Yd, Y, (size_t)T * D * 8, hipMemcpyHostToDevice, g->stream));
    // B = X'Y (FP64)
    rocblas_status s2 = rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, n, D, T, &one, Xs, n,
                                      Yd, D, &zero, B, n);
    rocblas_status s1 = rocblas_status_success;
    double* XW = nullptr;   // fast mode: X W [T x D] of the refinement residual
    double* dX = nullptr;   // fast mode: the last refinement step [n x D]
    if (fast) {
        // Gram from split-BF16 products, X'X ~ hi'hi + hi'lo + lo'hi, over chunks of
        // kGramChunk rows: each chunk accumulates in FP32 on the matrix cores, the
        // chunks add up in FP64 (the FP32 accumulation length stays kGramChunk).
        float* G1 = reinterpret_cast<float*>(fast_ws);
        float* Cx = reinterpret_cast<float*>(fast_ws + al(nn * 4));
        unsigned short* hi = reinterpret_cast<unsigned short*>(fast_ws + 2 * al(nn * 4));
        unsigned short* lo = reinterpret_cast<unsigned short*>(fast_ws + 2 * al(nn * 4) + al(cnt * 2));
        XW = reinterpret_cast<double*>(fast_ws + 2 * al(nn * 4) + 2 * al(cnt * 2));
        dX = reinterpret_cast<double*>(fast_ws + 2 * al(nn * 4) + 2 * al(cnt * 2) + al((size_t)T * D * 8));
        const float onef = 1.0f, zerof = 0.0f;
        for (int c0 = 0; c0 < T && s1 == rocblas_status_success; c0 += TC) {
            const int tc = std::min(TC, T - c0);
            const size_t cc = (size_t)n * tc;
            hipLaunchKernelGGL(k_split_bf16, dim3(nb(cc, 256)), dim3(256), 0, g->stream, cc, Xs + (size_t)c0 * n, hi,
                               lo);
            s1 = rocblas_gemm_ex(h, rocblas_operation_none, rocblas_operation_transpose, n, n, tc, &onef, hi,
                                 rocblas_datatype_bf16_r, n, hi, rocblas_datatype_bf16_r, n, &zerof, G1,
                                 rocblas_datatype_f32_r, n, G1, rocblas_datatype_f32_r, n, rocblas_datatype_f32_r,
                                 rocblas_gemm_algo_standard, 0, 0);
            if (s1 == rocblas_status_success)
                s1 = rocblas_gemm_ex(h, rocblas_operation_none, rocblas_operation_transpose, n, n, tc, &onef, hi,
                                     rocblas_datatype_bf16_r, n, lo, rocblas_datatype_bf16_r, n, &zerof, Cx,
                                     rocblas_datatype_f32_r, n, Cx, rocblas_datatype_f32_r, n,
                                     rocblas_datatype_f32_r, rocblas_gemm_algo_standard, 0, 0);
            hipLaunchKernelGGL(k_accum_gram, dim3(nb(nn, 256)), dim3(256), 0, g->stream, n, G1, Cx, A, c0 == 0 ? 1 : 0);
        }
        hipLaunchKernelGGL(k_add_diag, dim3(nb((size_t)n, 256)), dim3(256), 0, g->stream, n, A, lambda);
    } else {
        // A = X'X (lower) in FP64
        s1 = rocblas_dsyrk(h, rocblas_fill_lower, rocblas_operation_none, n, T, &one, Xs, n, &zero, A, n);
        hipLaunchKernelGGL(k_ridge_diag_sym, dim3(nb(nn, 256)), dim3(256), 0, g->stream, n, A, lambda);
    }
    if (s1 != rocblas_status_success || s2 != rocblas_status_success) {
        cleanup();
        set_error("Gram/cross products failed");
        return SPIRES_DEV_EINTERNAL;
    }
    H(hipMemcpyAsync(A0, A, nn * 8, hipMemcpyDeviceToDevice, g->stream));
    H(hipMemcpyAsync(X, B, nd * 8, hipMemcpyDeviceToDevice, g->stream));
    rocblas_status s3 = rocsolver_dpotrf(h, rocblas_fill_lower, n, A, n, info);
    rocblas_int hinfo = 0;
    H(hipMemcpyAsync(&hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost, g->stream));
    H(hipStreamSynchronize(g->stream));
    bool use_lu = (s3 != rocblas_status_success) || hinfo != 0;
    bool fp64_gram = !fast;
    if (use_lu && fast) {
        // The split-BF16 Gram (about 2^-16 relative) is not positive definite:
        // lambda is too small for it.  Form the FP64 Gram instead and factor that.
        fast = false;
        fp64_gram = true;
        rocblas_status s5 = rocblas_dsyrk(h, rocblas_fill_lower, rocblas_operation_none, n, T, &one, Xs, n, &zero, A, n);
        hipLaunchKernelGGL(k_ridge_diag_sym, dim3(nb(nn, 256)), dim3(256), 0, g->stream, n, A, lambda);
        H(hipMemcpyAsync(A0, A, nn * 8, hipMemcpyDeviceToDevice, g->stream));
        H(hipMemcpyAsync(X, B, nd * 8, hipMemcpyDeviceToDevice, g->stream));
        s3 = s5 == rocblas_status_success ? rocsolver_dpotrf(h, rocblas_fill_lower, n, A, n, info) : s5;
        H(hipMemcpyAsync(&hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost, g->stream));
        H(hipStreamSynchronize(g->stream));
        use_lu = (s3 != rocblas_status_success) || hinfo != 0;
    }
    if (use_lu) {
        H(hipMemcpyAsync(A, A0, nn * 8, hipMemcpyDeviceToDevice, g->stream));
        rocblas_status s4 = rocsolver_dgetrf(h, n, n, A, n, ipiv, info);
        rocblas_int i2 = 0;
        H(hipMemcpyAsync(&i2, info, sizeof(i2), hipMemcpyDeviceToHost, g->stream));
        H(hipStreamSynchronize(g->stream));
        if (s4 != rocblas_status_success || i2 != 0) {
            cleanup();
            set_error("ridge system is singular");
            if (info_out) info_out[0] = (double)i2;
            return SPIRES_DEV_ESINGULAR;   // W_out left unchanged, as the reference does
        }
        rocsolver_dgetrs(h, rocblas_operation_none, n, D, A, n, ipiv, X, n);
    } else {
        rocsolver_dpotrs(h, rocblas_fill_lower, n, D, A, n, X, n);
    }
    // Residual R = B - A W of the factored system (exact mode: the FP64 Gram;
    // fast mode: the split-BF16 Gram in FP64), or, with exact_gram in fast
    // mode, of the FP64 normal equations B - (X'X + lambda I) W formed from X
    // itself (two products with the state record).
    auto residual = [&](bool exact_gram) {
        H(hipMemcpyAsync(R, B, nd * 8, hipMemcpyDeviceToDevice, g->stream));
        if (fast && exact_gram) {
            rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, T, D, n, &one, Xs, n, X, n, &zero, XW,
                          T);
            rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, n, D, T, &mone, Xs, n, XW, T, &one, R, n);
            hipLaunchKernelGGL(k_axpby, dim3(nb(nd, 256)), dim3(256), 0, g->stream, nd, -lambda, X, R);
        } else {
            rocblas_dsymm(h, rocblas_side_left, rocblas_fill_lower, n, D, &mone, A0, n, X, n, &one, R, n);
        }
    };
    auto solve_r = [&]() {   // R <- A^-1 R with the factorisation
        if (use_lu) rocsolver_dgetrs(h, rocblas_operation_none, n, D, A, n, ipiv, R, n);
        else rocsolver_dpotrs(h, rocblas_fill_lower, n, D, A, n, R, n);
    };
    std::vector<double> hr(nd), hb(nd);
    auto rel_norm = [&]() {  // max over outputs of ||R|| / ||B|| (host)
        H(hipMemcpyAsync(hr.data(), R, nd * 8, hipMemcpyDeviceToHost, g->stream));
        H(hipMemcpyAsync(hb.data(), B, nd * 8, hipMemcpyDeviceToHost, g->stream));
        H(hipStreamSynchronize(g->stream));
        double rel = 0.0;
        for (int o = 0; o < D; ++o) {
            double a = 0, b = 0;
            for (int i = 0; i < n; ++i) { a += hr[(size_t)o * n + i] * hr[(size_t)o * n + i]; b += hb[(size_t)o * n + i] * hb[(size_t)o * n + i]; }
            if (b > 0) rel = fmax(rel, sqrt(a / b));
        }
        return rel;
    };
    // iterative refinement against the factored system: R = B - A X ; solve ; X += dX
    for (int it = 0; it < 2; ++it) {
        residual(false);
        solve_r();
        hipLaunchKernelGGL(k_axpy, dim3(nb(nd, 256)), dim3(256), 0, g->stream, nd, R, X);
    }
    // fast mode: further steps against the FP64 normal equations, each kept only
    // while it lowers their residual (the split-BF16 factor is only a
    // preconditioner there, which may not converge when lambda is small)
    if (fast) {
        residual(true);
        double cur = rel_norm();
        for (int it = 0; it < 4 && herr == hipSuccess; ++it) {
            solve_r();
            hipLaunchKernelGGL(k_axpy, dim3(nb(nd, 256)), dim3(256), 0, g->stream, nd, R, X);
            H(hipMemcpyAsync(dX, R, nd * 8, hipMemcpyDeviceToDevice, g->stream));   // keep the step
            residual(true);
            const double nxt = rel_norm();
            if (!(nxt < cur)) {   // undo the step
                hipLaunchKernelGGL(k_axpby, dim3(nb(nd, 256)), dim3(256), 0, g->stream, nd, -1.0, dX, X);
                break;
            }
            cur = nxt;
        }
    }
    double rel = 0.0;
    if (info_out) {
        residual(fast);
        rel = rel_norm();
        info_out[0] = (use_lu ? 1.0 : 0.0) + (fp64_gram && g->mode == SPIRES_DEV_MODE_FAST ? 2.0 : 0.0);
        info_out[1] = rel;
    }
    // X is column-major n x D == row-major W_out [D x n]
    H(hipMemcpyAsync(g->W_out, X, nd * 8, hipMemcpyDeviceToDevice, g->stream));
    H(hipStreamSynchronize(g->stream));
    cleanup();
    if (herr != hipSuccess) return hip_fail(herr, "ridge", __LINE__);
    return 0;
}

int spires_dev_ridge_rows_batch(spires_dev* const* gs, int count, size_t row0, size_t row1, const double* const* Y,
                                const double* lambda, double* info_out) {
    sdev::RandShield rand_shield_;
    if (!gs || count < 1 || !Y || !lambda) return SPIRES_DEV_EINVAL;
    spires_dev* g0 = gs[0];
    if (!g0) return SPIRES_DEV_EINVAL;
    const int n = g0->n, D = g0->dout;
    bool same = true;
    for (int r = 0; r < count; ++r) {
        if (!gs[r] || !Y[r]) return SPIRES_DEV_EINVAL;
        if (const int rc0 = settle(gs[r])) return rc0;
        if (!gs[r]->rec_state || row0 >= row1 || row1 > gs[r]->rec_rows) { set_error("no recorded states for ridge"); return SPIRES_DEV_EINVAL; }
        same = same && gs[r]->n == n && gs[r]->dout == D && gs[r]->mode == SPIRES_DEV_MODE_EXACT;
    }
    if (!same || (count == 1 && !chol_batch_ok(n))) {   // one system at a time
        for (int r = 0; r < count; ++r) {
            int rc = spires_dev_ridge_rows(gs[r], row0, row1, Y[r], lambda[r], info_out ? info_out + 2 * r : nullptr);
            if (rc) return rc;
        }
        return 0;
    }
    return ridge_batch_impl(gs, count, row0, row1, Y, lambda, info_out);
}

// Exact-mode systems of one order n and output count D, factorised together.
// Singular systems (LU fails too) keep their W_out and report -1.
static int ridge_batch_impl(spires_dev* const* gs, int count, size_t row0, size_t row1, const double* const* Y,
                            const double* lambda, double* info_out) {
    spires_dev* g0 = gs[0];
    const int n = g0->n, D = g0->dout;
    if (D == 0) return 0;
    hipError_t herr = hipSuccess;
    auto H = [&](hipError_t e) { if (e != hipSuccess && herr == hipSuccess) herr = e; };
    for (int r = 0; r < count; ++r) H(hipStreamSynchronize(gs[r]->stream));   // records complete
    if (herr != hipSuccess) return hip_fail(herr, "ridge batch", __LINE__);
    ReadoutLock lock(g0);
    rocblas_handle h = lock.h;
    if (!h) { set_error("rocblas_create_handle failed"); return SPIRES_DEV_EINTERNAL; }
    const int T = (int)(row1 - row0);
    const size_t nn = (size_t)n * n, nd = (size_t)n * D, td = (size_t)T * D;
    // workspace: A, A0 [count][n x n]; B, X, R [count][n x D]; Y [count][T x D];
    // then the batched-factorisation workspace; info [count] + pivots [n].
    // Orders that are multiples of 64 use the batched Gram / Cholesky / solve
    // kernels of dev_chol.hip, others rocBLAS + rocSOLVER.
    const bool own = chol_batch_ok(n);
    const bool own_rhs = own && D <= kCholRhsMax;   // cross products, solves, residuals too
    const size_t cw = own ? chol_batch_ws_doubles(count, n, T, D) : 0;
    int rc = lock.reserve(g0, (size_t)count * (2 * nn + 3 * nd + td) + cw, (size_t)count + n);
    if (rc) return rc;
    double* A = lock.c->ws;
    double* A0 = A + (size_t)count * nn;
    double* B = A0 + (size_t)count * nn;
    double* X = B + (size_t)count * nd;
    double* R = X + (size_t)count * nd;
    double* Yd = R + (size_t)count * nd;
    double* CW = Yd + (size_t)count * td;
    rocblas_int* info = lock.c->iw;
    rocblas_int* ipiv = info + count;
    hipStream_t st = g0->stream;
    const double one = 1.0, zero = 0.0, mone = -1.0;
    rocblas_status sb = rocblas_status_success;
    auto B_ok = [&](rocblas_status x) { if (x != rocblas_status_success && sb == rocblas_status_success) sb = x; };
    std::vector<const double*> xs(count);
    for (int r = 0; r < count; ++r) {
        xs[r] = gs[r]->rec_state + row0 * (size_t)n;   // column-major n x T == X'
        H(hipMemcpyAsync(Yd + r * td, Y[r], td * 8, hipMemcpyHostToDevice, st));
        if (own_rhs) continue;
        B_ok(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, n, D, T, &one, xs[r], n, Yd + r * td,
                           D, &zero, B + r * nd, n));
        if (own) continue;
        B_ok(rocblas_dsyrk(h, rocblas_fill_lower, rocblas_operation_none, n, T, &one, xs[r], n, &zero, A + r * nn, n));
        hipLaunchKernelGGL(k_ridge_diag_sym, dim3(nb(nn, 256)), dim3(256), 0, st, n, A + r * nn, lambda[r]);
    }
    if (own) H(gram_cross_batched(st, count, n, T, D, xs.data(), Yd, lambda, A, nn, B, nd, CW, own_rhs));
    H(hipMemcpyAsync(A0, A, (size_t)count * nn * 8, hipMemcpyDeviceToDevice, st));
    H(hipMemcpyAsync(X, B, (size_t)count * nd * 8, hipMemcpyDeviceToDevice, st));
    if (own) H(chol_batched(st, count, n, A, nn, CW, info));
    else B_ok(rocsolver_dpotrf_strided_batched(h, rocblas_fill_lower, n, A, n, (rocblas_stride)nn, info, count));
    std::vector<rocblas_int> hinfo(count, 0);
    H(hipMemcpyAsync(hinfo.data(), info, count * sizeof(rocblas_int), hipMemcpyDeviceToHost, st));
    H(hipStreamSynchronize(st));
    if (sb != rocblas_status_success || herr != hipSuccess) {
        set_error("batched ridge factorisation failed");
        return SPIRES_DEV_EINTERNAL;
    }
    auto solve_all = [&](double* V) {   // V <- A^-1 V with the Cholesky factors
        if (own_rhs) {
            H(chol_solve_batched(st, count, n, D, A, nn, CW, V, nd));
            return;
        }
        if (own) {                  // the kernels' factor: blocked solves over all D columns
            H(chol_solve_many(st, count, n, D, A, nn, CW, V, nd));
            return;
        }
        // L Z = V, then L' X = Z (BLAS3 triangular solves over all D columns)
        B_ok(rocblas_dtrsm_strided_batched(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none,
                                           rocblas_diagonal_non_unit, n, D, &one, A, n, (rocblas_stride)nn, V, n,
                                           (rocblas_stride)nd, count));
        B_ok(rocblas_dtrsm_strided_batched(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_transpose,
                                           rocblas_diagonal_non_unit, n, D, &one, A, n, (rocblas_stride)nn, V, n,
                                           (rocblas_stride)nd, count));
    };
    solve_all(X);
    // systems that are not positive definite: LU with partial pivoting, one at a time
    std::vector<char> lu(count, 0), singular(count, 0);
    for (int r = 0; r < count; ++r) {
        if (hinfo[r] == 0) continue;
        lu[r] = 1;
        H(hipMemcpyAsync(A + r * nn, A0 + r * nn, nn * 8, hipMemcpyDeviceToDevice, st));
        B_ok(rocsolver_dgetrf(h, n, n, A + r * nn, n, ipiv, info));
        rocblas_int i2 = 0;
        H(hipMemcpyAsync(&i2, info, sizeof(i2), hipMemcpyDeviceToHost, st));
        H(hipStreamSynchronize(st));
        if (i2 != 0) { singular[r] = 1; continue; }
        H(hipMemcpyAsync(X + r * nd, B + r * nd, nd * 8, hipMemcpyDeviceToDevice, st));
        B_ok(rocsolver_dgetrs(h, rocblas_operation_none, n, D, A + r * nn, n, ipiv, X + r * nd, n));
        // keep this system's pivots next to its factor for the refinement solves
        for (int it = 0; it < 2; ++it) {
            H(hipMemcpyAsync(R + r * nd, B + r * nd, nd * 8, hipMemcpyDeviceToDevice, st));
            B_ok(rocblas_dsymm(h, rocblas_side_left, rocblas_fill_lower, n, D, &mone, A0 + r * nn, n, X + r * nd, n, &one,
                               R + r * nd, n));
            B_ok(rocsolver_dgetrs(h, rocblas_operation_none, n, D, A + r * nn, n, ipiv, R + r * nd, n));
            hipLaunchKernelGGL(k_axpy, dim3(nb(nd, 256)), dim3(256), 0, st, nd, R + r * nd, X + r * nd);
        }
    }
    auto resid_all = [&]() {   // R = B - A0 X
        if (own_rhs) {
            H(resid_batched(st, count, n, D, A0, nn, X, B, R, nd, CW));
            return;
        }
        H(hipMemcpyAsync(R, B, (size_t)count * nd * 8, hipMemcpyDeviceToDevice, st));
        B_ok(rocblas_dsymm_strided_batched(h, rocblas_side_left, rocblas_fill_lower, n, D, &mone, A0, n,
                                           (rocblas_stride)nn, X, n, (rocblas_stride)nd, &one, R, n,
                                           (rocblas_stride)nd, count));
    };
    // iterative refinement of the Cholesky solutions: R = B - A0 X ; solve ; X += R
    for (int it = 0; it < 2; ++it) {
        resid_all();
        solve_all(R);
        for (int r = 0; r < count; ++r)
            if (!lu[r]) hipLaunchKernelGGL(k_axpy, dim3(nb(nd, 256)), dim3(256), 0, st, nd, R + r * nd, X + r * nd);
    }
    if (info_out) {
        resid_all();
        std::vector<double> hr((size_t)count * nd), hb((size_t)count * nd);
        H(hipMemcpyAsync(hr.data(), R, hr.size() * 8, hipMemcpyDeviceToHost, st));
        H(hipMemcpyAsync(hb.data(), B, hb.size() * 8, hipMemcpyDeviceToHost, st));
        H(hipStreamSynchronize(st));
        for (int r = 0; r < count; ++r) {
            double rel = 0.0;
            for (int o = 0; o < D; ++o) {
                double a = 0, b = 0;
                for (int i = 0; i < n; ++i) {
                    const double x = hr[r * nd + (size_t)o * n + i], y = hb[r * nd + (size_t)o * n + i];
                    a += x * x;
                    b += y * y;
                }
                if (b > 0) rel = fmax(rel, sqrt(a / b));
            }
            info_out[2 * r] = singular[r] ? -1.0 : (lu[r] ? 1.0 : 0.0);
            info_out[2 * r + 1] = rel;
        }
    }
    for (int r = 0; r < count; ++r)   // singular systems keep their W_out, as the reference does
        if (!singular[r]) H(hipMemcpyAsync(gs[r]->W_out, X + r * nd, nd * 8, hipMemcpyDeviceToDevice, st));
    H(hipStreamSynchronize(st));
    if (sb != rocblas_status_success) { set_error("batched ridge solve failed"); return SPIRES_DEV_EINTERNAL; }
    if (herr != hipSuccess) return hip_fail(herr, "ridge batch", __LINE__);
    return 0;
}

int spires_dev_predict_records(spires_dev* g, size_t row0, size_t rows, double* out) {
    sdev::RandShield rand_shield_;
    if (!g || (!out && rows)) return SPIRES_DEV_EINVAL;
    if (const int rc0 = settle(g)) return rc0;
    if (!g->rec_state || row0 + rows > g->rec_rows) { set_error("rows beyond the recorded states"); return SPIRES_DEV_EINVAL; }
    if (rows == 0 || g->dout == 0) return 0;
    ReadoutLock lock(g);
    rocblas_handle h = lock.h;
    if (!h) { set_error("rocblas_create_handle failed"); return SPIRES_DEV_EINTERNAL; }
    const size_t cnt = rows * (size_t)g->dout;
    int rc = lock.reserve(g, cnt, 0);
    if (rc) return rc;
    const double one = 1.0, zero = 0.0;
    SDEV_RB(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, g->dout, (int)rows, g->n, &one,
                          g->W_out, g->n, g->rec_state + row0 * (size_t)g->n, g->n, &zero, lock.c->ws, g->dout));
    hipError_t e = hipMemcpyAsync(out, lock.c->ws, cnt * 8, hipMemcpyDeviceToHost, g->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
    return e == hipSuccess ? 0 : hip_fail(e, "predict records", __LINE__);
}

int spires_dev_rls(spires_dev* g, const double* Y, size_t T, double delta, double lambda) {
    sdev::RandShield rand_shield_;
    hipError_t herr = hipSuccess;
    auto H = [&](hipError_t e) { if (e != hipSuccess && herr == hipSuccess) herr = e; };
    if (!g || (!Y && T)) return SPIRES_DEV_EINVAL;
    if (const int rc0 = settle(g)) return rc0;
    if (T > g->u_T && g->din > 0) { set_error("rls beyond uploaded inputs"); return SPIRES_DEV_EINVAL; }
    const int n = g->n, D = g->dout;
    const size_t nn = (size_t)n * n;
    double *P = nullptr, *k = nullptr, *e = nullptr, *den = nullptr, *Yd = nullptr;
    int rc = 0;
    auto cleanup = [&]() { dfree(P); dfree(k); dfree(e); dfree(den); dfree(Yd); };
    if ((rc = dalloc(&P, nn)) || (rc = dalloc(&k, (size_t)n)) || (rc = dalloc(&e, (size_t)D + 1)) ||
        (rc = dalloc(&den, 1)) || (rc = dalloc(&Yd, T * (size_t)D + 1))) {
        cleanup();
        return rc;
    }
    H(hipMemcpyAsync(Yd, Y, T * (size_t)D * 8, hipMemcpyHostToDevice, g->stream));
    hipLaunchKernelGGL(k_eye, dim3(nb(nn, 256)), dim3(256), 0, g->stream, n, P, 1.0 / delta);
    H(hipMemsetAsync(g->W_out, 0, (size_t)n * D * 8, g->stream));
    for (size_t t = 0; t < T; ++t) {
        const double* u = g->din > 0 ? g->u_series + t * g->din : nullptr;
        if ((rc = launch_step(g, u, nullptr, nullptr, nullptr))) { cleanup(); return rc; }
        hipLaunchKernelGGL(k_rls_px, dim3(n), dim3(256), 0, g->stream, n, P, g->V, k);
        hipLaunchKernelGGL(k_rls_gain, dim3(1), dim3(256), 0, g->stream, n, D, lambda, g->V, k, g->W_out,
                           Yd + t * D, e, den);
        const size_t tot = nn + (size_t)n * D;
        hipLaunchKernelGGL(k_rls_update, dim3(nb(tot, 256)), dim3(256), 0, g->stream, n, D, lambda, k, e, den,
                           P, g->W_out);
    }
    H(hipStreamSynchronize(g->stream));
    cleanup();
    if (herr != hipSuccess) return hip_fail(herr, "rls", __LINE__);
    return 0;
}

int spires_dev_online(spires_dev* g, const double* target, double lr) {
    sdev::RandShield rand_shield_;
    hipError_t herr = hipSuccess;
    auto H = [&](hipError_t e) { if (e != hipSuccess && herr == hipSuccess) herr = e; };
    if (!g || !target) return SPIRES_DEV_EINVAL;
    if (const int rc0 = settle(g)) return rc0;
    const int n = g->n, D = g->dout;
    if (D == 0) return 0;
    double* td = nullptr;
    int rc = dalloc(&td, (size_t)D);
    if (rc) return rc;
    H(hipMemcpyAsync(td, target, (size_t)D * 8, hipMemcpyHostToDevice, g->stream));
    if ((rc = launch_output(g, g->ybuf))) { dfree(td); return rc; }
    hipLaunchKernelGGL(k_online, dim3(nb((size_t)n * D, 256)), dim3(256), 0, g->stream, n, D, lr, td, g->ybuf,
                       g->V, g->W_out);
    H(hipStreamSynchronize(g->stream));
    dfree(td);
    if (herr != hipSuccess) return hip_fail(herr, "online", __LINE__);
    return 0;
}

}  // extern "C"
