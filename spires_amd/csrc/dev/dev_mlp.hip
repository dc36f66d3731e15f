// AGILE surrogate: an FP32 MLP in -> H -> H -> 1 (tanh hidden layers, linear
// output) trained with Adam on mean-squared error.  It predicts a trial's
// validation score from its hyperparameters and fidelity, so that the search
// can prune candidates between fidelity rungs.
//
// Parameters live in one device buffer, in this order:
//   W1 [H x in] (row-major), b1 [H], W2 [H x H], b2 [H], W3 [H], b3 [1].
// Gradients are data parallel: workgroup w copies the parameters into LDS and
// takes the batch's slices of kMlpSlice samples w, w + G, w + 2G, ... in that
// order, summing its per-parameter partial gradient in registers; a second
// kernel adds the G partials in workgroup order and applies Adam.  Every sum
// has a fixed order (no atomics), so results are deterministic run to run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "spires_ext.h"
#include "dev_internal.hpp"

namespace {

constexpr int kMlpThreads = 256;
constexpr int kMlpSlice = 8;         // samples per pass through the layers
constexpr int kMaxIn = 16;
constexpr int kMaxHidden = 64;
constexpr int kMaxParams = kMaxHidden * kMaxIn + kMaxHidden + kMaxHidden * kMaxHidden + 2 * kMaxHidden + 1;
constexpr int kMaxGroups = 128;      // workgroups of one gradient pass

struct MlpShape {
    int in, H;
    int oW1, ob1, oW2, ob2, oW3, ob3, count;
};

__host__ __device__ inline MlpShape mlp_shape(int in, int H) {
    MlpShape s;
    s.in = in; s.H = H;
    s.oW1 = 0;
    s.ob1 = s.oW1 + H * in;
    s.oW2 = s.ob1 + H;
    s.ob2 = s.oW2 + H * H;
    s.oW3 = s.ob2 + H;
    s.ob3 = s.oW3 + H;
    s.count = s.ob3 + 1;
    return s;
}

// Forward pass of one slice (samples s0 .. s0+ns-1) into LDS: h1, h2 [slice][H]
// and y [slice], with the parameters w already in LDS.
__device__ void mlp_forward_slice(const MlpShape sh, const float* __restrict__ w, const float* __restrict__ X,
                                  int s0, int ns, float* h1, float* h2, float* y) {
    const int tid = threadIdx.x, H = sh.H;
    const float* W1 = w + sh.oW1; const float* b1 = w + sh.ob1;
    const float* W2 = w + sh.oW2; const float* b2 = w + sh.ob2;
    const float* W3 = w + sh.oW3;
    for (int e = tid; e < ns * H; e += kMlpThreads) {
        const int s = e / H, j = e % H;
        float z = b1[j];
        for (int k = 0; k < sh.in; ++k) z += W1[j * sh.in + k] * X[(size_t)(s0 + s) * sh.in + k];
        h1[s * kMaxHidden + j] = tanhf(z);
    }
    __syncthreads();
    for (int e = tid; e < ns * H; e += kMlpThreads) {
        const int s = e / H, j = e % H;
        float z = b2[j];
        for (int i = 0; i < H; ++i) z += W2[j * H + i] * h1[s * kMaxHidden + i];
        h2[s * kMaxHidden + j] = tanhf(z);
    }
    __syncthreads();
    for (int s = tid; s < ns; s += kMlpThreads) {
        float z = w[sh.ob3];
        for (int j = 0; j < H; ++j) z += W3[j] * h2[s * kMaxHidden + j];
        y[s] = z;
    }
    __syncthreads();
}

__device__ void mlp_load_params(const MlpShape sh, const float* __restrict__ p, float* w) {
    for (int i = threadIdx.x; i < sh.count; i += kMlpThreads) w[i] = p[i];
    __syncthreads();
}

__global__ void __launch_bounds__(kMlpThreads) k_mlp_predict(MlpShape sh, const float* __restrict__ p,
                                                             const float* __restrict__ X, int B, float* out) {
    __shared__ float w[kMaxParams];
    __shared__ float h1[kMlpSlice * kMaxHidden], h2[kMlpSlice * kMaxHidden], y[kMlpSlice];
    mlp_load_params(sh, p, w);
    for (int s0 = blockIdx.x * kMlpSlice; s0 < B; s0 += gridDim.x * kMlpSlice) {
        const int ns = min(kMlpSlice, B - s0);
        mlp_forward_slice(sh, w, X, s0, ns, h1, h2, y);
        for (int s = threadIdx.x; s < ns; s += kMlpThreads) out[s0 + s] = y[s];
        __syncthreads();
    }
}

// Partial gradient of workgroup blockIdx.x over its slices: gpart[blockIdx.x][count],
// and its partial sum of squared errors lpart[blockIdx.x].
__global__ void __launch_bounds__(kMlpThreads) k_mlp_grad_part(MlpShape sh, const float* __restrict__ p,
                                                               const float* __restrict__ X,
                                                               const float* __restrict__ T, int B,
                                                               float* __restrict__ gpart, float* __restrict__ lpart) {
    __shared__ float w[kMaxParams];
    __shared__ float h1[kMlpSlice * kMaxHidden], h2[kMlpSlice * kMaxHidden], y[kMlpSlice];
    __shared__ float dz2[kMlpSlice * kMaxHidden], dz1[kMlpSlice * kMaxHidden], dy[kMlpSlice];
    __shared__ float xs[kMlpSlice * kMaxIn];
    const int tid = threadIdx.x, H = sh.H, in = sh.in;
    mlp_load_params(sh, p, w);
    const float* W2 = w + sh.oW2; const float* W3 = w + sh.oW3;
    constexpr int kPerThread = (kMaxParams + kMlpThreads - 1) / kMlpThreads;
    float acc[kPerThread];
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) acc[q] = 0.0f;
    float lsum = 0.0f;
    const float inv_b = 1.0f / (float)B;
    for (int s0 = blockIdx.x * kMlpSlice; s0 < B; s0 += gridDim.x * kMlpSlice) {
        const int ns = min(kMlpSlice, B - s0);
        mlp_forward_slice(sh, w, X, s0, ns, h1, h2, y);
        for (int e = tid; e < ns * in; e += kMlpThreads) xs[(e / in) * kMaxIn + e % in] = X[(size_t)s0 * in + e];
        for (int s = tid; s < ns; s += kMlpThreads) dy[s] = 2.0f * (y[s] - T[s0 + s]) * inv_b;
        if (tid == 0)
            for (int s = 0; s < ns; ++s) { const float d = y[s] - T[s0 + s]; lsum += d * d; }
        __syncthreads();
        for (int e = tid; e < ns * H; e += kMlpThreads) {            // through the output and tanh of layer 2
            const int s = e / H, j = e % H;
            const float a = h2[s * kMaxHidden + j];
            dz2[s * kMaxHidden + j] = dy[s] * W3[j] * (1.0f - a * a);
        }
        __syncthreads();
        for (int e = tid; e < ns * H; e += kMlpThreads) {            // back through W2 and tanh of layer 1
            const int s = e / H, i = e % H;
            float d = 0.0f;
            for (int j = 0; j < H; ++j) d += W2[j * H + i] * dz2[s * kMaxHidden + j];
            const float a = h1[s * kMaxHidden + i];
            dz1[s * kMaxHidden + i] = d * (1.0f - a * a);
        }
        __syncthreads();
        // parameter gradients, each summed over this slice's samples in order
#pragma unroll
        for (int q = 0; q < kPerThread; ++q) {
            const int idx = tid + q * kMlpThreads;
            if (idx >= sh.count) break;
            float s_acc = 0.0f;
            if (idx < sh.ob1) {                       // W1[j][k]
                const int j = idx / in, k = idx % in;
                for (int s = 0; s < ns; ++s) s_acc += dz1[s * kMaxHidden + j] * xs[s * kMaxIn + k];
            } else if (idx < sh.oW2) {                // b1[j]
                const int j = idx - sh.ob1;
                for (int s = 0; s < ns; ++s) s_acc += dz1[s * kMaxHidden + j];
            } else if (idx < sh.ob2) {                // W2[j][i]
                const int r = idx - sh.oW2, j = r / H, i = r % H;
                for (int s = 0; s < ns; ++s) s_acc += dz2[s * kMaxHidden + j] * h1[s * kMaxHidden + i];
            } else if (idx < sh.oW3) {                // b2[j]
                const int j = idx - sh.ob2;
                for (int s = 0; s < ns; ++s) s_acc += dz2[s * kMaxHidden + j];
            } else if (idx < sh.ob3) {                // W3[j]
                const int j = idx - sh.oW3;
                for (int s = 0; s < ns; ++s) s_acc += dy[s] * h2[s * kMaxHidden + j];
            } else {                                  // b3
                for (int s = 0; s < ns; ++s) s_acc += dy[s];
            }
            acc[q] += s_acc;
        }
        __syncthreads();
    }
    float* gp = gpart + (size_t)blockIdx.x * sh.count;
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) {
        const int idx = tid + q * kMlpThreads;
        if (idx < sh.count) gp[idx] = acc[q];
    }
    if (tid == 0) lpart[blockIdx.x] = lsum;
}

// g = sum of the G partial gradients in workgroup order; loss = sum of the
// partial squared errors / B.  With adam set, also one Adam step on p.
__global__ void k_mlp_reduce(int count, int G, const float* __restrict__ gpart, const float* __restrict__ lpart,
                             float inv_b, float* __restrict__ g, float* __restrict__ loss, bool adam,
                             float* __restrict__ p, float* __restrict__ m, float* __restrict__ v, float lr,
                             float beta1, float beta2, float eps, float bc1, float bc2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        float l = 0.0f;
        for (int w = 0; w < G; ++w) l += lpart[w];
        *loss = l * inv_b;
    }
    if (i >= count) return;
    float gi = 0.0f;
    for (int w = 0; w < G; ++w) gi += gpart[(size_t)w * count + i];
    g[i] = gi;
    if (adam) {
        const float mi = beta1 * m[i] + (1.0f - beta1) * gi;
        const float vi = beta2 * v[i] + (1.0f - beta2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    }
}

thread_local char g_mlp_err[256] = "";

void mlp_set_error(const char* msg) {
    strncpy(g_mlp_err, msg, sizeof(g_mlp_err) - 1);
    g_mlp_err[sizeof(g_mlp_err) - 1] = 0;
}

uint64_t splitmix64(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

}  // namespace

struct spires_mlp {
    MlpShape sh;
    float* p = nullptr;      // parameters
    float* g = nullptr;      // gradient
    float* m = nullptr;      // Adam first moment
    float* v = nullptr;      // Adam second moment
    float* loss = nullptr;   // [1]
    float* gpart = nullptr;  // [kMaxGroups x count] partial gradients
    float* lpart = nullptr;  // [kMaxGroups] partial squared errors
    float* X = nullptr;      // staging [cap x in]
    float* T = nullptr;      // staging [cap]
    float* y = nullptr;      // staging [cap]
    int cap = 0;
    long long step = 0;
    hipStream_t stream = nullptr;
};

#define MLP_CK(x)                                                   \
    do {                                                            \
        hipError_t e_ = (x);                                        \
        if (e_ != hipSuccess) {                                     \
            mlp_set_error(hipGetErrorString(e_));                   \
            return SPIRES_ERR_INTERNAL;                             \
        }                                                           \
    } while (0)

static int mlp_groups(int B) { return std::max(1, std::min(kMaxGroups, (B + kMlpSlice - 1) / kMlpSlice)); }

// Gradient of the staged batch into n->g and n->loss; with adam, one Adam step.
static void mlp_grad_launch(spires_mlp* n, int B, bool adam, float lr) {
    const int G = mlp_groups(B);
    hipLaunchKernelGGL(k_mlp_grad_part, dim3(G), dim3(kMlpThreads), 0, n->stream, n->sh, n->p, n->X, n->T, B,
                       n->gpart, n->lpart);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float bc1 = 1.0f, bc2 = 1.0f;
    if (adam) {
        ++n->step;
        bc1 = 1.0f - std::pow(b1, (float)n->step);
        bc2 = 1.0f - std::pow(b2, (float)n->step);
    }
    hipLaunchKernelGGL(k_mlp_reduce, dim3((n->sh.count + 255) / 256), dim3(256), 0, n->stream, n->sh.count, G,
                       n->gpart, n->lpart, 1.0f / (float)B, n->g, n->loss, adam, n->p, n->m, n->v, lr, b1, b2, eps,
                       bc1, bc2);
}

static spires_status mlp_stage(spires_mlp* n, int B) {
    if (B <= n->cap) return SPIRES_OK;
    (void)hipFree(n->X); (void)hipFree(n->T); (void)hipFree(n->y);
    n->X = n->T = n->y = nullptr;
    n->cap = 0;
    MLP_CK(hipMalloc(&n->X, (size_t)B * n->sh.in * sizeof(float)));
    MLP_CK(hipMalloc(&n->T, (size_t)B * sizeof(float)));
    MLP_CK(hipMalloc(&n->y, (size_t)B * sizeof(float)));
    n->cap = B;
    return SPIRES_OK;
}

extern "C" {

spires_status spires_ext_mlp_create(int in_dim, int hidden, unsigned long long seed, spires_mlp** out) {
    sdev::RandShield rand_shield_;
    if (!out || in_dim < 1 || in_dim > kMaxIn || hidden < 1 || hidden > kMaxHidden) {
        mlp_set_error("mlp: in_dim must be 1..16 and hidden 1..64");
        return SPIRES_ERR_INVALID_ARG;
    }
    *out = nullptr;
    spires_mlp* n = new spires_mlp;
    n->sh = mlp_shape(in_dim, hidden);
    // Glorot-uniform weights from a SplitMix64 stream, zero biases.
    std::vector<float> h(n->sh.count, 0.0f);
    uint64_t s = seed;
    auto uni = [&](float a) { return (float)((splitmix64(&s) >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0) * a; };
    const float a1 = std::sqrt(6.0f / (in_dim + hidden)), a2 = std::sqrt(6.0f / (2 * hidden)),
                a3 = std::sqrt(6.0f / (hidden + 1));
    for (int i = 0; i < hidden * in_dim; ++i) h[n->sh.oW1 + i] = uni(a1);
    for (int i = 0; i < hidden * hidden; ++i) h[n->sh.oW2 + i] = uni(a2);
    for (int i = 0; i < hidden; ++i) h[n->sh.oW3 + i] = uni(a3);
    const size_t bytes = (size_t)n->sh.count * sizeof(float);
    hipError_t e = hipSuccess;
    if ((e = hipStreamCreate(&n->stream)) != hipSuccess || (e = hipMalloc(&n->p, bytes)) != hipSuccess ||
        (e = hipMalloc(&n->g, bytes)) != hipSuccess || (e = hipMalloc(&n->m, bytes)) != hipSuccess ||
        (e = hipMalloc(&n->v, bytes)) != hipSuccess || (e = hipMalloc(&n->loss, sizeof(float))) != hipSuccess ||
        (e = hipMalloc(&n->gpart, bytes * kMaxGroups)) != hipSuccess ||
        (e = hipMalloc(&n->lpart, sizeof(float) * kMaxGroups)) != hipSuccess ||
        (e = hipMemcpy(n->p, h.data(), bytes, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemset(n->m, 0, bytes)) != hipSuccess || (e = hipMemset(n->v, 0, bytes)) != hipSuccess) {
        mlp_set_error(hipGetErrorString(e));
        spires_ext_mlp_destroy(n);
        return SPIRES_ERR_INTERNAL;
    }
    *out = n;
    return SPIRES_OK;
}

void spires_ext_mlp_destroy(spires_mlp* n) {
    sdev::RandShield rand_shield_;
    if (!n) return;
    (void)hipFree(n->p); (void)hipFree(n->g); (void)hipFree(n->m); (void)hipFree(n->v); (void)hipFree(n->loss);
    (void)hipFree(n->gpart); (void)hipFree(n->lpart);
    (void)hipFree(n->X); (void)hipFree(n->T); (void)hipFree(n->y);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    delete n;
}

int spires_ext_mlp_num_params(const spires_mlp* n) { return n ? n->sh.count : 0; }

spires_status spires_ext_mlp_get_params(spires_mlp* n, float* params) {
    sdev::RandShield rand_shield_;
    if (!n || !params) return SPIRES_ERR_INVALID_ARG;
    MLP_CK(hipMemcpyAsync(params, n->p, (size_t)n->sh.count * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    MLP_CK(hipStreamSynchronize(n->stream));
    return SPIRES_OK;
}

spires_status spires_ext_mlp_set_params(spires_mlp* n, const float* params) {
    sdev::RandShield rand_shield_;
    if (!n || !params) return SPIRES_ERR_INVALID_ARG;
    MLP_CK(hipMemcpyAsync(n->p, params, (size_t)n->sh.count * sizeof(float), hipMemcpyHostToDevice, n->stream));
    MLP_CK(hipMemsetAsync(n->m, 0, (size_t)n->sh.count * sizeof(float), n->stream));
    MLP_CK(hipMemsetAsync(n->v, 0, (size_t)n->sh.count * sizeof(float), n->stream));
    MLP_CK(hipStreamSynchronize(n->stream));
    n->step = 0;
    return SPIRES_OK;
}

spires_status spires_ext_mlp_predict(spires_mlp* n, const float* X, int B, float* y) {
    sdev::RandShield rand_shield_;
    if (!n || !X || !y || B < 0) return SPIRES_ERR_INVALID_ARG;
    if (B == 0) return SPIRES_OK;
    spires_status rc = mlp_stage(n, B);
    if (rc) return rc;
    MLP_CK(hipMemcpyAsync(n->X, X, (size_t)B * n->sh.in * sizeof(float), hipMemcpyHostToDevice, n->stream));
    hipLaunchKernelGGL(k_mlp_predict, dim3(mlp_groups(B)), dim3(kMlpThreads), 0, n->stream, n->sh, n->p, n->X, B,
                       n->y);
    MLP_CK(hipGetLastError());
    MLP_CK(hipMemcpyAsync(y, n->y, (size_t)B * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    MLP_CK(hipStreamSynchronize(n->stream));
    return SPIRES_OK;
}

spires_status spires_ext_mlp_gradient(spires_mlp* n, const float* X, const float* T, int B, float* grad,
                                      float* loss) {
    sdev::RandShield rand_shield_;
    if (!n || !X || !T || B < 1) return SPIRES_ERR_INVALID_ARG;
    spires_status rc = mlp_stage(n, B);
    if (rc) return rc;
    MLP_CK(hipMemcpyAsync(n->X, X, (size_t)B * n->sh.in * sizeof(float), hipMemcpyHostToDevice, n->stream));
    MLP_CK(hipMemcpyAsync(n->T, T, (size_t)B * sizeof(float), hipMemcpyHostToDevice, n->stream));
    mlp_grad_launch(n, B, false, 0.0f);
    MLP_CK(hipGetLastError());
    if (grad)
        MLP_CK(hipMemcpyAsync(grad, n->g, (size_t)n->sh.count * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    if (loss) MLP_CK(hipMemcpyAsync(loss, n->loss, sizeof(float), hipMemcpyDeviceToHost, n->stream));
    MLP_CK(hipStreamSynchronize(n->stream));
    return SPIRES_OK;
}

spires_status spires_ext_mlp_train(spires_mlp* n, const float* X, const float* T, int B, int epochs, float lr,
                                   float* loss) {
    sdev::RandShield rand_shield_;
    if (!n || !X || !T || B < 1 || epochs < 0) return SPIRES_ERR_INVALID_ARG;
    spires_status rc = mlp_stage(n, B);
    if (rc) return rc;
    MLP_CK(hipMemcpyAsync(n->X, X, (size_t)B * n->sh.in * sizeof(float), hipMemcpyHostToDevice, n->stream));
    MLP_CK(hipMemcpyAsync(n->T, T, (size_t)B * sizeof(float), hipMemcpyHostToDevice, n->stream));
    for (int e = 0; e < epochs; ++e) mlp_grad_launch(n, B, true, lr);
    MLP_CK(hipGetLastError());
    if (loss) MLP_CK(hipMemcpyAsync(loss, n->loss, sizeof(float), hipMemcpyDeviceToHost, n->stream));
    MLP_CK(hipStreamSynchronize(n->stream));
    return SPIRES_OK;
}

const char* spires_ext_mlp_last_error(void) { return g_mlp_err; }

}  // extern "C"
