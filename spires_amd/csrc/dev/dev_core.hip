// Device layer: reservoir state in HBM and the per-micro-step kernels.
//
// One thread per neuron.  The step fuses, for neuron i:
//   ext   = gain * (W_in[i,:] . u)                         reservoir.c:132-136
//   rec   = row i of W applied to the presynaptic signal    sparse.c:68-76 and synapses/*.c
//           (SIMPLE: CSR row-pull over the u8 spike vector, summing only the
//            spiking columns, which is bit-identical to the full sum because
//            products with 0/1 spikes are exact and the partial sum is never -0;
//            PSC traces are advanced inside the pull)
//   update of the neuron model with total = ext + rec     reservoir.c:158-160, neurons/*.c
// All FP64 operations appear in the reference's order and the unit is built
// with -ffp-contract=off, so results match the CPU path bit for bit.
#include "dev_internal.hpp"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

namespace sdev {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

namespace {
std::mutex g_shield_mu;
int g_shield_depth = 0;
char* g_shield_prev = nullptr;
char g_shield_state[256];   // the private generator state (glibc TYPE_4)
}  // namespace

RandShield::RandShield() {
    std::lock_guard<std::mutex> lk(g_shield_mu);
    if (g_shield_depth++ == 0) g_shield_prev = initstate(0x5eedu, g_shield_state, sizeof(g_shield_state));
}

RandShield::~RandShield() {
    std::lock_guard<std::mutex> lk(g_shield_mu);
    if (--g_shield_depth == 0 && g_shield_prev) setstate(g_shield_prev);
}
int hip_fail(hipError_t e, const char* what, int line) {
    set_error(std::string(what) + " failed (line " + std::to_string(line) + "): " + hipGetErrorString(e));
    (void)hipGetLastError();
    return e == hipErrorOutOfMemory ? SPIRES_DEV_EALLOC : SPIRES_DEV_EINTERNAL;
}

// ---------------------------------------------------------------- device code
__device__ __forceinline__ double syn_rec(const StepArgs& a, int i) {
    double s = 0.0;
    const int n = a.n;
    switch (a.stype) {
    case ST_SIMPLE:
        if (a.backend == BK_SPARSE) {
            for (int64_t k = a.rp[i]; k < a.rp[i + 1]; ++k)
                if (a.S_prev[a.col[k]]) s = s + a.val[k];
        } else {
            for (int j = 0; j < n; ++j)
                if (a.S_prev[j]) s = s + a.WT[(size_t)j * n + i];
        }
        break;
    case ST_PSC_HOM: {  // psc_homogeneous.c:71-84, trace recomputed per edge
        const double d = a.syn_decay, om = 1.0 - d;
        if (a.backend == BK_SPARSE) {
            for (int64_t k = a.rp[i]; k < a.rp[i + 1]; ++k) {
                const int j = a.col[k];
                const double t = a.tr_prev[j] * d + om * (double)a.S_prev[j];
                s = s + a.val[k] * t;
            }
        } else {
            for (int j = 0; j < n; ++j) {
                const double t = a.tr_prev[j] * d + om * (double)a.S_prev[j];
                s = s + a.WT[(size_t)j * n + i] * t;
            }
        }
        break;
    }
    case ST_PSC_HET: {  // psc_heterogeneous.c:124-139, 206-229 (per-edge trace owned by the row)
        if (a.backend == BK_SPARSE) {
            for (int64_t k = a.rp[i]; k < a.rp[i + 1]; ++k) {
                const double d = a.edge_decay[k];
                const double t = a.edge_tr[k] * d + (1.0 - d) * (double)a.S_prev[a.col[k]];
                a.edge_tr[k] = t;
                s = s + a.val[k] * t;
            }
        } else {
            for (int j = 0; j < n; ++j) {
                const size_t k = (size_t)j * n + i;
                const double d = a.edge_decay[k];
                const double t = a.edge_tr[k] * d + (1.0 - d) * (double)a.S_prev[j];
                a.edge_tr[k] = t;
                s = s + a.WT[k] * t;
            }
        }
        break;
    }
    default:  // fractional / multi-exponential: signal prepared by a separate pass
        if (a.backend == BK_SPARSE) {
            for (int64_t k = a.rp[i]; k < a.rp[i + 1]; ++k) s = s + a.val[k] * a.eff[a.col[k]];
        } else {
            for (int j = 0; j < n; ++j) s = s + a.WT[(size_t)j * n + i] * a.eff[j];
        }
        break;
    }
    return s;
}

// SplitMix64 (flif_diffusive.c:8-19).
__device__ __forceinline__ unsigned long long sm64(unsigned long long* x) {
    unsigned long long z = (*x += 0x9E3779B97f4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double sm_u01(unsigned long long* s) {
    return (double)(sm64(s) >> 11) * (1.0 / 9007199254740992.0);
}

// GL history sum H = sum_{k=1..lim} c_k * ring[(head-k) mod L][i], ascending k,
// one dependent add per lag (flif_gl.c:66-72).  Two contiguous slot ranges.
__device__ __forceinline__ double gl_history(const double* __restrict__ ring, const double* __restrict__ c,
                                             int n, int L, int head, int lim, int i) {
    double h = 0.0;
    int k = 1;
    const int r1 = lim < head ? lim : head;  // k in [1, r1] -> slot head-k
    for (; k + 7 <= r1; k += 8) {
        double v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = ring[(size_t)(head - k - j) * n + i];
#pragma unroll
        for (int j = 0; j < 8; ++j) h = h + c[k + j] * v[j];
    }
    for (; k <= r1; ++k) h = h + c[k] * ring[(size_t)(head - k) * n + i];
    for (; k + 7 <= lim; k += 8) {  // k in [head+1, lim] -> slot head-k+L
        double v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = ring[(size_t)(head - k - j + L) * n + i];
#pragma unroll
        for (int j = 0; j < 8; ++j) h = h + c[k + j] * v[j];
    }
    for (; k <= lim; ++k) h = h + c[k] * ring[(size_t)(head - k + L) * n + i];
    return h;
}

template <int NT>
__global__ void __launch_bounds__(256) k_step(StepArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    // external input (reservoir.c:132-136)
    double su = 0.0;
    if (a.has_u) {
        if (a.u) {
            for (int j = 0; j < a.din; ++j) su = su + a.W_in[(size_t)i * a.din + j] * a.u[j];
        } else {
            for (int j = 0; j < a.din; ++j) su = su + a.W_in[(size_t)i * a.din + j] * a.uin.v[j];
        }
    }
    const double ext = a.gain * su;
    const double rec = syn_rec(a, i);
    if (a.stype == ST_PSC_HOM) {
        const double d = a.syn_decay;
        a.tr_cur[i] = a.tr_prev[i] * d + (1.0 - d) * (double)a.S_prev[i];
    }
    const double x = ext + rec;
    const double* p = a.np;
    double V = a.V[i];
    bool spk = false;
    if constexpr (NT == NT_LIF_DISCRETE) {  // lif_discrete.c:22-34
        V = (1.0 - p[2]) * V + x + p[3];
        if (V > p[1]) { V = p[0]; spk = true; }
    } else if constexpr (NT == NT_LIF_BIO) {  // lif_bio.c:16-29
        const double dV = (-V + x + p[3]) / p[2];
        V = V + dV * a.dt;
        if (V >= p[1]) { V = p[0]; spk = true; }
    } else if constexpr (NT == NT_FLIF_GL) {  // flif_gl.c:50-100
        double tp = a.tprev[i];
        const int L = a.L, n = a.n, head = a.head;
        if (tp > p[6]) {
            const int sp = head == 0 ? L - 1 : head - 1;
            const double Vp = a.ring[(size_t)sp * n + i];
            const double H = gl_history(a.ring, a.coef, n, L, head, a.lim, i);
            const double rhs = (-(Vp - p[2]) / p[3]) + x + p[5];
            V = p[4] * rhs - H;
            if (V >= p[0]) { V = p[1]; spk = true; tp = 0.0; }
        } else {
            V = p[1];
        }
        a.ring[(size_t)head * n + i] = V;
        a.tprev[i] = tp + a.dt;
    } else if constexpr (NT == NT_FLIF_CAPUTO) {  // flif_caputo.c:43-71
        double tp = a.tprev[i];
        if (tp > p[6]) {
            const double dV = (-p[1] * (V - p[2]) + x) / p[0];
            const double mk = p[5] * dV;
            double mem = 0.0;
            const int m = a.m, n = a.n;
            if (m > 0) {
                int hd = a.chead[i] + 1;
                if (hd == m) hd = 0;
                a.chead[i] = hd;
                a.ring[(size_t)hd * n + i] = V - p[4];
                for (int j = 0; j < m; ++j) {
                    int s = hd - j;
                    if (s < 0) s += m;
                    mem = mem + a.coef[j] * a.ring[(size_t)s * n + i];
                }
            }
            V = V + (mk - mem);
            if (V >= p[3]) { spk = true; V = p[4]; tp = 0.0; }
        } else {
            V = p[4];
        }
        a.tprev[i] = tp + a.dt;
    } else {  // NT_FLIF_DIFFUSIVE, flif_diffusive.c:140-175
        const double Vp = V;
        const double rhs = -(Vp - p[2]) + x + p[4];
        V = Vp + p[3] * rhs;
        if (a.st >= a.until[i] && V >= p[0]) {
            V = p[1];
            spk = true;
            long long steps;
            if (p[5] < 1.0) {
                unsigned long long st = a.rng[i];
                double W;
                const double u0 = sm_u01(&st);
                if (u0 < p[9]) {
                    W = p[6];
                } else {
                    const double u1 = sm_u01(&st);
                    const double tmin = p[7], tmax = p[8], q = p[10];
                    if (tmax <= tmin) W = tmin;
                    else if (q <= 1e-12) W = tmin + u1 * (tmax - tmin);
                    else if (fabs(q - 1.0) < 1e-12) W = tmin * pow(tmax / tmin, u1);
                    else W = pow((1.0 - u1) * p[11] + u1 * p[12], p[13]);
                }
                a.rng[i] = st;
                steps = (W <= 0.0) ? 0 : (long long)ceil(W / a.dt);
                if (steps < 0) steps = 0;
                if (steps < 1) steps = 1;
            } else {
                const double W = p[6];
                steps = (W <= 0.0) ? 0 : (long long)ceil(W / a.dt);
                if (steps < 0) steps = 0;
            }
            a.until[i] = a.st + steps;
        }
    }
    a.V[i] = V;
    a.S_cur[i] = spk ? 1 : 0;
    if (a.rec_V) a.rec_V[i] = V;
    if (a.rec_S) a.rec_S[i] = spk ? 1 : 0;
}

// Fractional synapse current, fractional.c:103-142 (history ring [L][n]).
__global__ void k_prepare_frac(int n, int L, int head, int lim, double da, double tau,
                               const double* __restrict__ c, double* hist, const uint8_t* S,
                               double* I) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int sp = head == 0 ? L - 1 : head - 1;
    const double Ip = hist[(size_t)sp * n + i];
    double H = 0.0;
    for (int k = 1; k <= lim; ++k) {
        int s = head - k;
        if (s < 0) s += L;
        H = H + c[k] * hist[(size_t)s * n + i];
    }
    const double rhs = (-Ip + (double)S[i]) / tau;
    const double In = da * rhs - H;
    hist[(size_t)head * n + i] = In;
    I[i] = In;
}

struct MexpArgs { int n, nc; double d[64]; double c[64]; };
// Multi-exponential synapse, fractional_multiexp.c:136-157.
__global__ void k_prepare_mexp(MexpArgs a, double* tr, const uint8_t* S, double* comb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const double s = (double)S[i];
    for (int c = 0; c < a.nc; ++c) {
        const double d = a.d[c];
        double* t = &tr[(size_t)c * a.n + i];
        *t = *t * d + (1.0 - d) * s;
    }
    double sum = 0.0;
    for (int c = 0; c < a.nc; ++c) sum = sum + a.c[c] * tr[(size_t)c * a.n + i];
    comb[i] = sum;
}

// y[o] = W_out[o,:] . V  (reservoir.c:169-192); one block per output.
__global__ void __launch_bounds__(256) k_output(int n, const double* __restrict__ W,
                                                const double* __restrict__ V, double* y) {
    __shared__ double red[256];
    const int o = blockIdx.x;
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += W[(size_t)o * n + i] * V[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) y[o] = red[0];
}

// Neuron (re)initialisation, init_* of neurons/*.c.
__global__ void k_init(int n, int ntype, int L, int m, double v0, double tp0,
                       double ring_fill, const unsigned long long* rng0, double* V,
                       uint8_t* S0, uint8_t* S1, double* tprev, double* ring, int* chead,
                       long long* until, unsigned long long* rng) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    V[i] = v0;
    S0[i] = 0;
    S1[i] = 0;
    if (ntype == NT_FLIF_GL) {
        tprev[i] = tp0;
        for (int k = 0; k < L; ++k) ring[(size_t)k * n + i] = ring_fill;
    } else if (ntype == NT_FLIF_CAPUTO) {
        tprev[i] = tp0;
        chead[i] = 0;
        for (int k = 0; k < m; ++k) ring[(size_t)k * n + i] = 0.0;
    } else if (ntype == NT_FLIF_DIFFUSIVE) {
        until[i] = -1;
        rng[i] = rng0[i];
    }
}

__global__ void k_fill_u8(uint8_t* p, size_t n, uint8_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

static inline unsigned nblk(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

// ---------------------------------------------------------------- host side
static int init_neurons(spires_dev* g, const spires_dev_desc* d, bool full) {
    if (full) {
        g->d = *d;
    } else {  // reset: only the neuron model changes; synapse and input scalars stay
        memcpy(g->d.np, d->np, sizeof(g->d.np));
        g->d.L = d->L;
    }
    const int n = g->n;
    double v0 = 0.0, tp0 = 0.0, fill = 0.0;
    switch (g->ntype) {
    case NT_LIF_DISCRETE: case NT_LIF_BIO: v0 = d->np[0]; break;
    case NT_FLIF_GL: v0 = d->np[2]; fill = d->np[2]; tp0 = d->np[6] + d->dt; break;
    case NT_FLIF_CAPUTO: v0 = d->np[4]; tp0 = d->np[7]; break;
    case NT_FLIF_DIFFUSIVE: v0 = d->np[2]; break;
    }
    // (re)size the neuron history
    int L = 0, m = 0;
    size_t need = 0;
    if (g->ntype == NT_FLIF_GL) { L = d->L; need = (size_t)L * n; }
    if (g->ntype == NT_FLIF_CAPUTO) { L = d->L; m = L - 2 > 0 ? L - 2 : 0; need = (size_t)m * n; }
    if (need > g->ring_elems) {
        dfree(g->ring);
        int rc = dalloc(&g->ring, need);
        if (rc) return rc;
        g->ring_elems = need;
    }
    g->L = L;
    g->m = m;
    // model table: GL coefficients (kept from creation by the host) or Caputo weights
    int nco = g->ntype == NT_FLIF_GL ? L : (g->ntype == NT_FLIF_CAPUTO ? m : 0);
    if (nco > 0 && d->coef) {
        if (nco > g->ncoef) {
            dfree(g->coef);
            int rc = dalloc(&g->coef, (size_t)nco);
            if (rc) return rc;
            g->ncoef = nco;
        }
        SDEV_CK(hipMemcpyAsync(g->coef, d->coef, (size_t)nco * sizeof(double), hipMemcpyHostToDevice, g->stream));
    }
    unsigned long long* rng0 = nullptr;
    if (g->ntype == NT_FLIF_DIFFUSIVE) {
        int rc = dalloc(&rng0, (size_t)n);
        if (rc) return rc;
        SDEV_CK(hipMemcpyAsync(rng0, d->rng_state, (size_t)n * 8, hipMemcpyHostToDevice, g->stream));
    }
    hipLaunchKernelGGL(k_init, dim3(nblk(n, 256)), dim3(256), 0, g->stream, n, g->ntype, L, m, v0,
                       tp0, fill, rng0, g->V, g->S[0], g->S[1], g->tprev, g->ring, g->chead,
                       g->until, g->rng);
    SDEV_CK(hipGetLastError());
    if (rng0) {
        SDEV_CK(hipStreamSynchronize(g->stream));
        dfree(rng0);
    }
    if (g->xch) SDEV_CK(hipMemsetAsync(g->xch, 0, g->xch_words * 8, g->stream));
    g->ring32_valid = 0;
    g->ring64_valid = 1;
    g->cur = 0;
    g->step = 0;
    g->msteps = 0;
    return 0;
}

int launch_step(spires_dev* g, const double* u_dev, const double* u_host, double* rec_V,
                uint8_t* rec_S) {
    if (!g->ring64_valid) {
        int rcs = fast_sync_rings(g, false);
        if (rcs) return rcs;
        g->ring64_valid = 1;
    }
    g->ring32_valid = 0;
    StepArgs a;
    memset(&a, 0, sizeof(a));
    a.n = g->n; a.din = g->din; a.L = g->L; a.m = g->m;
    a.backend = g->backend; a.stype = g->stype;
    a.gain = g->d.gain; a.dt = g->d.dt;
    memcpy(a.np, g->d.np, sizeof(a.np));
    a.W_in = g->W_in;
    if (u_dev) { a.u = u_dev; a.has_u = 1; }
    else if (u_host) {
        if (g->din <= kMaxInlineInputs) {
            for (int j = 0; j < g->din; ++j) a.uin.v[j] = u_host[j];
        } else {
            SDEV_CK(hipMemcpyAsync(g->ustep, u_host, (size_t)g->din * 8, hipMemcpyHostToDevice, g->stream));
            SDEV_CK(hipStreamSynchronize(g->stream));
            a.u = g->ustep;
        }
        a.has_u = 1;
    }
    a.rp = g->rp; a.col = g->col; a.val = g->val; a.WT = g->WT;
    a.V = g->V; a.tprev = g->tprev; a.ring = g->ring; a.coef = g->coef;
    a.chead = g->chead; a.until = g->until; a.rng = g->rng;
    a.syn_decay = g->d.syn_decay;
    a.edge_decay = g->edge_decay; a.edge_tr = g->edge_tr;
    const unsigned B = 64;
    for (int mi = 0; mi < g->micro; ++mi) {
        const uint8_t* Sp = g->S[g->cur];
        uint8_t* Sc = g->S[g->cur ^ 1];
        if (g->stype == ST_FRACTIONAL) {
            const int L = g->frac_L;
            const long long fs = g->frac_step;
            const int head = (int)(fs % L);
            const int lim = fs < L ? (int)fs : L - 1;
            hipLaunchKernelGGL(k_prepare_frac, dim3(nblk(g->n, 256)), dim3(256), 0, g->stream, g->n, L,
                               head, lim, g->frac_da, g->frac_tau, g->frac_coef, g->frac_hist, Sp,
                               g->frac_I);
            SDEV_CK(hipGetLastError());
            g->frac_step++;
            a.eff = g->frac_I;
        } else if (g->stype == ST_MULTIEXP) {
            MexpArgs ma;
            ma.n = g->n; ma.nc = g->me_n;
            for (int c = 0; c < g->me_n; ++c) { ma.d[c] = g->me_decay[c]; ma.c[c] = g->me_c[c]; }
            hipLaunchKernelGGL(k_prepare_mexp, dim3(nblk(g->n, 256)), dim3(256), 0, g->stream, ma,
                               g->me_tr, Sp, g->me_comb);
            SDEV_CK(hipGetLastError());
            a.eff = g->me_comb;
        }
        a.S_prev = Sp; a.S_cur = Sc;
        a.tr_prev = g->tr[g->trcur]; a.tr_cur = g->tr[g->trcur ^ 1];
        a.st = g->step;
        if (g->ntype == NT_FLIF_GL) {
            a.head = (int)(g->step % g->L);
            a.lim = g->step < g->L ? (int)g->step : g->L - 1;
        }
        const bool last = (mi == g->micro - 1);
        a.rec_V = last ? rec_V : nullptr;
        a.rec_S = last ? rec_S : nullptr;
        const dim3 grid(nblk(g->n, B)), blk(B);
        switch (g->ntype) {
        case NT_LIF_DISCRETE: hipLaunchKernelGGL(k_step<NT_LIF_DISCRETE>, grid, blk, 0, g->stream, a); break;
        case NT_LIF_BIO: hipLaunchKernelGGL(k_step<NT_LIF_BIO>, grid, blk, 0, g->stream, a); break;
        case NT_FLIF_CAPUTO: hipLaunchKernelGGL(k_step<NT_FLIF_CAPUTO>, grid, blk, 0, g->stream, a); break;
        case NT_FLIF_GL: hipLaunchKernelGGL(k_step<NT_FLIF_GL>, grid, blk, 0, g->stream, a); break;
        default: hipLaunchKernelGGL(k_step<NT_FLIF_DIFFUSIVE>, grid, blk, 0, g->stream, a); break;
        }
        SDEV_CK(hipGetLastError());
        g->cur ^= 1;
        if (g->stype == ST_PSC_HOM) g->trcur ^= 1;
        g->step++;
    }
    g->msteps++;
    return 0;
}

int launch_output(spires_dev* g, double* y_dev) {
    if (g->dout <= 0) return 0;
    hipLaunchKernelGGL(k_output, dim3(g->dout), dim3(256), 0, g->stream, g->n, g->W_out, g->V, y_dev);
    SDEV_CK(hipGetLastError());
    return 0;
}

// Record buffers of at least `rows` rows; when one grows, its first `keep`
// rows are carried over (appending runs).
template <typename T>
static int grow_record(spires_dev* g, T** buf, size_t* cap, size_t need, size_t keep_elems) {
    if (need <= *cap) return 0;
    T* nb = nullptr;
    int rc = dalloc(&nb, need);
    if (rc) return rc;
    if (*buf && keep_elems) {
        hipError_t e = hipMemcpyAsync(nb, *buf, keep_elems * sizeof(T), hipMemcpyDeviceToDevice, g->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
        if (e != hipSuccess) { dfree(nb); return hip_fail(e, "record copy", __LINE__); }
    } else if (*buf) {
        (void)hipStreamSynchronize(g->stream);   // a queued launch may still write the old record
    }
    dfree(*buf);
    *buf = nb;
    *cap = need;
    return 0;
}

static int ensure_records(spires_dev* g, size_t rows, unsigned flags, size_t keep = 0) {
    int rc;
    keep = std::min(keep, g->rec_rows);
    if ((flags & SPIRES_DEV_REC_STATE) &&
        (rc = grow_record(g, &g->rec_state, &g->rec_cap_s, rows * g->n, keep * g->n)))
        return rc;
    if ((flags & SPIRES_DEV_REC_RASTER) &&
        (rc = grow_record(g, &g->rec_raster, &g->rec_cap_r, rows * g->n, keep * g->n)))
        return rc;
    if ((flags & SPIRES_DEV_REC_OUTPUT) &&
        (rc = grow_record(g, &g->rec_out, &g->rec_cap_o, rows * (size_t)g->dout, keep * (size_t)g->dout)))
        return rc;
    g->rec_rows = rows;
    return 0;
}

}  // namespace sdev

using namespace sdev;

extern "C" {

const char* spires_dev_last_error(void) { return g_err.c_str(); }

void spires_dev_destroy(spires_dev* g) {
    sdev::RandShield rand_shield_;
    if (!g) return;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    if (g->abort_host) (void)hipHostFree(g->abort_host);
    g->abort_host = nullptr;
    readout_release(g);
    dfree(g->W_in); dfree(g->W_out); dfree(g->rp); dfree(g->col); dfree(g->val); dfree(g->WT);
    dfree(g->coef); dfree(g->V); dfree(g->S[0]); dfree(g->S[1]); dfree(g->tprev); dfree(g->ring);
    dfree(g->chead); dfree(g->until); dfree(g->rng); dfree(g->tr[0]); dfree(g->tr[1]);
    dfree(g->edge_decay); dfree(g->edge_tr); dfree(g->frac_coef); dfree(g->frac_hist);
    dfree(g->frac_I); dfree(g->me_tr); dfree(g->me_comb); dfree(g->u_series);
    dfree(g->rec_state); dfree(g->rec_raster); dfree(g->rec_out); dfree(g->ybuf); dfree(g->ustep);
    dfree(g->xch); dfree(g->abort_flag); dfree(g->ring32); dfree(g->rbm); dfree(g->fseg);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
}

int spires_dev_create(const spires_dev_desc* d, spires_dev** out) {
    sdev::RandShield rand_shield_;
    if (!d || !out || d->n <= 0 || d->micro <= 0 || d->din < 0 || d->dout < 0) {
        set_error("invalid descriptor");
        return SPIRES_DEV_EINVAL;
    }
    if (d->synapse_type == ST_MULTIEXP && (d->me_n < 1 || d->me_n > 64)) {
        set_error("multi-exponential synapse supports 1..64 components");
        return SPIRES_DEV_EINVAL;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        set_error("no HIP device available");
        return SPIRES_DEV_EINTERNAL;
    }
    spires_dev* g = new spires_dev();
    g->d = *d;
    g->n = d->n; g->din = d->din; g->dout = d->dout; g->micro = d->micro;
    g->ntype = d->neuron_type; g->stype = d->synapse_type; g->backend = d->backend;
    int rc = 0;
    const size_t n = (size_t)g->n;
    auto fail = [&](int code) { spires_dev_destroy(g); return code; };
    if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        set_error("hipStreamCreate failed");
        delete g;
        return SPIRES_DEV_EINTERNAL;
    }
    if ((rc = dalloc(&g->W_in, n * g->din)) || (rc = dalloc(&g->W_out, n * g->dout)) ||
        (rc = dalloc(&g->V, n)) || (rc = dalloc(&g->S[0], n)) || (rc = dalloc(&g->S[1], n)) ||
        (rc = dalloc(&g->tprev, n)) || (rc = dalloc(&g->chead, n)) || (rc = dalloc(&g->until, n)) ||
        (rc = dalloc(&g->rng, n)) || (rc = dalloc(&g->ybuf, (size_t)g->dout + 1)) ||
        (rc = dalloc(&g->ustep, (size_t)g->din + 1)))
        return fail(rc);
    auto up = [&](void* dst, const void* src, size_t bytes) -> int {
        if (!bytes) return 0;
        hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g->stream);
        return e == hipSuccess ? 0 : hip_fail(e, "upload", __LINE__);
    };
    if ((rc = up(g->W_in, d->W_in, n * g->din * 8))) return fail(rc);
    if (d->W_out) { if ((rc = up(g->W_out, d->W_out, n * g->dout * 8))) return fail(rc); }
    else if (n * g->dout) { if (hipMemsetAsync(g->W_out, 0, n * g->dout * 8, g->stream) != hipSuccess) return fail(hip_fail(hipGetLastError(), "W_out clear", __LINE__)); }
    if (g->backend == BK_SPARSE) {
        g->nnz = d->nnz;
        g->rp_host.assign(d->row_ptr, d->row_ptr + n + 1);
        g->col_host.assign(d->col, d->col + d->nnz);
        if ((rc = dalloc(&g->rp, n + 1)) || (rc = dalloc(&g->col, (size_t)d->nnz)) || (rc = dalloc(&g->val, (size_t)d->nnz)))
            return fail(rc);
        if ((rc = up(g->rp, d->row_ptr, (n + 1) * 8)) || (rc = up(g->col, d->col, (size_t)d->nnz * 4)) ||
            (rc = up(g->val, d->val, (size_t)d->nnz * 8)))
            return fail(rc);
    } else {
        // store transposed: WT[j*n + i] = W[i*n + j], so a warp reads one column coalesced
        std::vector<double> WT(n * n);
        for (size_t i = 0; i < n; ++i)
            for (size_t j = 0; j < n; ++j) WT[j * n + i] = d->dense[i * n + j];
        if ((rc = dalloc(&g->WT, n * n))) return fail(rc);
        if ((rc = up(g->WT, WT.data(), n * n * 8))) return fail(rc);
        SDEV_CK(hipStreamSynchronize(g->stream));
    }
    // synapse state
    if (g->stype == ST_PSC_HOM && g->backend == BK_SPARSE) {
        // per-edge copies of the presynaptic trace: every copy of neuron j's trace sees
        // the same operations (tr*d + (1-d)*s_j from the same start), so each equals the
        // reference's per-neuron trace bit for bit, and the row pull needs only spike bits
        g->stype = ST_PSC_HET;
        const size_t m = (size_t)d->nnz;
        std::vector<double> dec(m ? m : 1, d->syn_decay);
        if ((rc = dalloc(&g->edge_decay, m)) || (rc = dalloc(&g->edge_tr, m))) return fail(rc);
        if ((rc = up(g->edge_decay, dec.data(), m * 8))) return fail(rc);
        if (m) SDEV_CK(hipMemsetAsync(g->edge_tr, 0, m * 8, g->stream));
        SDEV_CK(hipStreamSynchronize(g->stream));
    } else if (g->stype == ST_PSC_HOM) {
        if ((rc = dalloc(&g->tr[0], n)) || (rc = dalloc(&g->tr[1], n))) return fail(rc);
        SDEV_CK(hipMemsetAsync(g->tr[0], 0, n * 8, g->stream));
        SDEV_CK(hipMemsetAsync(g->tr[1], 0, n * 8, g->stream));
    } else if (g->stype == ST_PSC_HET) {
        const size_t m = g->backend == BK_SPARSE ? (size_t)d->nnz : n * n;
        if ((rc = dalloc(&g->edge_decay, m)) || (rc = dalloc(&g->edge_tr, m))) return fail(rc);
        if (g->backend == BK_SPARSE) {
            if ((rc = up(g->edge_decay, d->edge_decay, m * 8))) return fail(rc);
        } else {
            std::vector<double> T(n * n);
            for (size_t i = 0; i < n; ++i)
                for (size_t j = 0; j < n; ++j) T[j * n + i] = d->edge_decay[i * n + j];
            if ((rc = up(g->edge_decay, T.data(), m * 8))) return fail(rc);
            SDEV_CK(hipStreamSynchronize(g->stream));
        }
        if (m) SDEV_CK(hipMemsetAsync(g->edge_tr, 0, m * 8, g->stream));
    } else if (g->stype == ST_FRACTIONAL) {
        g->frac_L = d->frac_L; g->frac_da = d->frac_dt_alpha; g->frac_tau = d->frac_tau;
        if ((rc = dalloc(&g->frac_coef, (size_t)d->frac_L)) || (rc = dalloc(&g->frac_hist, n * d->frac_L)) ||
            (rc = dalloc(&g->frac_I, n)))
            return fail(rc);
        if ((rc = up(g->frac_coef, d->frac_coef, (size_t)d->frac_L * 8))) return fail(rc);
        SDEV_CK(hipMemsetAsync(g->frac_hist, 0, n * d->frac_L * 8, g->stream));
        SDEV_CK(hipMemsetAsync(g->frac_I, 0, n * 8, g->stream));
    } else if (g->stype == ST_MULTIEXP) {
        g->me_n = d->me_n;
        for (int c = 0; c < d->me_n; ++c) { g->me_decay[c] = d->me_decay[c]; g->me_c[c] = d->me_c[c]; }
        if ((rc = dalloc(&g->me_tr, n * d->me_n)) || (rc = dalloc(&g->me_comb, n))) return fail(rc);
        SDEV_CK(hipMemsetAsync(g->me_tr, 0, n * d->me_n * 8, g->stream));
    }
    if ((rc = init_neurons(g, d, true))) return fail(rc);
    {
        hipError_t e = hipStreamSynchronize(g->stream);
        if (e != hipSuccess) return fail(hip_fail(e, "create sync", __LINE__));
    }
    *out = g;
    return 0;
}

int spires_dev_reset_neurons(spires_dev* g, const spires_dev_desc* d) {
    sdev::RandShield rand_shield_;
    if (!g || !d) return SPIRES_DEV_EINVAL;
    if (const int rc = settle(g)) return rc;
    return init_neurons(g, d, false);
}

int spires_dev_step(spires_dev* g, const double* u) {
    sdev::RandShield rand_shield_;
    if (!g) return SPIRES_DEV_EINVAL;
    if (const int rc = settle(g)) return rc;
    static const double zero = 0.0;
    (void)zero;
    return launch_step(g, nullptr, u, nullptr, nullptr);
}

int spires_dev_upload_inputs(spires_dev* g, const double* u, size_t T) {
    sdev::RandShield rand_shield_;
    if (!g || (!u && T)) return SPIRES_DEV_EINVAL;
    if (const int rc = settle(g)) return rc;
    const size_t cnt = T * (size_t)g->din;
    if (T > g->u_T) {
        dfree(g->u_series);
        int rc = dalloc(&g->u_series, cnt + 1);
        if (rc) { g->u_T = 0; return rc; }
        g->u_T = T;
    }
    if (cnt) SDEV_CK(hipMemcpyAsync(g->u_series, u, cnt * 8, hipMemcpyHostToDevice, g->stream));
    SDEV_CK(hipStreamSynchronize(g->stream));
    return 0;
}

int spires_dev_run(spires_dev* g, size_t t0, size_t T, unsigned flags) {
    sdev::RandShield rand_shield_;
    if (!g) return SPIRES_DEV_EINVAL;
    if (t0 + T > g->u_T && g->din > 0) { set_error("run beyond uploaded inputs"); return SPIRES_DEV_EINVAL; }
    if (const int rc0 = settle_ready(g)) return rc0;
    const size_t base = (flags & SPIRES_DEV_REC_APPEND) ? t0 : 0;   // first record row of this run
    const size_t bn = base * (size_t)g->n;
    if (g->mode == SPIRES_DEV_MODE_FAST) {
        if (const int rcs = settle(g)) return rcs;
        if (!fast_supported(g)) { set_error("fast mode supports FLIF-GL with SIMPLE/PSC sparse synapses only"); return SPIRES_DEV_EINVAL; }
        unsigned needf = flags | ((flags & SPIRES_DEV_REC_OUTPUT) ? SPIRES_DEV_REC_STATE : 0u);
        int rcf = ensure_records(g, base + T, needf, base);
        if (rcf) return rcf;
        if (T == 0) return 0;
        const double* u = g->din > 0 ? g->u_series : nullptr;
        if ((rcf = launch_fast(g, u, (long long)t0, (long long)T,
                               (needf & SPIRES_DEV_REC_STATE) ? g->rec_state + bn : nullptr,
                               (flags & SPIRES_DEV_REC_RASTER) ? g->rec_raster + bn : nullptr)))
            return rcf;
        if (flags & SPIRES_DEV_REC_OUTPUT)
            return outputs_from_states(g, g->rec_state + bn, T, g->rec_out + base * g->dout);
        return 0;
    }
    if (!g->ring64_valid) {
        int rcs = fast_sync_rings(g, false);
        if (rcs) return rcs;
        g->ring64_valid = 1;
    }
    g->ring32_valid = 0;
    const bool persist = g->mode == SPIRES_DEV_MODE_EXACT && g->path != 1 && T > 0 && persist_supported(g);
    if (g->path >= 2 && !persist) { set_error("persistent kernel cannot run this reservoir"); return SPIRES_DEV_EINVAL; }
    if (!persist)   // the per-step path checks its launches as it goes
        if (const int rcs = settle(g)) return rcs;
    unsigned need = flags;
    if (persist && (flags & SPIRES_DEV_REC_OUTPUT)) need |= SPIRES_DEV_REC_STATE;   // outputs from the state record
    int rc = ensure_records(g, base + T, need, base);
    if (rc) return rc;
    if (persist) {
        const double* u = g->din > 0 ? g->u_series : nullptr;
        double* rv = (need & SPIRES_DEV_REC_STATE) ? g->rec_state + bn : nullptr;
        uint8_t* rs = (flags & SPIRES_DEV_REC_RASTER) ? g->rec_raster + bn : nullptr;
        if ((rc = launch_persist(g, u, (long long)t0, (long long)T, rv, rs))) return rc;
        if (flags & SPIRES_DEV_REC_OUTPUT)
            return outputs_from_states(g, g->rec_state + bn, T, g->rec_out + base * g->dout);
        return 0;
    }
    for (size_t t = 0; t < T; ++t) {
        const double* u = g->din > 0 ? g->u_series + (t0 + t) * g->din : nullptr;
        double* rv = (flags & SPIRES_DEV_REC_STATE) ? g->rec_state + (base + t) * g->n : nullptr;
        uint8_t* rs = (flags & SPIRES_DEV_REC_RASTER) ? g->rec_raster + (base + t) * g->n : nullptr;
        if ((rc = launch_step(g, u, nullptr, rv, rs))) return rc;
        if (flags & SPIRES_DEV_REC_OUTPUT)
            if ((rc = launch_output(g, g->rec_out + (base + t) * g->dout))) return rc;
    }
    return 0;
}

int spires_dev_run_batch(spires_dev* const* gs, int count, size_t t0, size_t T, unsigned flags) {
    sdev::RandShield rand_shield_;
    if (!gs || count < 1) return SPIRES_DEV_EINVAL;
    std::vector<double*> rv(count);
    std::vector<uint8_t*> rs(count);
    unsigned need = flags | ((flags & SPIRES_DEV_REC_OUTPUT) ? SPIRES_DEV_REC_STATE : 0u);
    const size_t base = (flags & SPIRES_DEV_REC_APPEND) ? t0 : 0;
    for (int r = 0; r < count; ++r) {
        spires_dev* g = gs[r];
        if (!g) return SPIRES_DEV_EINVAL;
        if (const int rcs = settle(g)) return rcs;
        if (g->mode != SPIRES_DEV_MODE_EXACT || g->path == 1) { set_error("batched runs are exact-mode persistent runs"); return SPIRES_DEV_EINVAL; }
        if (t0 + T > g->u_T && g->din > 0) { set_error("run beyond uploaded inputs"); return SPIRES_DEV_EINVAL; }
        if (!g->ring64_valid) {
            int rcs = fast_sync_rings(g, false);
            if (rcs) return rcs;
            g->ring64_valid = 1;
        }
        g->ring32_valid = 0;
        int rc = ensure_records(g, base + T, need, base);
        if (rc) return rc;
        rv[r] = (need & SPIRES_DEV_REC_STATE) ? g->rec_state + base * (size_t)g->n : nullptr;
        rs[r] = (flags & SPIRES_DEV_REC_RASTER) ? g->rec_raster + base * (size_t)g->n : nullptr;
    }
    if (T == 0) return 0;
    int rc = launch_persist_batch(gs, count, (long long)t0, (long long)T, rv.data(), rs.data());
    if (rc) return rc;
    if (flags & SPIRES_DEV_REC_OUTPUT)
        for (int r = 0; r < count; ++r)
            if ((rc = outputs_from_states(gs[r], rv[r], T, gs[r]->rec_out + base * gs[r]->dout))) return rc;
    return 0;
}

int spires_dev_batch_workgroups(spires_dev* const* gs, int count) {
    sdev::RandShield rand_shield_;
    int total = 0;
    if (!gs || count < 1 || batch_plan(gs, count, nullptr, &total)) return 0;
    return total;
}

int spires_dev_reserve_records(spires_dev* g, size_t rows, unsigned flags) {
    sdev::RandShield rand_shield_;
    if (!g) return SPIRES_DEV_EINVAL;
    if (const int rc = settle(g)) return rc;
    return ensure_records(g, rows, flags);
}

int spires_dev_read_records(spires_dev* g, size_t row0, size_t rows, double* states,
                            uint8_t* raster, double* outputs) {
    sdev::RandShield rand_shield_;
    if (!g || row0 + rows > g->rec_rows) return SPIRES_DEV_EINVAL;
    if (const int rc = settle(g)) return rc;
    const size_t n = g->n;
    if (states && g->rec_state)
        SDEV_CK(hipMemcpyAsync(states, g->rec_state + row0 * n, rows * n * 8, hipMemcpyDeviceToHost, g->stream));
    if (raster && g->rec_raster)
        SDEV_CK(hipMemcpyAsync(raster, g->rec_raster + row0 * n, rows * n, hipMemcpyDeviceToHost, g->stream));
    if (outputs && g->rec_out)
        SDEV_CK(hipMemcpyAsync(outputs, g->rec_out + row0 * g->dout, rows * g->dout * 8, hipMemcpyDeviceToHost, g->stream));
    SDEV_CK(hipStreamSynchronize(g->stream));
    return 0;
}

int spires_dev_read_state(spires_dev* g, double* V) {
    sdev::RandShield rand_shield_;
    if (!g || !V) return SPIRES_DEV_EINVAL;
    if (const int rc = settle(g)) return rc;
    SDEV_CK(hipMemcpyAsync(V, g->V, (size_t)g->n * 8, hipMemcpyDeviceToHost, g->stream));
    SDEV_CK(hipStreamSynchronize(g->stream));
    return 0;
}

int spires_dev_set_state(spires_dev* g, const double* V) {
    sdev::RandShield rand_shield_;
    if (!g || !V) return SPIRES_DEV_EINVAL;
    if (const int rc = settle(g)) return rc;
    SDEV_CK(hipMemcpyAsync(g->V, V, (size_t)g->n * 8, hipMemcpyHostToDevice, g->stream));
    SDEV_CK(hipStreamSynchronize(g->stream));
    return 0;
}

int spires_dev_read_spikes(spires_dev* g, double* S) {
    sdev::RandShield rand_shield_;
    if (!g || !S) return SPIRES_DEV_EINVAL;
    if (const int rc = settle(g)) return rc;
    std::vector<uint8_t> s(g->n);
    SDEV_CK(hipMemcpyAsync(s.data(), g->S[g->cur], (size_t)g->n, hipMemcpyDeviceToHost, g->stream));
    SDEV_CK(hipStreamSynchronize(g->stream));
    for (int i = 0; i < g->n; ++i) S[i] = s[i] ? 1.0 : 0.0;
    return 0;
}

int spires_dev_compute_output(spires_dev* g, double* y) {
    sdev::RandShield rand_shield_;
    if (!g || !y) return SPIRES_DEV_EINVAL;
    if (const int rc = settle(g)) return rc;
    int rc = launch_output(g, g->ybuf);
    if (rc) return rc;
    if (g->dout) SDEV_CK(hipMemcpyAsync(y, g->ybuf, (size_t)g->dout * 8, hipMemcpyDeviceToHost, g->stream));
    SDEV_CK(hipStreamSynchronize(g->stream));
    return 0;
}

int spires_dev_get_readout(spires_dev* g, double* W) {
    sdev::RandShield rand_shield_;
    if (!g || !W) return SPIRES_DEV_EINVAL;
    if (const int rc = settle(g)) return rc;
    const size_t cnt = (size_t)g->n * g->dout;
    if (cnt) SDEV_CK(hipMemcpyAsync(W, g->W_out, cnt * 8, hipMemcpyDeviceToHost, g->stream));
    SDEV_CK(hipStreamSynchronize(g->stream));
    return 0;
}

int spires_dev_set_readout(spires_dev* g, const double* W) {
    sdev::RandShield rand_shield_;
    if (!g || !W) return SPIRES_DEV_EINVAL;
    if (const int rc = settle(g)) return rc;
    const size_t cnt = (size_t)g->n * g->dout;
    if (cnt) SDEV_CK(hipMemcpyAsync(g->W_out, W, cnt * 8, hipMemcpyHostToDevice, g->stream));
    SDEV_CK(hipStreamSynchronize(g->stream));
    return 0;
}

int spires_dev_set_mode(spires_dev* g, int mode, int block_k) {
    sdev::RandShield rand_shield_;
    if (!g || (mode != SPIRES_DEV_MODE_EXACT && mode != SPIRES_DEV_MODE_FAST)) return SPIRES_DEV_EINVAL;
    if (const int rc = settle(g)) return rc;
    g->mode = mode;
    g->block_k = block_k;
    return 0;
}

int spires_dev_set_path(spires_dev* g, int path) {
    sdev::RandShield rand_shield_;
    if (!g || path < 0 || path > 4) return SPIRES_DEV_EINVAL;
    g->path = path;
    return 0;
}

int spires_dev_sync(spires_dev* g) {
    sdev::RandShield rand_shield_;
    if (!g) return SPIRES_DEV_EINVAL;
    if (const int rc = settle(g)) return rc;
    SDEV_CK(hipStreamSynchronize(g->stream));
    return 0;
}

void* spires_dev_stream(spires_dev* g) { return g ? (void*)g->stream : nullptr; }
size_t spires_dev_steps(const spires_dev* g) { return g ? g->msteps : 0; }

}  // extern "C"
