// Internal definitions shared by the device-layer translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <string>
#include <vector>

#include "spires_dev.h"

namespace sdev {

enum { NT_LIF_DISCRETE = 0, NT_LIF_BIO, NT_FLIF_CAPUTO, NT_FLIF_GL, NT_FLIF_DIFFUSIVE };
enum { ST_SIMPLE = 0, ST_PSC_HOM, ST_PSC_HET, ST_FRACTIONAL, ST_MULTIEXP };
enum { BK_SPARSE = 0, BK_DENSE = 1 };

void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what, int line);

// glibc's rand() sequence belongs to the caller: the reference's construction
// draws the weights from it (callers seed it with srand()), and the HIP runtime
// and rocBLAS draw from it on first use.  While a RandShield is alive, rand()
// reads a private state; the last one to end switches the caller's state back,
// at the position it had.  Every device-layer entry point holds one, so no GPU
// call moves the caller's sequence.  (Another thread's rand() calls made while
// a device call is inside the runtime also read the private state.)
class RandShield {
  public:
    RandShield();
    ~RandShield();
    RandShield(const RandShield&) = delete;
    RandShield& operator=(const RandShield&) = delete;
};

#define SDEV_CK(x)                                                        \
    do {                                                                  \
        hipError_t e_ = (x);                                              \
        if (e_ != hipSuccess) return ::sdev::hip_fail(e_, #x, __LINE__);  \
    } while (0)

template <typename T>
int dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        set_error(std::string("hipMalloc failed: ") + hipGetErrorString(e));
        return SPIRES_DEV_EALLOC;
    }
    return 0;
}
template <typename T>
void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

// Input vector passed by value for host-driven steps (no copy needed).
constexpr int kMaxInlineInputs = 32;
struct InlineU { double v[kMaxInlineInputs]; };

// Arguments of the per-micro-step kernel (passed by value).
struct StepArgs {
    int n, din, L, m;           // m: Caputo register length (L-2)
    int backend, stype;
    long long st;               // micro-step index since reset
    int head, lim;              // GL ring head and lag limit for this step
    double gain, dt;
    double np[16];
    const double* W_in;
    const double* u;            // device input row (nullptr -> inline)
    InlineU uin;
    int has_u;
    const int64_t* rp; const int32_t* col; const double* val; const double* WT;
    const uint8_t* S_prev; uint8_t* S_cur;
    const double* tr_prev; double* tr_cur; double syn_decay;
    const double* edge_decay; double* edge_tr;
    const double* eff;
    double* V; double* tprev; double* ring; const double* coef;
    int* chead; long long* until; unsigned long long* rng;
    double* rec_V; uint8_t* rec_S;
};

}  // namespace sdev

struct spires_dev {
    spires_dev_desc d;          // scalars (pointer members not retained)
    hipStream_t stream = nullptr;
    int n = 0, din = 0, dout = 0, micro = 1;
    int ntype = 0, stype = 0, backend = 0;
    int L = 0, m = 0;
    // weights
    double* W_in = nullptr; double* W_out = nullptr;
    int64_t* rp = nullptr; int32_t* col = nullptr; double* val = nullptr; int64_t nnz = 0;
    std::vector<int64_t> rp_host;   // host copy of the row pointers (launch planning)
    std::vector<int32_t> col_host;  // host copy of the column indices (row bitmaps)
    unsigned* rbm = nullptr;        // per-row column bitmaps [n][ceil(n/32)] (bitmap row pull)
    double* WT = nullptr;       // dense backend, transposed (column j contiguous over rows)
    // neuron state
    double* coef = nullptr; int ncoef = 0;
    double* V = nullptr; uint8_t* S[2] = {nullptr, nullptr}; int cur = 0;
    double* tprev = nullptr; double* ring = nullptr; size_t ring_elems = 0;
    int* chead = nullptr; long long* until = nullptr; unsigned long long* rng = nullptr;
    // synapse state (survives reset)
    double* tr[2] = {nullptr, nullptr}; int trcur = 0;
    double* edge_decay = nullptr; double* edge_tr = nullptr;
    double* frac_coef = nullptr; double* frac_hist = nullptr; double* frac_I = nullptr;
    int frac_L = 0; long long frac_step = 0; double frac_da = 0, frac_tau = 1;
    int me_n = 0; double me_decay[64]; double me_c[64]; double* me_tr = nullptr; double* me_comb = nullptr;
    // inputs and records
    double* u_series = nullptr; size_t u_T = 0;
    double* rec_state = nullptr; uint8_t* rec_raster = nullptr; double* rec_out = nullptr;
    size_t rec_rows = 0, rec_cap_s = 0, rec_cap_r = 0, rec_cap_o = 0;
    double* ybuf = nullptr;
    double* ustep = nullptr;    // device row for host steps with din > inline limit
    long long step = 0;         // micro-steps since reset
    size_t msteps = 0;          // macro steps since reset
    int mode = SPIRES_DEV_MODE_EXACT, block_k = 0;
    void* readout_ws = nullptr; // rocBLAS/rocSOLVER handles (readout unit)
    int path = 0;               // 0 auto, 1 per-step kernels, 2 persistent kernel
    unsigned long long* xch = nullptr; size_t xch_words = 0; int xch_plan_G = 0;
    unsigned* abort_flag = nullptr;   // raised by a persistent kernel whose exchange wait timed out
    unsigned* abort_host = nullptr;   // pinned copy of abort_flag after the last asynchronous launch
    bool abort_pending = false;       // an asynchronous launch's abort flag is still to be checked
    bool abort_dirty = true;          // abort_flag must be cleared before the next launch
    float* ring32 = nullptr; size_t ring32_cap = 0;   // FP32 history of the fast mode
    int* fseg = nullptr;        // fast mode's row-pull segments (dev_fast.hip)
    int fseg_max = 0;           // most segments in one workgroup
    int ring32_valid = 0, ring64_valid = 1;
};

namespace sdev {
int launch_step(spires_dev* g, const double* u_dev, const double* u_host, double* rec_V,
                uint8_t* rec_S);
int launch_output(spires_dev* g, double* y_dev);
void readout_release(spires_dev* g);
int persist_supported(spires_dev* g);
int launch_persist(spires_dev* g, const double* u_series, long long t0, long long T, double* rec_V,
                   uint8_t* rec_S);
// Several exact-mode FLIF-GL reservoirs in one launch of the product-wave kernel.
int batch_plan(spires_dev* const* gs, int count, int* nb_out, int* wgs_out);
int launch_persist_batch(spires_dev* const* gs, int count, long long t0, long long T, double* const* rec_V,
                         uint8_t* const* rec_S);
int outputs_from_states(spires_dev* g, const double* states, size_t T, double* out);
// Waits for g's stream when an asynchronous persistent launch is still to be
// checked, and reports its exchange timeout (then the flag is cleared before
// the next launch).  Every entry point that reads results, changes state or
// launches other kernels on g calls it first.
int settle(spires_dev* g);
// The same check without waiting: reports a timeout already copied back.
int settle_ready(spires_dev* g);
// Batched FP64 ridge systems of one order n (dev_chol.hip): A [count][n x n]
// column-major (the Gram fills both triangles; the factorisation keeps L in the
// lower and L' in the upper one), B / X / R [count][n x D], workspace of
// chol_batch_ws_doubles() doubles shared by the calls.
constexpr int kCholMaxN = 6144;
constexpr int kCholRhsMax = 8;     // right-hand sides handled by the cross / solve / residual kernels
bool chol_batch_ok(int n);
size_t chol_batch_ws_doubles(int count, int n, int T, int D);
hipError_t gram_cross_batched(hipStream_t st, int count, int n, int T, int D, const double* const* Xs,
                              const double* Yd, const double* lam, double* A, size_t strideA, double* B,
                              size_t strideB, double* ws, bool with_cross);
hipError_t chol_batched(hipStream_t st, int count, int n, double* A, size_t strideA, double* ws, int* info);
hipError_t chol_solve_batched(hipStream_t st, int count, int n, int D, const double* A, size_t strideA,
                              const double* ws, double* X, size_t strideX);
// The same solve for any number of right-hand sides: blocked by 64 rows x 64 columns,
// two launches per block and direction.
hipError_t chol_solve_many(hipStream_t st, int count, int n, int D, const double* A, size_t strideA,
                           const double* ws, double* X, size_t strideX);
// R = B - A X with A the full symmetric (unfactored) matrices.
hipError_t resid_batched(hipStream_t st, int count, int n, int D, const double* A, size_t strideA, const double* X,
                         const double* B, double* R, size_t strideX, double* ws);
// Split-BF16 Gram on the bf16 matrix cores (dev_gram.hip): lower triangle of
// hi'hi + hi'lo + lo'hi + lambda I for the state record X [T x n]; ws holds
// gram_split_ws_elems(T, n) bf16.  n must be a multiple of 128.
int gram_split_bf16(hipStream_t st, const double* X, int T, int n, double lambda, double* G, unsigned short* ws);
size_t gram_split_ws_elems(int T, int n);
int fast_supported(spires_dev* g);
int fast_sync_rings(spires_dev* g, bool to_fast);
int launch_fast(spires_dev* g, const double* u_series, long long t0, long long T, double* rec_V,
                uint8_t* rec_S);
}  // namespace sdev
