// FP64 ridge systems in exact mode: one or many independent normal equations
// (X_r'X_r + lambda_r I) W_r = X_r'Y_r of one order n (a multiple of 64), solved
// together so that every launch spans all systems (the readout fits of a
// hyperparameter search, or a single readout fit).
//
//  - Gram: lower-triangle tiles of X'X on the FP64 matrix cores
//    (v_mfma_f64_16x16x4f64, operands straight from the recorded state rows,
//    which lie n-contiguous per time step), lambda added on the diagonal; each
//    tile is stored twice (lower and mirrored upper), so the unfactored copy
//    is a full symmetric matrix for the refinement residual and LU fallback.
//  - Cross products X'Y: per-thread row sums over chunks of time steps, the
//    chunks then added in a fixed order (no atomics: results do not depend on
//    scheduling).
//  - Cholesky: right-looking, blocks of 64.  The diagonal block is factored by
//    one wave in registers and inverted (forward substitution against I), so the
//    panel below it is one 64x64x64 product with the inverse; the trailing lower
//    triangle is updated tile by tile.
//  - The factor kernels also store L' in the (otherwise unused) upper
//    triangle, so both triangular solves stream coalesced column segments.
//  - Triangular solves, up to 8 right-hand sides: one workgroup per (system,
//    right-hand side); the diagonal blocks apply their inverses.  More
//    right-hand sides: blocked by 64 rows x 64 columns, one launch applying a
//    diagonal block's inverse and one subtracting that block's product from
//    the row blocks still to solve.
//  - Residual B - A X of the refinement: a batched GEMV over the full
//    symmetric copy, column chunks summed in a fixed order.
//
// Factorisation order and rounding differ from LAPACK's; callers refine the
// solution against the unfactored matrix.  Reference: the normal equations of
// reservoir.c:573-679 (ridge readout); the reference solves each output
// separately with LU.
#include "dev_internal.hpp"

#include <algorithm>
#include <cmath>

namespace sdev {

namespace {

constexpr int CB = 64;                 // Cholesky block order
constexpr int kGramBatch = 16;         // systems per Gram launch (pointers passed by value)

typedef double d4 __attribute__((ext_vector_type(4)));

struct GramArgs {
    const double* X[kGramBatch];       // state rows of system s: X[t * n + i]
    double lam[kGramBatch];
};

__device__ __forceinline__ void tri_decode(int t, int& bi, int& bj) {
    // lower-triangle index t -> (bi >= bj), row-major over the triangle
    int b = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((b + 1) * (b + 2) / 2 <= t) ++b;
    while (b * (b + 1) / 2 > t) --b;
    bi = b;
    bj = t - b * (b + 1) / 2;
}

// G = X'X over T rows, lower triangle, TILE x TILE per workgroup (4 waves, each
// a (TILE/2)^2 quarter).  MFMA operands: A[p][k] = X[t+k][j0+p] and B[k][q] =
// X[t+k][i0+q] (lane l holds row k = l/16, column l%16 of each), so the result
// D[p][q] = G[j0+p][i0+q]: register r of lane l holds p = 4r + l/16, q = l%16,
// and its stores to column-major A (element (i, j) at j*n + i) run along i.
template <int TILE>
__global__ __launch_bounds__(256, 2) void k_gram_lower(GramArgs ga, int n, int T, double* A, size_t strideA) {
    constexpr int WT = TILE / 2, M = WT / 16;
    const int sys = blockIdx.y;
    int bi, bj;
    tri_decode(blockIdx.x, bi, bj);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i0 = bi * TILE + (wave >> 1) * WT;
    const int j0 = bj * TILE + (wave & 1) * WT;
    if (j0 > i0 + WT - 1) return;      // quarter entirely above the diagonal
    const int li = lane & 15, lk = lane >> 4;
    const double* X = ga.X[sys];
    const double* pa = X + (size_t)lk * n + j0 + li;
    const double* pb = X + (size_t)lk * n + i0 + li;
    d4 acc[M][M];
#pragma unroll
    for (int m = 0; m < M; ++m)
#pragma unroll
        for (int q = 0; q < M; ++q) acc[m][q] = d4{0.0, 0.0, 0.0, 0.0};
    int t = 0;
    for (; t + 4 <= T; t += 4) {
        const size_t off = (size_t)t * n;
        double a[M], b[M];
#pragma unroll
        for (int m = 0; m < M; ++m) {
            a[m] = pa[off + 16 * m];
            b[m] = pb[off + 16 * m];
        }
#pragma unroll
        for (int m = 0; m < M; ++m)
#pragma unroll
            for (int q = 0; q < M; ++q) acc[m][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[m], b[q], acc[m][q], 0, 0, 0);
    }
    if (t < T) {                       // last 1-3 rows: missing rows contribute zero
        const bool ok = t + lk < T;
        const size_t off = ok ? (size_t)t * n : 0;
        double a[M], b[M];
#pragma unroll
        for (int m = 0; m < M; ++m) {
            a[m] = ok ? pa[off + 16 * m] : 0.0;
            b[m] = ok ? pb[off + 16 * m] : 0.0;
        }
#pragma unroll
        for (int m = 0; m < M; ++m)
#pragma unroll
            for (int q = 0; q < M; ++q) acc[m][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[m], b[q], acc[m][q], 0, 0, 0);
    }
    double* As = A + (size_t)sys * strideA;
    const double lam = ga.lam[sys];
    const bool diag = i0 == j0;        // a quarter on the diagonal holds both triangles
#pragma unroll
    for (int m = 0; m < M; ++m)
#pragma unroll
        for (int q = 0; q < M; ++q)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + 16 * q + li, j = j0 + 16 * m + 4 * r + lk;
                const double v = acc[m][q][r] + (i == j ? lam : 0.0);
                As[(size_t)j * n + i] = v;
                if (!diag) As[(size_t)i * n + j] = v;   // mirror (i > j here)
            }
}

// Partial cross products P[s][c][d][i] = sum over rows t of chunk c of X[t][i] Y[t][d],
// for outputs d0 .. d0+DP-1.
constexpr int DP = kCholRhsMax;
__global__ __launch_bounds__(256) void k_cross_part(GramArgs ga, int n, int T, int D, int d0, int chunk,
                                                    const double* Y, size_t strideY, double* P, int nch) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int c = blockIdx.y, sys = blockIdx.z;
    if (i >= n) return;
    const double* X = ga.X[sys];
    const double* Ys = Y + (size_t)sys * strideY;
    const int dn = min(DP, D - d0);
    double acc[DP];
#pragma unroll
    for (int d = 0; d < DP; ++d) acc[d] = 0.0;
    const int t1 = min(T, (c + 1) * chunk);
    for (int t = c * chunk; t < t1; ++t) {
        const double x = X[(size_t)t * n + i];
#pragma unroll
        for (int d = 0; d < DP; ++d)
            if (d < dn) acc[d] = fma(x, Ys[(size_t)t * D + d0 + d], acc[d]);
    }
    double* Ps = P + ((size_t)sys * nch + c) * DP * n;
#pragma unroll
    for (int d = 0; d < DP; ++d)
        if (d < dn) Ps[(size_t)d * n + i] = acc[d];
}

// B[s][(d0+d) * n + i] = sum over chunks (in order) of the partials.
__global__ __launch_bounds__(256) void k_cross_sum(int n, int D, int d0, const double* P, int nch, double* B,
                                                   size_t strideB) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int d = blockIdx.y, sys = blockIdx.z;
    if (i >= n || d0 + d >= D) return;
    double s = 0.0;
    for (int c = 0; c < nch; ++c) s += P[(((size_t)sys * nch + c) * DP + d) * n + i];
    B[(size_t)sys * strideB + (size_t)(d0 + d) * n + i] = s;
}

__global__ void k_zero_int(int* p, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0;
}

__device__ __forceinline__ double lane_bcast(double x, int l) {   // x of lane l (l wave-uniform)
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Factor the diagonal block kb of each system (lower, in place; L' into the upper half
// of the block) and store the inverse of its factor, row-major, in Linv[s][kb].  One
// wave: lane r keeps row r of the block in registers and takes column entries of other
// rows by readlane, so the 64 column steps need no barrier.  Entries above the
// diagonal are updated too but never stored, and never feed an entry below it.  The
// inverse: lane c computes column c, Li[i][c] = (delta_ic - sum_{m<i} L[i][m] Li[m][c])
// / L[i][i], with the factor's rows broadcast from LDS.  info[s] records the first
// non-positive pivot (1-based column) and is never cleared here.
__global__ __launch_bounds__(CB) void k_chol_diag(int n, int kb, double* A, size_t strideA, double* Linv, int* info) {
    __shared__ double L[CB][CB + 1];
    const int sys = blockIdx.x, lane = threadIdx.x;
    double* As = A + (size_t)sys * strideA;
    const int k = kb * CB;
    double row[CB];
#pragma unroll
    for (int c = 0; c < CB; ++c) row[c] = As[(size_t)(k + c) * n + k + lane];
    int bad = 0;
#pragma unroll
    for (int j = 0; j < CB; ++j) {
        const double d = lane_bcast(row[j], j);
        const bool ok = d > 0.0;                     // false for NaN too
        const double s = ok ? sqrt(d) : 1.0;
        if (!ok && bad == 0) bad = k + j + 1;
        row[j] = lane == j ? s : row[j] / s;
#pragma unroll
        for (int c = j + 1; c < CB; ++c) row[c] = fma(-row[j], lane_bcast(row[j], c), row[c]);
    }
#pragma unroll
    for (int c = 0; c < CB; ++c) L[lane][c] = c <= lane ? row[c] : 0.0;
    __syncthreads();
    for (int e = lane; e < CB * CB; e += CB) {
        const int c = e / CB, r = e % CB;
        As[(size_t)(k + c) * n + k + r] = r >= c ? L[r][c] : L[c][r];   // L below, L' above
    }
    double li[CB];
#pragma unroll
    for (int i = 0; i < CB; ++i) {
        double s = i == lane ? 1.0 : 0.0;
#pragma unroll
        for (int m = 0; m < i; ++m) s = fma(-L[i][m], li[m], s);
        li[i] = s / L[i][i];
    }
    double* Lv = Linv + ((size_t)sys * (n / CB) + kb) * CB * CB;
#pragma unroll
    for (int i = 0; i < CB; ++i) Lv[(size_t)i * CB + lane] = li[i];
    if (lane == 0 && bad != 0 && info[sys] == 0) info[sys] = bad;
}

// Panel below diagonal block kb: L[ib][kb] = A[ib][kb] * Linv_kb^T for ib > kb.
__global__ __launch_bounds__(256) void k_chol_panel(int n, int kb, double* A, size_t strideA, const double* Linv) {
    __shared__ double P[CB][CB + 1];
    __shared__ double V[CB][CB + 1];
    const int sys = blockIdx.y, tid = threadIdx.x;
    const int k = kb * CB, i0 = (kb + 1 + blockIdx.x) * CB;
    double* As = A + (size_t)sys * strideA;
    const double* Lv = Linv + ((size_t)sys * (n / CB) + kb) * CB * CB;
    for (int e = tid; e < CB * CB; e += 256) {
        const int c = e / CB, r = e % CB;
        P[r][c] = As[(size_t)(k + c) * n + i0 + r];
        V[c][r] = Lv[e];                             // V[row][col] = Linv[row][col]
    }
    __syncthreads();
    const int tr = tid & 15, tc = tid >> 4;
    double acc[4][4] = {};
    for (int m = 0; m < CB; ++m) {
        double p[4], v[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) p[a] = P[tr + 16 * a][m];
#pragma unroll
        for (int b = 0; b < 4; ++b) v[b] = V[tc + 16 * b][m];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = fma(p[a], v[b], acc[a][b]);
    }
    __syncthreads();                                 // P is reused for the transposed store
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            As[(size_t)(k + tc + 16 * b) * n + i0 + tr + 16 * a] = acc[a][b];
            P[tr + 16 * a][tc + 16 * b] = acc[a][b];
        }
    __syncthreads();
    for (int e = tid; e < CB * CB; e += 256) {      // L' block: row k + c, column i0 + r
        const int r = e / CB, c = e % CB;
        As[(size_t)(i0 + r) * n + k + c] = P[r][c];
    }
}

// Trailing update after block kb: A[ib][jb] -= L[ib][kb] L[jb][kb]^T for kb < jb <= ib, on
// the FP64 matrix cores, operands straight from the factor (column-major: element
// (row, col) at col * n + row).  Wave w takes the 32x32 quarter rows r0 = 32(w/2), columns
// c0 = 32(w%2) of the tile.  MFMA operands A[p][k] = L[j0+c0+p][k+m0+k'] and
// B[k][q] = L[i0+r0+q][k+m0+k'], so D[p][q] is the product for column c0+p, row r0+q:
// register t of lane l holds p = 4t + l/16, q = l%16, and the read-modify-write of each
// register runs along 16 contiguous rows.
__global__ __launch_bounds__(256) void k_chol_update(int n, int kb, double* A, size_t strideA) {
    const int sys = blockIdx.y;
    int bi, bj;
    tri_decode(blockIdx.x, bi, bj);
    const int k = kb * CB, i0 = (kb + 1 + bi) * CB, j0 = (kb + 1 + bj) * CB;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r0 = (wave >> 1) * 32, c0 = (wave & 1) * 32;
    if (i0 == j0 && c0 > r0) return;            // quarter above the diagonal of a diagonal tile
    const int li = lane & 15, lk = lane >> 4;
    double* As = A + (size_t)sys * strideA;
    const double* pa = As + (size_t)(k + lk) * n + j0 + c0 + li;   // L[j0+c0+p][k+m], p = li (+16)
    const double* pb = As + (size_t)(k + lk) * n + i0 + r0 + li;   // L[i0+r0+q][k+m], q = li (+16)
    d4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
    for (int m0 = 0; m0 < CB; m0 += 4) {
        const size_t off = (size_t)m0 * n;
        const double a0 = pa[off], a1 = pa[off + 16], b0 = pb[off], b1 = pb[off + 16];
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int c = c0 + 16 * a + 4 * t + lk, r = r0 + 16 * b + li;
                if (i0 != j0 || r >= c) {
                    double* q = As + (size_t)(j0 + c) * n + i0 + r;
                    *q = *q - acc[a][b][t];
                }
            }
}

constexpr int kSolveThreads = 1024;

// x <- L^-1 x for column d = blockIdx.y of system blockIdx.x.
__global__ __launch_bounds__(kSolveThreads) void k_chol_fwd(int n, const double* A, size_t strideA,
                                                            const double* Linv, double* Xv, size_t strideX) {
    extern __shared__ double z[];
    __shared__ double part[kSolveThreads / CB][CB];
    __shared__ double rb[CB];
    const int sys = blockIdx.x, tid = threadIdx.x;
    const int nbk = n / CB;
    const double* As = A + (size_t)sys * strideA;
    double* x = Xv + (size_t)sys * strideX + (size_t)blockIdx.y * n;
    constexpr int PH = kSolveThreads / CB;
    const int r = tid % CB, ph = tid / CB;
    for (int kb = 0; kb < nbk; ++kb) {
        const int k = kb * CB;
        double s = 0.0;
#pragma unroll 4
        for (int c = ph; c < k; c += PH) s = fma(As[(size_t)c * n + k + r], z[c], s);
        part[ph][r] = s;
        __syncthreads();
        if (tid < CB) {
            double v = x[k + tid];
            for (int p = 0; p < PH; ++p) v -= part[p][tid];
            rb[tid] = v;
        }
        __syncthreads();
        if (tid < CB) {
            const double* Lv = Linv + ((size_t)sys * nbk + kb) * CB * CB + (size_t)tid * CB;
            double v = 0.0;
            for (int m = 0; m <= tid; ++m) v = fma(Lv[m], rb[m], v);
            z[k + tid] = v;
        }
        __syncthreads();
    }
    for (int i = tid; i < n; i += kSolveThreads) x[i] = z[i];
}

// x <- L^-T x for column d = blockIdx.y of system blockIdx.x, reading L' from
// the upper triangle (element (i, c), i < c, at c * n + i).
__global__ __launch_bounds__(kSolveThreads) void k_chol_bwd(int n, const double* A, size_t strideA,
                                                            const double* Linv, double* Xv, size_t strideX) {
    extern __shared__ double w[];
    constexpr int PH = kSolveThreads / CB;
    __shared__ double part[PH][CB];
    __shared__ double rb[CB];
    const int sys = blockIdx.x, tid = threadIdx.x;
    const int nbk = n / CB;
    const double* As = A + (size_t)sys * strideA;
    double* x = Xv + (size_t)sys * strideX + (size_t)blockIdx.y * n;
    const int r = tid % CB, ph = tid / CB;
    for (int kb = nbk - 1; kb >= 0; --kb) {
        const int k = kb * CB;
        // sum over c >= k + CB of L'[k + r][c] w[c]
        double s = 0.0;
#pragma unroll 4
        for (int c = k + CB + ph; c < n; c += PH) s = fma(As[(size_t)c * n + k + r], w[c], s);
        part[ph][r] = s;
        __syncthreads();
        if (tid < CB) {
            double v = x[k + tid];
            for (int p = 0; p < PH; ++p) v -= part[p][tid];
            rb[tid] = v;
        }
        __syncthreads();
        if (tid < CB) {
            const double* Lv = Linv + ((size_t)sys * nbk + kb) * CB * CB;
            double v = 0.0;
            for (int m = tid; m < CB; ++m) v = fma(Lv[(size_t)m * CB + tid], rb[m], v);
            w[k + tid] = v;
        }
        __syncthreads();
    }
    for (int i = tid; i < n; i += kSolveThreads) x[i] = w[i];
}

// Blocked triangular solves for many right-hand sides.  k_trsm_diag: rows of block kb,
// 64 columns per workgroup, X <- M X with M = Linv_kb (forward) or Linv_kb' (backward).
__global__ __launch_bounds__(256) void k_trsm_diag(int n, int D, int kb, int trans, const double* Linv, double* X,
                                                   size_t strideX) {
    __shared__ double Bs[CB][CB + 1];                // Bs[m][c] = X(k + m, d0 + c)
    __shared__ double V[CB][CB + 1];                 // V[i][m] = M[i][m]
    const int d0 = blockIdx.x * CB, sys = blockIdx.y, tid = threadIdx.x;
    const int nbk = n / CB, k = kb * CB;
    const double* Lv = Linv + ((size_t)sys * nbk + kb) * CB * CB;
    double* Xs = X + (size_t)sys * strideX;
    for (int e = tid; e < CB * CB; e += 256) {
        const int c = e / CB, m = e % CB;
        Bs[m][c] = d0 + c < D ? Xs[(size_t)(d0 + c) * n + k + m] : 0.0;
        V[c][m] = trans ? Lv[(size_t)m * CB + c] : Lv[e];
    }
    __syncthreads();
    const int tr = tid & 15, tc = tid >> 4;
    double acc[4][4] = {};
    for (int m = 0; m < CB; ++m) {
        double p[4], v[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) p[a] = V[tr + 16 * a][m];
#pragma unroll
        for (int b = 0; b < 4; ++b) v[b] = Bs[m][tc + 16 * b];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = fma(p[a], v[b], acc[a][b]);
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const int c = tc + 16 * b;
        if (d0 + c < D)
#pragma unroll
            for (int a = 0; a < 4; ++a) Xs[(size_t)(d0 + c) * n + k + tr + 16 * a] = acc[a][b];
    }
}

// X rows of block ib = ib0 + blockIdx.x minus A-block (ib, kb) times X rows of block kb,
// 64 columns per workgroup.  The A-block is L[ib][kb] below the diagonal (forward) and
// L'[ib][kb] above it (backward); both sit at column-major (kb*64 + m) * n + ib*64 + r.
__global__ __launch_bounds__(256) void k_trsm_update(int n, int D, int kb, int ib0, const double* A, size_t strideA,
                                                     double* X, size_t strideX) {
    __shared__ double P[CB][CB + 1];                 // P[r][m] = A(ib*64 + r, kb*64 + m)
    __shared__ double Bs[CB][CB + 1];                // Bs[m][c] = X(kb*64 + m, d0 + c)
    const int ib = ib0 + blockIdx.x, d0 = blockIdx.y * CB, sys = blockIdx.z, tid = threadIdx.x;
    const int k = kb * CB, i0 = ib * CB;
    const double* As = A + (size_t)sys * strideA;
    double* Xs = X + (size_t)sys * strideX;
    for (int e = tid; e < CB * CB; e += 256) {
        const int c = e / CB, r = e % CB;
        P[r][c] = As[(size_t)(k + c) * n + i0 + r];
        Bs[r][c] = d0 + c < D ? Xs[(size_t)(d0 + c) * n + k + r] : 0.0;
    }
    __syncthreads();
    const int tr = tid & 15, tc = tid >> 4;
    double acc[4][4] = {};
    for (int m = 0; m < CB; ++m) {
        double p[4], v[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) p[a] = P[tr + 16 * a][m];
#pragma unroll
        for (int b = 0; b < 4; ++b) v[b] = Bs[m][tc + 16 * b];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = fma(p[a], v[b], acc[a][b]);
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const int c = tc + 16 * b;
        if (d0 + c < D)
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                double* q = Xs + (size_t)(d0 + c) * n + i0 + tr + 16 * a;
                *q = *q - acc[a][b];
            }
    }
}

// Partial products Q[s][ch][d][i] = sum over j in column chunk ch of A[j][i] X[d0 + d][j]
// (A full symmetric, column-major).
constexpr int kResChunk = 256;
__global__ __launch_bounds__(256) void k_resid_part(int n, int D, int d0, const double* A, size_t strideA,
                                                    const double* X, size_t strideX, double* Q, int nchj) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int ch = blockIdx.y, sys = blockIdx.z;
    if (i >= n) return;
    const double* As = A + (size_t)sys * strideA;
    const double* Xs = X + (size_t)sys * strideX;
    const int dn = min(DP, D - d0);
    double acc[DP];
#pragma unroll
    for (int d = 0; d < DP; ++d) acc[d] = 0.0;
    const int j1 = min(n, (ch + 1) * kResChunk);
    for (int j = ch * kResChunk; j < j1; ++j) {
        const double a = As[(size_t)j * n + i];
#pragma unroll
        for (int d = 0; d < DP; ++d)
            if (d < dn) acc[d] = fma(a, Xs[(size_t)(d0 + d) * n + j], acc[d]);
    }
    double* Qs = Q + ((size_t)sys * nchj + ch) * DP * n;
#pragma unroll
    for (int d = 0; d < DP; ++d)
        if (d < dn) Qs[(size_t)d * n + i] = acc[d];
}

// R = B - sum over chunks (in order) of Q.
__global__ __launch_bounds__(256) void k_resid_sum(int n, int D, int d0, const double* Q, int nchj, const double* B,
                                                   double* R, size_t strideB) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int d = blockIdx.y, sys = blockIdx.z;
    if (i >= n || d0 + d >= D) return;
    double s = 0.0;
    for (int c = 0; c < nchj; ++c) s += Q[(((size_t)sys * nchj + c) * DP + d) * n + i];
    const size_t o = (size_t)sys * strideB + (size_t)(d0 + d) * n + i;
    R[o] = B[o] - s;
}

int cross_chunks(int T) { return std::max(1, std::min(32, (T + 127) / 128)); }
int resid_chunks(int n) { return (n + kResChunk - 1) / kResChunk; }

}  // namespace

bool chol_batch_ok(int n) { return n >= CB && n % CB == 0 && n <= kCholMaxN; }

// Workspace: Linv blocks [count][n / 64][64 x 64], residual partials
// [count][resid_chunks][DP][n], cross-product partials [count][cross_chunks][DP][n].
size_t chol_batch_ws_doubles(int count, int n, int T, int D) {
    (void)D;
    return (size_t)count * n * CB + (size_t)count * resid_chunks(n) * DP * n + (size_t)count * cross_chunks(T) * DP * n;
}

hipError_t gram_cross_batched(hipStream_t st, int count, int n, int T, int D, const double* const* Xs,
                              const double* Yd, const double* lam, double* A, size_t strideA, double* B,
                              size_t strideB, double* ws, bool with_cross) {
    const int nch = cross_chunks(T);
    const int chunk = (T + nch - 1) / nch;
    double* P = ws + (size_t)count * n * CB + (size_t)count * resid_chunks(n) * DP * n;
    for (int s0 = 0; s0 < count; s0 += kGramBatch) {
        const int cs = std::min(kGramBatch, count - s0);
        GramArgs ga{};
        for (int s = 0; s < cs; ++s) {
            ga.X[s] = Xs[s0 + s];
            ga.lam[s] = lam[s0 + s];
        }
        double* As = A + (size_t)s0 * strideA;
        if (n % 128 == 0) {
            const int nt = n / 128;
            hipLaunchKernelGGL(k_gram_lower<128>, dim3(nt * (nt + 1) / 2, cs), dim3(256), 0, st, ga, n, T, As, strideA);
        } else {
            const int nt = n / 64;
            hipLaunchKernelGGL(k_gram_lower<64>, dim3(nt * (nt + 1) / 2, cs), dim3(256), 0, st, ga, n, T, As, strideA);
        }
        const size_t strideY = (size_t)T * D;      // Y of system s at Yd + s * T * D, row-major [T x D]
        for (int d0 = 0; with_cross && d0 < D; d0 += DP) {
            hipLaunchKernelGGL(k_cross_part, dim3((n + 255) / 256, nch, cs), dim3(256), 0, st, ga, n, T, D, d0, chunk,
                               Yd + (size_t)s0 * strideY, strideY, P, nch);
            hipLaunchKernelGGL(k_cross_sum, dim3((n + 255) / 256, DP, cs), dim3(256), 0, st, n, D, d0, P, nch,
                               B + (size_t)s0 * strideB, strideB);
        }
    }
    return hipGetLastError();
}

hipError_t chol_batched(hipStream_t st, int count, int n, double* A, size_t strideA, double* ws, int* info) {
    const int nbk = n / CB;
    hipLaunchKernelGGL(k_zero_int, dim3((count + 255) / 256), dim3(256), 0, st, info, count);
    for (int kb = 0; kb < nbk; ++kb) {
        hipLaunchKernelGGL(k_chol_diag, dim3(count), dim3(CB), 0, st, n, kb, A, strideA, ws, info);
        const int rest = nbk - kb - 1;
        if (rest == 0) break;
        hipLaunchKernelGGL(k_chol_panel, dim3(rest, count), dim3(256), 0, st, n, kb, A, strideA, ws);
        hipLaunchKernelGGL(k_chol_update, dim3(rest * (rest + 1) / 2, count), dim3(256), 0, st, n, kb, A, strideA);
    }
    return hipGetLastError();
}

hipError_t resid_batched(hipStream_t st, int count, int n, int D, const double* A, size_t strideA, const double* X,
                         const double* B, double* R, size_t strideX, double* ws) {
    const int nchj = resid_chunks(n);
    double* Q = ws + (size_t)count * n * CB;
    for (int d0 = 0; d0 < D; d0 += DP) {
        hipLaunchKernelGGL(k_resid_part, dim3((n + 255) / 256, nchj, count), dim3(256), 0, st, n, D, d0, A, strideA, X,
                           strideX, Q, nchj);
        hipLaunchKernelGGL(k_resid_sum, dim3((n + 255) / 256, DP, count), dim3(256), 0, st, n, D, d0, Q, nchj, B, R,
                           strideX);
    }
    return hipGetLastError();
}

hipError_t chol_solve_many(hipStream_t st, int count, int n, int D, const double* A, size_t strideA,
                           const double* ws, double* X, size_t strideX) {
    const int nbk = n / CB, dc = (D + CB - 1) / CB;
    for (int kb = 0; kb < nbk; ++kb) {               // L Z = X
        hipLaunchKernelGGL(k_trsm_diag, dim3(dc, count), dim3(256), 0, st, n, D, kb, 0, ws, X, strideX);
        if (kb + 1 < nbk)
            hipLaunchKernelGGL(k_trsm_update, dim3(nbk - kb - 1, dc, count), dim3(256), 0, st, n, D, kb, kb + 1, A,
                               strideA, X, strideX);
    }
    for (int kb = nbk - 1; kb >= 0; --kb) {          // L' X = Z
        hipLaunchKernelGGL(k_trsm_diag, dim3(dc, count), dim3(256), 0, st, n, D, kb, 1, ws, X, strideX);
        if (kb > 0)
            hipLaunchKernelGGL(k_trsm_update, dim3(kb, dc, count), dim3(256), 0, st, n, D, kb, 0, A, strideA, X,
                               strideX);
    }
    return hipGetLastError();
}

hipError_t chol_solve_batched(hipStream_t st, int count, int n, int D, const double* A, size_t strideA,
                              const double* ws, double* X, size_t strideX) {
    const size_t lds = (size_t)n * sizeof(double);
    hipLaunchKernelGGL(k_chol_fwd, dim3(count, D), dim3(kSolveThreads), lds, st, n, A, strideA, ws, X, strideX);
    hipLaunchKernelGGL(k_chol_bwd, dim3(count, D), dim3(kSolveThreads), lds, st, n, A, strideA, ws, X, strideX);
    return hipGetLastError();
}

}  // namespace sdev
