// Split-BF16 Gram of the state record on the bf16 matrix cores (fast-mode
// readout; the reference forms X'X in FP64 with one dgemm, math_utils.c:154-162,
// reservoir.c:600-620).
//
// Each state value x (FP64) is split once into hi = bf16(x) and
// lo = bf16(fl32(x) - hi), so x ~ hi + lo to about 2^-16 relative, and
// X'X ~ hi'hi + hi'lo + lo'hi (the lo'lo term, ~2^-16, is dropped).  The split
// pass writes hi and lo neuron-major ([n][Tp] bf16, Tp = T rounded up to 32,
// zero padded), so that one lane's operand fragment -- 8 consecutive rows of
// one neuron -- is a single 16-byte load.
//
// The Gram kernel computes one 128 x 128 lower-triangle tile per workgroup
// with v_mfma_f32_16x16x32_bf16: four waves of 64 x 64 (4 x 4 blocks of
// 16 x 16), three products per block per 32-row chunk, accumulated in FP32
// over segments of kSegChunks chunks and then added into the FP64 tile in
// HBM (the FP32 accumulation length stays kSegChunks x 32 rows).  Operand
// layout (checked with an in-bounds probe on gfx950): A[i][k] is element k%8
// of lane i + 16 (k/8), B[k][j] element k%8 of lane j + 16 (k/8); register r
// of lane l holds D[4 (l/16) + r][l % 16].
//
// Tiles are ordered so that the workgroups resident on one XCD at a time
// (workgroup ids alike mod 8 are dispatched to the same XCD) work on a
// compact block of tiles: their row and column panels are shared in that
// XCD's L2 instead of being fetched from HBM by every tile.
#include "dev_internal.hpp"

#include <cstdint>

namespace sdev {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kGramTile = 128;      // tile edge (neurons)
constexpr int kSegChunks = 64;      // 32-row chunks per FP32 segment (2048 rows)

__device__ __forceinline__ unsigned short bf16_rne(float f) {
    const unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);   // inf / nan
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// hi / lo [n][Tp] from X [T x n] (row t contiguous over neurons): 64 x 64
// blocks through LDS so that both the reads (along neurons) and the writes
// (along rows) are contiguous.  Rows T..Tp-1 are written as zeros.
__global__ void __launch_bounds__(256) k_split_transpose(const double* __restrict__ X, int T, int n, int Tp,
                                                        unsigned short* __restrict__ hi,
                                                        unsigned short* __restrict__ lo) {
    __shared__ unsigned short sh[64][65];
    __shared__ unsigned short sl[64][65];
    const int t0 = blockIdx.x * 64, i0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;   // 64 x 4
    for (int r = ty; r < 64; r += 4) {
        const int t = t0 + r, i = i0 + tx;
        float f = 0.0f;
        if (t < T && i < n) f = (float)X[(size_t)t * n + i];
        const unsigned short h = bf16_rne(f);
        const float hf = __uint_as_float((unsigned)h << 16);
        sh[r][tx] = h;
        sl[r][tx] = bf16_rne(f - hf);
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 4) {
        const int i = i0 + r, t = t0 + tx;
        if (i < n && t < Tp) {
            hi[(size_t)i * Tp + t] = sh[tx][r];
            lo[(size_t)i * Tp + t] = sl[tx][r];
        }
    }
}

// Lower-triangle tile (I, J), I >= J, of workgroup b: consecutive groups of 8
// workgroup ids (one per XCD) walk a block of 8 x 8 tiles column by column, so
// the ~32 workgroups an XCD holds at once share about four row and four
// column panels.
__device__ __forceinline__ void tile_of(int b, int nt, int* I, int* J) {
    const int xcd = b & 7, k = b >> 3;            // k-th tile handed to this XCD
    const int per = (nt * (nt + 1) / 2 + 7) / 8;  // tiles per XCD, rounded up
    int q = xcd * per + k;                        // this XCD's contiguous share of the tile order
    // tile order: 8 x 8 super-blocks of the lower triangle, row-major over super-blocks,
    // column-major inside one
    const int ns = (nt + 7) / 8;
    for (int sI = 0; sI < ns; ++sI) {
        for (int sJ = 0; sJ <= sI; ++sJ) {
            const int r0 = sI * 8, c0 = sJ * 8;
            const int rh = min(8, nt - r0), cw = min(8, nt - c0);
            int cnt = 0;
            if (sI != sJ) cnt = rh * cw;
            else cnt = rh * (rh + 1) / 2;
            if (q < cnt) {
                if (sI != sJ) { *J = c0 + q / rh; *I = r0 + q % rh; }
                else {
                    int c = 0, left = q;
                    while (left >= rh - c) { left -= rh - c; ++c; }
                    *J = c0 + c; *I = r0 + c + left;
                }
                return;
            }
            q -= cnt;
        }
    }
    *I = -1; *J = -1;
}

__global__ void __launch_bounds__(256) k_gram_split(const unsigned short* __restrict__ hi,
                                                   const unsigned short* __restrict__ lo, int n, int Tp,
                                                   double* __restrict__ G, double lambda) {
    const int nt = n / kGramTile;
    int I, J;
    tile_of(blockIdx.x, nt, &I, &J);
    if (I < 0) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;   // this wave's 64 x 64 part of the tile
    const int row0 = I * kGramTile + wr, col0 = J * kGramTile + wc;
    const int li = lane & 15, lk = (lane >> 4) * 8;
    // operand rows of this lane: neuron row0 + 16 a + li (A) and col0 + 16 b + li (B)
    const unsigned short* ha = hi + (size_t)(row0 + li) * Tp + lk;
    const unsigned short* la = lo + (size_t)(row0 + li) * Tp + lk;
    const unsigned short* hb = hi + (size_t)(col0 + li) * Tp + lk;
    const unsigned short* lb = lo + (size_t)(col0 + li) * Tp + lk;
    const size_t blk = (size_t)16 * Tp;                       // next 16-neuron block
    const int chunks = Tp / 32;
    bool first = true;
    for (int c0 = 0; c0 < chunks; c0 += kSegChunks) {
        f32x4 acc[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        const int c1 = min(chunks, c0 + kSegChunks);
        for (int c = c0; c < c1; ++c) {
            const size_t off = (size_t)c * 32;
            bf16x8 Ah[4], Al[4], Bh[4], Bl[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                Ah[a] = *reinterpret_cast<const bf16x8*>(ha + a * blk + off);
                Al[a] = *reinterpret_cast<const bf16x8*>(la + a * blk + off);
                Bh[a] = *reinterpret_cast<const bf16x8*>(hb + a * blk + off);
                Bl[a] = *reinterpret_cast<const bf16x8*>(lb + a * blk + off);
            }
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah[a], Bh[b], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah[a], Bl[b], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Al[a], Bh[b], acc[a][b], 0, 0, 0);
                }
        }
        // segment -> FP64 tile (column-major n x n: element (row, col) at row + col n);
        // each element belongs to one lane, which also adds lambda on the diagonal
        // after the last segment
        const bool last = c1 == chunks;
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int col = col0 + 16 * b + li;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = row0 + 16 * a + (lane >> 4) * 4 + r;
                    double* p = G + (size_t)col * n + row;
                    double v = (double)acc[a][b][r];
                    if (!first) v = *p + v;
                    if (last && row == col) v = v + lambda;
                    *p = v;
                }
            }
        first = false;
    }
}

// Lower triangle (column-major, ld n) of hi'hi + hi'lo + lo'hi + lambda I for the
// state record X [T x n]; the upper triangle is not written.  ws holds 2 n Tp
// bf16 (Tp = T rounded up to 32).  n must be a multiple of 128.
int gram_split_bf16(hipStream_t st, const double* X, int T, int n, double lambda, double* G, unsigned short* ws) {
    if (n % kGramTile != 0 || T < 1) { set_error("split-BF16 Gram: n must be a multiple of 128"); return SPIRES_DEV_EINVAL; }
    const int Tp = (T + 31) / 32 * 32;
    unsigned short* hi = ws;
    unsigned short* lo = ws + (size_t)n * Tp;
    hipLaunchKernelGGL(k_split_transpose, dim3((Tp + 63) / 64, (n + 63) / 64), dim3(256), 0, st, X, T, n, Tp, hi, lo);
    SDEV_CK(hipGetLastError());
    const int nt = n / kGramTile;
    const int tiles = nt * (nt + 1) / 2;
    const int grid = (tiles + 7) / 8 * 8;         // every XCD's share, rounded up
    hipLaunchKernelGGL(k_gram_split, dim3(grid), dim3(256), 0, st, hi, lo, n, Tp, G, lambda);
    SDEV_CK(hipGetLastError());
    return 0;
}

size_t gram_split_ws_elems(int T, int n) { return 2 * (size_t)n * (size_t)((T + 31) / 32 * 32); }

}  // namespace sdev
