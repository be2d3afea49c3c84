// Tall-by-small products y = x W for training the DP nets through autograd:
// x [K][a] FP32 (row stride ldx, K = neighbor slots or atoms), W [a][b] FP32
// (element (j, c) at W[j * w_row + c * w_col], so a transposed view needs no
// copy), y [K][b] FP32 contiguous, 1 <= a, b <= 64. With lowp the operands
// are rounded to bfloat16 and so is the result (products and sums in FP32):
// the arithmetic of a BF16 product with FP32 accumulation whose BF16 output
// is read back as FP32, without the separate conversion passes.
//
// Reference: the layer products torch.matmul(x, weight) of EmbeddingNet and
// FittingNet (src/model/dp_embedding.py:73-94, 140-152), which the reference
// leaves to the BLAS library.
//
// These products are bound by memory and latency, not by arithmetic (a and b
// are 25 or 50 for the DP nets). A block of 256 threads takes 64 rows. It
// stages W and its 64 x a rows of x in LDS (sized to a and b, so that many
// blocks share a CU) with coalesced loads, at odd row strides, so that the
// 64 lanes reading one column of the tile hit distinct banks. Each thread forms the outputs g, g + 4, ... (g = thread / 64, the
// same for the whole wavefront) of one row, and the 64 x b results leave
// through LDS as one contiguous store.
#include "common.h"

namespace matpl {
namespace {

constexpr int TM_ROWS = 64, TM_THREADS = 256, TM_MAX = 64;

__device__ __forceinline__ float bf16_round(float v) { return (float)(__bf16)v; }

// Row and column of element e of a row-major tile with n columns, from a float reciprocal: for
// e < 2^13 and n <= 64, (e + 1/2) / n lies at least 1/(2n) from an integer and the float product's
// error is below 1e-3, so the truncation is exact (no integer division).
__device__ __forceinline__ int tile_row(int e, float inv_n) { return (int)(((float)e + 0.5f) * inv_n); }

template <int NO, bool LOWP>
__global__ void __launch_bounds__(TM_THREADS)
tallmm_kernel(const float *__restrict__ x, int64_t ldx, const float *__restrict__ W, int64_t w0, int64_t w1,
              float *__restrict__ y, int64_t K, int a, int b, float inv_a, float inv_b) {
    // LDS: W [a][b], the x tile [64][a | 1], the y tile [64][b | 1] (odd strides)
    extern __shared__ float smem[];
    const int ap = a | 1, bp = b | 1;
    float *ws = smem, *xs = smem + a * b, *ys = xs + TM_ROWS * ap;
    const int t = threadIdx.x;
    for (int e = t; e < a * b; e += TM_THREADS) {
        const int j = tile_row(e, inv_b), c = e - j * b;
        const float v = W[j * w0 + c * w1];
        ws[e] = LOWP ? bf16_round(v) : v;
    }
    const int64_t row0 = (int64_t)blockIdx.x * TM_ROWS;
    const int nr = K - row0 < TM_ROWS ? (int)(K - row0) : TM_ROWS;
    const float *xg = x + row0 * ldx;
    for (int e = t; e < nr * a; e += TM_THREADS) {
        const int r = tile_row(e, inv_a), j = e - r * a;
        const float v = xg[r * ldx + j];
        xs[r * ap + j] = LOWP ? bf16_round(v) : v;
    }
    __syncthreads();
    const int r = t & (TM_ROWS - 1), g = t / TM_ROWS;
    float acc[NO];
#pragma unroll
    for (int i = 0; i < NO; i++) acc[i] = 0.0f;
    if (r < nr) {
        for (int j = 0; j < a; j++) {
            const float xv = xs[r * ap + j];
            const float *wr = ws + j * b;
#pragma unroll
            for (int i = 0; i < NO; i++) {
                const int c = g + 4 * i;
                if (c < b) acc[i] += xv * wr[c];
            }
        }
#pragma unroll
        for (int i = 0; i < NO; i++) {
            const int c = g + 4 * i;
            if (c < b) ys[r * bp + c] = LOWP ? bf16_round(acc[i]) : acc[i];
        }
    }
    __syncthreads();
    float *yg = y + row0 * b;
    for (int e = t; e < nr * b; e += TM_THREADS) {
        const int rr = tile_row(e, inv_b), c = e - rr * b;
        yg[e] = ys[rr * bp + c];
    }
}

template <int NO>
int launch_tallmm(const float *x, int64_t ldx, const float *W, int64_t w0, int64_t w1, float *y, int64_t K, int a,
                  int b, bool lowp, hipStream_t st) {
    const dim3 grid((unsigned)((K + TM_ROWS - 1) / TM_ROWS));
    const size_t lds = sizeof(float) * ((size_t)a * b + (size_t)TM_ROWS * ((a | 1) + (b | 1)));
    const float ia = 1.0f / (float)a, ib = 1.0f / (float)b;
    if (lowp)
        hipLaunchKernelGGL((tallmm_kernel<NO, true>), grid, dim3(TM_THREADS), lds, st, x, ldx, W, w0, w1, y, K, a, b,
                           ia, ib);
    else
        hipLaunchKernelGGL((tallmm_kernel<NO, false>), grid, dim3(TM_THREADS), lds, st, x, ldx, W, w0, w1, y, K, a, b,
                           ia, ib);
    return check_launch("tallmm_kernel");
}

}  // namespace
}  // namespace matpl

using namespace matpl;

MATPL_EXPORT int matpl_tall_mm(const float *x, int64_t ldx, const float *W, int64_t w_row, int64_t w_col, float *y,
                               int64_t K, int a, int b, int lowp, void *stream) {
    MATPL_CHECK_ARG(a >= 1 && a <= TM_MAX && b >= 1 && b <= TM_MAX, "tall_mm: a = %d, b = %d outside [1, %d]", a, b,
                    TM_MAX);
    MATPL_CHECK_ARG(K >= 0 && K / TM_ROWS < (int64_t)0x7fffffff, "tall_mm: K = %lld out of range", (long long)K);
    if (K == 0) return OK;
    MATPL_CHECK_ARG(x && W && y, "tall_mm: null pointer");
    MATPL_CHECK_ARG(ldx >= a, "tall_mm: row stride %lld of x below a = %d", (long long)ldx, a);
    hipStream_t st = (hipStream_t)stream;
    const int no = (b + 3) / 4;
    if (no <= 1) return launch_tallmm<1>(x, ldx, W, w_row, w_col, y, K, a, b, lowp != 0, st);
    if (no <= 2) return launch_tallmm<2>(x, ldx, W, w_row, w_col, y, K, a, b, lowp != 0, st);
    if (no <= 4) return launch_tallmm<4>(x, ldx, W, w_row, w_col, y, K, a, b, lowp != 0, st);
    if (no <= 8) return launch_tallmm<8>(x, ldx, W, w_row, w_col, y, K, a, b, lowp != 0, st);
    return launch_tallmm<16>(x, ldx, W, w_row, w_col, y, K, a, b, lowp != 0, st);
}
