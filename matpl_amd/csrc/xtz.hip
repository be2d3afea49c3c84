// Y_t = X_t^T Z_t for tall, thin operands: the weight-gradient sums over atoms of the DP training
// step's fitting nets (per centre type t, X_t the rows of the layer inputs and their tangents, Z_t
// the pre-activation adjoints; K = twice the type's atoms, tens of thousands of rows, M <= 512,
// N <= 64). Library GEMMs pick small tiles for these shapes and run far below the HBM rate; here a
// split over K: each block reduces a chunk of rows for a 64 x 64 output tile (thread: 4 x 4
// outputs, operand slabs staged in LDS), writes its partial tile, and a second kernel adds the
// chunks of each type in chunk order (deterministic, no atomics).
//
// Rows come in pairs (an atom's value and its tangent): row r of a type lives at
// base + (r / 2) * ld2 + (r % 2) * ld1, so interleaved per-atom layouts need no copy.
#include "common.h"

namespace matpl {
namespace {

constexpr int XTZ_MAXT = 16;
constexpr int XT_TILE = 64;      // output tile (M and N)
constexpr int XT_SLAB = 16;      // rows staged per step
constexpr int XT_CHUNK = 1024;   // rows per block

struct XtzArgs {
    const float *X, *Z;
    int64_t ldx2, ldx1, ldz2, ldz1;
    int64_t start[XTZ_MAXT + 1];   // type t's atoms [start[t], start[t+1]) (rows 2 start .. 2 end)
    int M, N, nchunk;
    float *part;                   // [T][nchunk][M][N]
};

__device__ __forceinline__ int64_t row_off(int64_t r, int64_t ld2, int64_t ld1) { return (r >> 1) * ld2 + (r & 1) * ld1; }

__global__ void __launch_bounds__(256) xtz_partial_kernel(XtzArgs a) {
    __shared__ float xs[XT_SLAB][XT_TILE + 1];
    __shared__ float zs[XT_SLAB][XT_TILE + 1];
    const int t = blockIdx.z, c = blockIdx.y, m0 = blockIdx.x * XT_TILE;
    const int tid = threadIdx.x, ty = tid / 16, tx = tid % 16;
    const int64_t r0 = 2 * a.start[t] + (int64_t)c * XT_CHUNK;
    const int64_t r1e = 2 * a.start[t + 1];
    const int64_t r1 = r0 + XT_CHUNK < r1e ? r0 + XT_CHUNK : r1e;
    float acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = 0.0f;
    for (int64_t rs = r0; rs < r1; rs += XT_SLAB) {
        // stage X[rs .. rs+16][m0 .. m0+64] and Z[rs .. rs+16][0 .. 64]: 1024 values each, 4 per thread
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int e = tid + u * 256, rr = e / XT_TILE, cc = e % XT_TILE;
            const int64_t r = rs + rr;
            const bool ok = r < r1;
            xs[rr][cc] = (ok && m0 + cc < a.M) ? a.X[row_off(r, a.ldx2, a.ldx1) + m0 + cc] : 0.0f;
            zs[rr][cc] = (ok && cc < a.N) ? a.Z[row_off(r, a.ldz2, a.ldz1) + cc] : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < XT_SLAB; k++) {
            float xv[4], zv[4];
#pragma unroll
            for (int i = 0; i < 4; i++) { xv[i] = xs[k][ty * 4 + i]; zv[i] = zs[k][tx * 4 + i]; }
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] += xv[i] * zv[j];
        }
        __syncthreads();
    }
    float *p = a.part + (((int64_t)t * a.nchunk + c) * a.M) * a.N;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int m = m0 + ty * 4 + i;
        if (m >= a.M) continue;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int n = tx * 4 + j;
            if (n < a.N) p[(int64_t)m * a.N + n] = acc[i][j];
        }
    }
}

__global__ void xtz_reduce_kernel(const float *part, int nchunk, int64_t MN, int T, float *Y) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int t = blockIdx.y;
    if (e >= MN || t >= T) return;
    float s = 0.0f;
    for (int c = 0; c < nchunk; c++) s += part[((int64_t)t * nchunk + c) * MN + e];
    Y[(int64_t)t * MN + e] = s;
}

int nchunk_of(int64_t max_atoms) {
    const int64_t rows = 2 * max_atoms;
    int64_t n = (rows + XT_CHUNK - 1) / XT_CHUNK;
    return (int)(n < 1 ? 1 : n);
}

}  // namespace
}  // namespace matpl

using namespace matpl;

// Workspace elements of matpl_tall_xtz for types of at most max_atoms atoms.
MATPL_EXPORT int64_t matpl_tall_xtz_work(int64_t max_atoms, int ntypes, int M, int N) {
    if (max_atoms < 0 || ntypes < 1 || M < 1 || N < 1) return -1;
    return (int64_t)ntypes * nchunk_of(max_atoms) * M * N;
}

MATPL_EXPORT int matpl_tall_xtz(const float *X, int64_t ldx2, int64_t ldx1, const float *Z, int64_t ldz2, int64_t ldz1,
                                const int64_t *start, int ntypes, int M, int N, float *Y, float *work,
                                int64_t work_elems, void *stream) {
    MATPL_CHECK_ARG(ntypes >= 1 && ntypes <= XTZ_MAXT, "tall_xtz: ntypes in [1, %d]", XTZ_MAXT);
    MATPL_CHECK_ARG(M >= 1 && N >= 1 && N <= XT_TILE, "tall_xtz: M >= 1, 1 <= N <= %d", XT_TILE);
    MATPL_CHECK_ARG(start && X && Z && Y && work, "tall_xtz: null pointer");
    XtzArgs a{};
    int64_t maxa = 0;
    for (int t = 0; t <= ntypes; t++) a.start[t] = start[t];
    for (int t = 0; t < ntypes; t++) {
        MATPL_CHECK_ARG(start[t + 1] >= start[t], "tall_xtz: decreasing type ranges");
        if (start[t + 1] - start[t] > maxa) maxa = start[t + 1] - start[t];
    }
    a.X = X; a.Z = Z; a.ldx2 = ldx2; a.ldx1 = ldx1; a.ldz2 = ldz2; a.ldz1 = ldz1;
    a.M = M; a.N = N; a.nchunk = nchunk_of(maxa); a.part = work;
    MATPL_CHECK_ARG(work_elems >= (int64_t)ntypes * a.nchunk * M * N, "tall_xtz: workspace too small (%lld < %lld)",
                    (long long)work_elems, (long long)((int64_t)ntypes * a.nchunk * M * N));
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(xtz_partial_kernel, dim3((M + XT_TILE - 1) / XT_TILE, a.nchunk, ntypes), dim3(256), 0, st, a);
    int rc = check_launch("xtz_partial_kernel");
    if (rc) return rc;
    const int64_t MN = (int64_t)M * N;
    hipLaunchKernelGGL(xtz_reduce_kernel, dim3((unsigned)((MN + 255) / 256), ntypes), dim3(256), 0, st, work, a.nchunk,
                       MN, ntypes, Y);
    return check_launch("xtz_reduce_kernel");
}
