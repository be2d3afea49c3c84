// Y_t = X_t^T Z_t for tall, thin operands: the weight-gradient sums over atoms of the DP training
// step's fitting nets (per centre type t, X_t the rows of the layer inputs and their tangents, Z_t
// the pre-activation adjoints; K = twice the type's atoms, tens of thousands of rows, M <= 512,
// N <= 64). Library GEMMs pick small tiles for these shapes and run far below the HBM rate; here a
// split over K on the FP32 matrix cores: a 4-wavefront block owns a 64 x 64 output tile and a
// chunk of 4 x 128 rows, each wavefront accumulates its 128 rows with v_mfma_f32_32x32x2_f32 (two
// rows, an atom's value and its tangent, per instruction, operands loaded straight from global
// memory a batch ahead), the four wavefronts' tiles are added in LDS in a fixed order, and a
// second kernel adds the chunks of each type in chunk order (deterministic, no atomics).
//
// Rows come in pairs (an atom's value and its tangent): row r of a type lives at
// base + (r / 2) * ld2 + (r % 2) * ld1, so interleaved per-atom layouts need no copy.
#include "common.h"

namespace matpl {
namespace {

constexpr int XTZ_MAXT = 16;
constexpr int XT_TILE = 64;          // output tile (M and N)
constexpr int XT_WAVES = 4;          // wavefronts per block, each on its own rows
constexpr int XT_WROWS = 128;        // rows per wavefront
constexpr int XT_CHUNK = XT_WAVES * XT_WROWS;
constexpr int XT_BATCH = 8;          // atoms (row pairs) loaded per batch

typedef __attribute__((__vector_size__(16 * sizeof(float)))) float floatx16;

struct XtzArgs {
    const float *X, *Z;
    int64_t ldx2, ldx1, ldz2, ldz1;
    int64_t start[XTZ_MAXT + 1];     // type t's atoms [start[t], start[t+1]) (rows 2 start .. 2 end)
    int M, N, nchunk, mtiles, ntypes;
    int64_t nblocks;                 // T * nchunk * mtiles (the grid may carry up to 7 more)
    float *part;                     // [T][nchunk][M][N]
};

__device__ __forceinline__ int64_t row_off(int64_t r, int64_t ld2, int64_t ld1) { return (r >> 1) * ld2 + (r & 1) * ld1; }

struct Batch {
    float x[XT_BATCH][2], z[XT_BATCH][2];
};

// Operands of XT_BATCH atoms from row r (even, r < r1): lane (kk = lane / 32, c = lane % 32)
// takes row r + 2u + kk, columns m0 + c, m0 + 32 + c of X and c, 32 + c of Z. Columns past M or N
// read column m0 or 0 instead (inside the operands): they only reach output rows and columns that
// are never written. Rows past r1 (FULL = false, the last batch of a range) read row r + kk and
// are zeroed; full batches load without any select, so that the compiler can leave the loads in
// flight while the matrix cores work on the previous batch.
template <bool FULL>
__device__ __forceinline__ void load_batch(const XtzArgs &a, int64_t r, int64_t r1, int m0, int kk, int c, Batch &b) {
    const int cA = m0 + c < a.M ? c : 0, cB = m0 + 32 + c < a.M ? 32 + c : 0;
    const int zA = c < a.N ? c : 0, zB = 32 + c < a.N ? 32 + c : 0;
#pragma unroll
    for (int u = 0; u < XT_BATCH; u++) {
        const bool ok = FULL || r + 2 * u < r1;
        const int64_t row = (ok ? r + 2 * u : r) + kk;
        const float *xr = a.X + row_off(row, a.ldx2, a.ldx1) + m0;
        const float *zr = a.Z + row_off(row, a.ldz2, a.ldz1);
        b.x[u][0] = xr[cA];
        b.x[u][1] = xr[cB];
        b.z[u][0] = ok ? zr[zA] : 0.0f;
        b.z[u][1] = ok ? zr[zB] : 0.0f;
    }
}

__device__ __forceinline__ void mma_batch(const Batch &b, floatx16 acc[2][2]) {
#pragma unroll
    for (int u = 0; u < XT_BATCH; u++)
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(b.x[u][i], b.z[u][j], acc[i][j], 0, 0, 0);
}

__global__ void __launch_bounds__(XT_WAVES * 64) xtz_partial_kernel(XtzArgs a) {
    __shared__ float red[(XT_WAVES - 1) * 64][64];
    // XCD-aware order: hardware block b runs on XCD b % 8; logical index L keeps consecutive
    // logical blocks (the M tiles of one chunk, which read the same Z rows) on one XCD's L2
    const int64_t G = gridDim.x, b = blockIdx.x;
    const int64_t L = (G % 8 == 0) ? (b % 8) * (G / 8) + b / 8 : b;
    if (L >= a.nblocks) return;
    const int mt = (int)(L % a.mtiles);
    const int64_t tc = L / a.mtiles;
    const int c = (int)(tc % a.nchunk), t = (int)(tc / a.nchunk);
    const int m0 = mt * XT_TILE;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);   // wave-uniform: scalar loop bounds
    const int lane = threadIdx.x % 64, kk = lane / 32, col = lane % 32;
    const int64_t r1e = 2 * a.start[t + 1];
    const int64_t rw0 = 2 * a.start[t] + (int64_t)c * XT_CHUNK + (int64_t)w * XT_WROWS;
    const int64_t rw1 = rw0 + XT_WROWS < r1e ? rw0 + XT_WROWS : r1e;
    floatx16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int v = 0; v < 16; v++) acc[i][j][v] = 0.0f;
    // full batches, two in flight: the next one loads while the matrix cores work on the current
    // one. The loads are unconditional (a branch around them would make the compiler wait for
    // them before the matrix products): the last pair reloads its second batch, unused.
    const int nb = rw1 > rw0 ? (int)((rw1 - rw0) / (2 * XT_BATCH)) : 0;
    const int64_t rfull = rw0 + (int64_t)nb * 2 * XT_BATCH;
    if (nb > 0) {
        Batch b0, b1;
        load_batch<true>(a, rw0, rfull, m0, kk, col, b0);
        int k = 0;
        for (; k + 1 < nb; k += 2) {
            load_batch<true>(a, rw0 + (int64_t)(k + 1) * 2 * XT_BATCH, rfull, m0, kk, col, b1);
            mma_batch(b0, acc);
            const int k2 = k + 2 < nb ? k + 2 : nb - 1;
            load_batch<true>(a, rw0 + (int64_t)k2 * 2 * XT_BATCH, rfull, m0, kk, col, b0);
            mma_batch(b1, acc);
        }
        if (k < nb) mma_batch(b0, acc);     // odd count: b0 holds batch nb - 1
    }
    if (rfull < rw1) {
        Batch bt;
        load_batch<false>(a, rfull, rw1, m0, kk, col, bt);
        mma_batch(bt, acc);
    }
    // the block's four partial tiles, added in wavefront order
    if (w > 0) {
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
#pragma unroll
                for (int v = 0; v < 16; v++) red[(w - 1) * 64 + (2 * i + j) * 16 + v][lane] = acc[i][j][v];
    }
    __syncthreads();
    if (w != 0) return;
    float *p = a.part + (((int64_t)t * a.nchunk + c) * a.M) * a.N;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int v = 0; v < 16; v++) {
                float s = acc[i][j][v];
#pragma unroll
                for (int q = 1; q < XT_WAVES; q++) s += red[(q - 1) * 64 + (2 * i + j) * 16 + v][lane];
                // accumulator layout of the 32 x 32 tile: row 8 (v / 4) + 4 (lane / 32) + v % 4, column lane % 32
                const int m = m0 + 32 * i + 8 * (v / 4) + 4 * kk + (v % 4);
                const int n = 32 * j + col;
                if (m < a.M && n < a.N) p[(int64_t)m * a.N + n] = s;
            }
}

// Y[t][e] = sum over chunks c of part[t][c][e]: a block takes 64 entries, its four wavefronts
// the chunks c = g (mod 4) in order, then the four sums are added in g order.
__global__ void __launch_bounds__(256) xtz_reduce_kernel(const float *part, int nchunk, int64_t MN, int T, float *Y) {
    __shared__ float red[4][64];
    const int el = threadIdx.x % 64, g = threadIdx.x / 64;
    const int64_t e = (int64_t)blockIdx.x * 64 + el;
    const int t = blockIdx.y;
    float s = 0.0f;
    if (e < MN) {
        const float *p = part + (int64_t)t * nchunk * MN + e;
#pragma unroll 4
        for (int c = g; c < nchunk; c += 4) s += p[(int64_t)c * MN];
    }
    red[g][el] = s;
    __syncthreads();
    if (g == 0 && e < MN) Y[(int64_t)t * MN + e] = ((s + red[1][el]) + red[2][el]) + red[3][el];
}

// Column sums of row segments: for each buffer s and type t, out[t][coff_s + j] = sum over the
// type's positions p in [start[t], start[t+1]) of ptr_s[p * ld_s + j], j < cols_s (the bias,
// resnet_dt and output-layer gradients of the training step's fitting nets). A block takes 64
// columns of one buffer and a chunk of SR_CHUNK positions, its four wavefronts the positions
// p = g (mod 4) in order, their sums added in g order; xtz_reduce_kernel then adds the chunks.
constexpr int SR_MAXSEG = 8;
constexpr int SR_CHUNK = 1024;

struct SegArgs {
    const float *ptr[SR_MAXSEG];
    int64_t ld[SR_MAXSEG];
    int cols[SR_MAXSEG], coff[SR_MAXSEG], cb0[SR_MAXSEG + 1];   // cb0: first 64-column block of each buffer
    int nseg, totc, nchunk;
    int64_t start[XTZ_MAXT + 1];
    float *part;                                                  // [T][nchunk][totc]
};

__global__ void __launch_bounds__(256) seg_rowsum_partial_kernel(SegArgs a) {
    __shared__ float red[4][64];
    const int cb = blockIdx.x, c = blockIdx.y, t = blockIdx.z;
    int s = 0;
    while (s + 1 < a.nseg && cb >= a.cb0[s + 1]) s++;
    const int el = threadIdx.x % 64, g = threadIdx.x / 64;
    const int col = (cb - a.cb0[s]) * 64 + el;
    const int64_t p0 = a.start[t] + (int64_t)c * SR_CHUNK;
    const int64_t p1 = p0 + SR_CHUNK < a.start[t + 1] ? p0 + SR_CHUNK : a.start[t + 1];
    float sum = 0.0f;
    if (col < a.cols[s]) {
        const float *q = a.ptr[s] + col;
        const int64_t ld = a.ld[s];
#pragma unroll 8
        for (int64_t p = p0 + g; p < p1; p += 4) sum += q[p * ld];
    }
    red[g][el] = sum;
    __syncthreads();
    if (g == 0 && col < a.cols[s])
        a.part[((int64_t)t * a.nchunk + c) * a.totc + a.coff[s] + col] = ((sum + red[1][el]) + red[2][el]) + red[3][el];
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
    a.mtiles = (M + XT_TILE - 1) / XT_TILE;
    a.ntypes = ntypes;
    a.nblocks = (int64_t)ntypes * a.nchunk * a.mtiles;
    MATPL_CHECK_ARG(work_elems >= (int64_t)ntypes * a.nchunk * M * N, "tall_xtz: workspace too small (%lld < %lld)",
                    (long long)work_elems, (long long)((int64_t)ntypes * a.nchunk * M * N));
    MATPL_CHECK_ARG(a.nblocks + 7 < (int64_t)0x7fffffff, "tall_xtz: too many blocks");
    hipStream_t st = (hipStream_t)stream;
    // grid rounded up to a multiple of 8 so that the XCD-aware order is a permutation of it
    const int64_t grid = a.nblocks >= 8 ? (a.nblocks + 7) / 8 * 8 : a.nblocks;
    hipLaunchKernelGGL(xtz_partial_kernel, dim3((unsigned)grid), dim3(XT_WAVES * 64), 0, st, a);
    int rc = check_launch("xtz_partial_kernel");
    if (rc) return rc;
    const int64_t MN = (int64_t)M * N;
    hipLaunchKernelGGL(xtz_reduce_kernel, dim3((unsigned)((MN + 63) / 64), ntypes), dim3(256), 0, st, work, a.nchunk,
                       MN, ntypes, Y);
    return check_launch("xtz_reduce_kernel");
}

// Workspace elements of matpl_seg_rowsum for types of at most max_atoms positions and total_cols
// output columns.
MATPL_EXPORT int64_t matpl_seg_rowsum_work(int64_t max_atoms, int ntypes, int total_cols) {
    if (max_atoms < 0 || ntypes < 1 || total_cols < 1) return -1;
    const int64_t nchunk = max_atoms > 0 ? (max_atoms + SR_CHUNK - 1) / SR_CHUNK : 1;
    return (int64_t)ntypes * nchunk * total_cols;
}

MATPL_EXPORT int matpl_seg_rowsum(int nseg, const uint64_t *ptrs, const int64_t *lds, const int32_t *cols,
                                  const int64_t *start, int ntypes, float *out, float *work, int64_t work_elems,
                                  void *stream) {
    MATPL_CHECK_ARG(nseg >= 1 && nseg <= SR_MAXSEG, "seg_rowsum: 1 <= nseg <= %d", SR_MAXSEG);
    MATPL_CHECK_ARG(ntypes >= 1 && ntypes <= XTZ_MAXT, "seg_rowsum: ntypes in [1, %d]", XTZ_MAXT);
    MATPL_CHECK_ARG(ptrs && lds && cols && start && out && work, "seg_rowsum: null pointer");
    SegArgs a{};
    int64_t maxa = 0;
    for (int t = 0; t <= ntypes; t++) a.start[t] = start[t];
    for (int t = 0; t < ntypes; t++) {
        MATPL_CHECK_ARG(start[t + 1] >= start[t], "seg_rowsum: decreasing type ranges");
        if (start[t + 1] - start[t] > maxa) maxa = start[t + 1] - start[t];
    }
    a.nseg = nseg;
    int totc = 0, cb = 0;
    for (int s = 0; s < nseg; s++) {
        MATPL_CHECK_ARG(cols[s] >= 1 && lds[s] >= 1 && ptrs[s], "seg_rowsum: buffer %d: cols >= 1, ld >= 1, non-null",
                        s);
        a.ptr[s] = reinterpret_cast<const float *>(ptrs[s]);
        a.ld[s] = lds[s];
        a.cols[s] = cols[s];
        a.coff[s] = totc;
        a.cb0[s] = cb;
        totc += cols[s];
        cb += (cols[s] + 63) / 64;
    }
    a.cb0[nseg] = cb;
    a.totc = totc;
    a.nchunk = maxa > 0 ? (int)((maxa + SR_CHUNK - 1) / SR_CHUNK) : 1;
    a.part = work;
    MATPL_CHECK_ARG(work_elems >= (int64_t)ntypes * a.nchunk * totc, "seg_rowsum: workspace too small (%lld < %lld)",
                    (long long)work_elems, (long long)((int64_t)ntypes * a.nchunk * totc));
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(seg_rowsum_partial_kernel, dim3(cb, a.nchunk, ntypes), dim3(256), 0, st, a);
    int rc = check_launch("seg_rowsum_partial_kernel");
    if (rc) return rc;
    hipLaunchKernelGGL(xtz_reduce_kernel, dim3((unsigned)((totc + 63) / 64), ntypes), dim3(256), 0, st, work, a.nchunk,
                       (int64_t)totc, ntypes, out);
    return check_launch("xtz_reduce_kernel");
}
