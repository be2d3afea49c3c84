// Batched products of small matrices for the training paths' per-atom
// contractions (A = Ri^T G over a bucket of neighbor slots, DR = A^T A', and
// the products their first and second derivatives need). One wavefront owns
// one batch element at a time: it copies both operands into its LDS slice
// (contiguous, coalesced), then lanes compute output entries, each a short
// dot product read from LDS, and store them contiguously.
//
//   mode 0 (TN): out[n][a][b] = sum_k P1[n][k][a] * P2[n][k][b]   P1 [K][a], P2 [K][b]
//   mode 1 (NT): out[n][k][a] = sum_j P1[n][k][j] * P2[n][a][j]   P1 [K][b], P2 [a][b]
//   mode 2 (NN): out[n][k][b] = sum_i P1[n][k][i] * P2[n][i][b]   P1 [K][a], P2 [a][b]
#include "common.h"

#pragma clang fp contract(off)

namespace matpl {
namespace {

constexpr int WAVE = 64;

template <typename T, int MODE>
__global__ void __launch_bounds__(256)
bsmm_kernel(const T *__restrict__ P1, const T *__restrict__ P2, T *__restrict__ out, int64_t n, int K, int a, int b,
            int waves) {
    extern __shared__ unsigned char smem_raw[];
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    const int s1 = MODE == 1 ? K * b : K * a;
    const int s2 = MODE == 0 ? K * b : a * b;
    const int so = MODE == 0 ? a * b : (MODE == 1 ? K * a : K * b);
    T *L1 = reinterpret_cast<T *>(smem_raw) + (size_t)wid * (s1 + s2);
    T *L2 = L1 + s1;
    for (int64_t e = (int64_t)blockIdx.x * waves + wid; e < n; e += (int64_t)gridDim.x * waves) {
        const T *g1 = P1 + e * s1, *g2 = P2 + e * s2;
        for (int x = lane; x < s1; x += WAVE) L1[x] = g1[x];
        for (int x = lane; x < s2; x += WAVE) L2[x] = g2[x];
        __builtin_amdgcn_wave_barrier();
        T *o = out + e * so;
        for (int x = lane; x < so; x += WAVE) {
            T acc = T(0);
            if (MODE == 0) {
                const int i = x / b, j = x - i * b;
                for (int k = 0; k < K; k++) acc += L1[k * a + i] * L2[k * b + j];
            } else if (MODE == 1) {
                const int k = x / a, i = x - k * a;
                for (int j = 0; j < b; j++) acc += L1[k * b + j] * L2[i * b + j];
            } else {
                const int k = x / b, j = x - k * b;
                for (int i = 0; i < a; i++) acc += L1[k * a + i] * L2[i * b + j];
            }
            o[x] = acc;
        }
        __builtin_amdgcn_wave_barrier();
    }
}

template <typename T, int MODE>
int bsmm_launch(const T *P1, const T *P2, T *out, int64_t n, int K, int a, int b, hipStream_t st) {
    const size_t s1 = MODE == 1 ? (size_t)K * b : (size_t)K * a;
    const size_t s2 = MODE == 0 ? (size_t)K * b : (size_t)a * b;
    int waves = 4;
    while (waves > 1 && (size_t)waves * (s1 + s2) * sizeof(T) > 64 * 1024) waves /= 2;
    const size_t lds = (size_t)waves * (s1 + s2) * sizeof(T);
    MATPL_CHECK_ARG(lds <= 64 * 1024, "bsmm: operands of %zu + %zu elements exceed LDS", s1, s2);
    int64_t blocks = (n + waves - 1) / waves;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL((bsmm_kernel<T, MODE>), dim3((unsigned)blocks), dim3(waves * WAVE), lds, st, P1, P2, out, n,
                       K, a, b, waves);
    return check_launch("bsmm_kernel");
}

template <typename T>
int bsmm_dispatch(int mode, const void *P1, const void *P2, void *out, int64_t n, int K, int a, int b,
                  hipStream_t st) {
    const T *p1 = (const T *)P1, *p2 = (const T *)P2;
    T *o = (T *)out;
    if (mode == 0) return bsmm_launch<T, 0>(p1, p2, o, n, K, a, b, st);
    if (mode == 1) return bsmm_launch<T, 1>(p1, p2, o, n, K, a, b, st);
    return bsmm_launch<T, 2>(p1, p2, o, n, K, a, b, st);
}

}  // namespace
}  // namespace matpl

using namespace matpl;

MATPL_EXPORT int matpl_bsmm(int f64, int mode, const void *P1, const void *P2, void *out, int64_t n, int K, int a,
                            int b, void *stream) {
    MATPL_CHECK_ARG(mode >= 0 && mode <= 2, "bsmm: mode is 0 (TN), 1 (NT) or 2 (NN)");
    MATPL_CHECK_ARG(n >= 0 && K >= 0 && a >= 0 && b >= 0, "bsmm: negative size");
    const int64_t so = mode == 0 ? (int64_t)a * b : (mode == 1 ? (int64_t)K * a : (int64_t)K * b);
    if (n == 0 || so == 0) return OK;
    MATPL_CHECK_ARG(out && P1 && P2, "bsmm: null pointer");
    hipStream_t st = (hipStream_t)stream;
    return f64 ? bsmm_dispatch<double>(mode, P1, P2, out, n, K, a, b, st)
               : bsmm_dispatch<float>(mode, P1, P2, out, n, K, a, b, st);
}
