// Tabulated (compressed) DP embedding net: the calculateCompress operator.
//
// G[s][i] = sum_q coef[s][i][q] * f2[s]^(K-1-q) for K = coe_num = 4 (cubic)
// or 6 (quintic) coefficients per node, highest power first, evaluated by
// Horner's rule. coef [n][L][K] holds, for each slot s, the table row of the
// interval its s_ij falls in (gathered by the caller), f2 [n] the offset of
// s_ij from that interval's left knot.
//
// Forward: one lane per (slot, node) output, so a wavefront reads 64
// consecutive K-vectors of coef (16 / 24 / 32 / 48 bytes each) and writes 64
// consecutive G values: fully coalesced; f2 is a broadcast within a slot.
// Backward: gf2[s] = sum_i gG[s][i] * dG[s][i]/df2, one lane per slot
// summing its L nodes in order (deterministic); the slot's coefficient rows
// are contiguous, so each lane streams one contiguous L*K block.
#include "common.h"

namespace matpl {
namespace {

template <typename T, int K>
__device__ __forceinline__ T horner(const T *__restrict__ c, T x) {
    T r = c[0];
#pragma unroll
    for (int q = 1; q < K; q++) r = r * x + c[q];
    return r;
}

template <typename T, int K>
__device__ __forceinline__ T horner_d(const T *__restrict__ c, T x) {
    T r = T(K - 1) * c[0];
#pragma unroll
    for (int q = 1; q < K - 1; q++) r = r * x + T(K - 1 - q) * c[q];
    return r;
}

template <typename T, int K>
__global__ void compress_kernel(const T *__restrict__ f2, const T *__restrict__ coef, T *__restrict__ G, int64_t n,
                                int L) {
    const int64_t total = n * L;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = t / L;
        G[t] = horner<T, K>(coef + t * K, f2[s]);
    }
}

template <typename T, int K>
__global__ void compress_grad_kernel(const T *__restrict__ f2, const T *__restrict__ coef, const T *__restrict__ gG,
                                     T *__restrict__ gf2, int64_t n, int L) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < n; s += (int64_t)gridDim.x * blockDim.x) {
        const T x = f2[s];
        T acc = T(0);
        for (int i = 0; i < L; i++) acc += gG[s * L + i] * horner_d<T, K>(coef + (s * L + i) * K, x);
        gf2[s] = acc;
    }
}

template <typename T>
int compress_launch(const T *f2, const T *coef, T *G, int64_t n, int L, int K, hipStream_t st) {
    MATPL_CHECK_ARG(n >= 0 && L >= 1 && (K == 4 || K == 6), "dp_compress: coe_num must be 4 or 6, L >= 1");
    if (n == 0) return OK;
    MATPL_CHECK_ARG(f2 && coef && G, "dp_compress: null pointer");
    const unsigned grid = grid_cap(n * L);
    if (K == 4) hipLaunchKernelGGL((compress_kernel<T, 4>), dim3(grid), dim3(256), 0, st, f2, coef, G, n, L);
    else hipLaunchKernelGGL((compress_kernel<T, 6>), dim3(grid), dim3(256), 0, st, f2, coef, G, n, L);
    return check_launch("compress_kernel");
}

template <typename T>
int compress_grad_launch(const T *f2, const T *coef, const T *gG, T *gf2, int64_t n, int L, int K, hipStream_t st) {
    MATPL_CHECK_ARG(n >= 0 && L >= 1 && (K == 4 || K == 6), "dp_compress_grad: coe_num must be 4 or 6, L >= 1");
    if (n == 0) return OK;
    MATPL_CHECK_ARG(f2 && coef && gG && gf2, "dp_compress_grad: null pointer");
    const unsigned grid = grid_cap(n);
    if (K == 4)
        hipLaunchKernelGGL((compress_grad_kernel<T, 4>), dim3(grid), dim3(256), 0, st, f2, coef, gG, gf2, n, L);
    else
        hipLaunchKernelGGL((compress_grad_kernel<T, 6>), dim3(grid), dim3(256), 0, st, f2, coef, gG, gf2, n, L);
    return check_launch("compress_grad_kernel");
}

}  // namespace
}  // namespace matpl

using namespace matpl;

MATPL_EXPORT int matpl_dp_compress_f32(const float *f2, const float *coef, int64_t n, int L, int coe_num, float *G,
                                       void *stream) {
    return compress_launch<float>(f2, coef, G, n, L, coe_num, (hipStream_t)stream);
}
MATPL_EXPORT int matpl_dp_compress_f64(const double *f2, const double *coef, int64_t n, int L, int coe_num, double *G,
                                       void *stream) {
    return compress_launch<double>(f2, coef, G, n, L, coe_num, (hipStream_t)stream);
}
MATPL_EXPORT int matpl_dp_compress_grad_f32(const float *f2, const float *coef, const float *gG, int64_t n, int L,
                                            int coe_num, float *gf2, void *stream) {
    return compress_grad_launch<float>(f2, coef, gG, gf2, n, L, coe_num, (hipStream_t)stream);
}
MATPL_EXPORT int matpl_dp_compress_grad_f64(const double *f2, const double *coef, const double *gG, int64_t n, int L,
                                            int coe_num, double *gf2, void *stream) {
    return compress_grad_launch<double>(f2, coef, gG, gf2, n, L, coe_num, (hipStream_t)stream);
}
