// Shared device helpers of the matrix-core DP fitting kernels (dp_fitting.hip: energies and
// dEi/dA; dp_fit_train.hip: the training step's second derivatives). Activations are
// [feature x atom] tiles in the 32x32 accumulator layout (lane l: atom n = l % 32, register r of
// block b: feature 32 b + f(r, h), f(r,h) = 8 (r/4) + 4 h + r%4, h = l / 32), hidden widths padded
// to 64 (two blocks).
#pragma once
#include "common.h"

namespace matpl {
namespace {

constexpr int WAVE = 64;
constexpr int E_MAX = 32;
constexpr int M2 = 16;

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float floatx16;

__device__ __forceinline__ constexpr int feat(int r, int h) { return 8 * (r >> 2) + 4 * h + (r & 3); }

// Per-type packed parameters (see matpl_amd/dp_fitting.py for the host packer).
//   w0f  [2 ob][64 lanes][Dh]   first layer, lane's K indices in order
//   w0b  [nb0 blk][64 lanes][32] first layer backward (rows = DR index)
//   whf  [L-1][2 ob][64][32]    hidden layers forward
//   whb  [L-1][2 ob][64][32]    hidden layers backward
//   vec  [L][3 (b, rdt, -)][2 blk][2 h][16] + out weights [2][2][16] + b_out
struct FitParams {
    const float *A;          // [N][4][E]
    float *dA;               // [N][4][E]
    float *Ei;               // [N]
    const int32_t *tiles;    // [ntiles][3] = (type, first atom, count), or (type, first position, count)
    const int32_t *perm;     // null: tiles index atoms; else [Npos] the atom at each position
    const void *w0f, *w0b, *whf, *whb;
    const float *vec;
    int64_t type_stride_w0f, type_stride_w0b, type_stride_wh, type_stride_vec;  // elements per type
    int ntiles, E, D, Dh, nb0, L;
};

// tanh: in BF16 mode 1 - 2 / (2^(2x log2 e) + 1) with the hardware exp2 and
// reciprocal (relative error ~1e-7, far below the bf16 operand rounding);
// in FP32 mode the library tanhf.
template <bool BF16>
__device__ __forceinline__ float act(float x) {
    if constexpr (BF16) {
        const float e = __builtin_amdgcn_exp2f(x * 2.8853900817779268f);
        return 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
    } else {
        return tanhf(x);
    }
}

// Per-layer tanh gates (1 - t^2) [rdt] kept for the backward, FP32 (packed
// bf16 pairs cost dA accuracy beyond the bf16-operand level and saved no
// registers here, so the kernel uses FP32).
template <bool PACK>
struct Gates16;
template <>
struct Gates16<true> {
    uint32_t v[8];
    __device__ __forceinline__ void set(int r, float g0, float g1) {
        const __bf16 a = (__bf16)g0, b = (__bf16)g1;
        v[r >> 1] = (uint32_t)__builtin_bit_cast(uint16_t, a) | ((uint32_t)__builtin_bit_cast(uint16_t, b) << 16);
    }
    __device__ __forceinline__ float get(int r) const {
        const uint32_t w = v[r >> 1];
        return (r & 1) ? __uint_as_float(w & 0xffff0000u) : __uint_as_float(w << 16);
    }
};
template <>
struct Gates16<false> {
    float v[16];
    __device__ __forceinline__ void set(int r, float g0, float g1) { v[r] = g0; v[r + 1] = g1; }
    __device__ __forceinline__ float get(int r) const { return v[r]; }
};

// FP16 pairs (BF16 mode): a gate is (1 - t^2) or rdt (1 - t^2), and multiplies a gradient that is
// rounded to a bf16 MFMA operand right after, so fp16's 11-bit significand adds error eight times
// below that rounding. Halving the gate registers brings the kernel under 256 VGPRs (two waves per
// SIMD instead of one), which hides the weight-load latency of the dependent layer chain.
struct GatesH {
    typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
    half2_t v[8];
    __device__ __forceinline__ void set(int r, float g0, float g1) {
        v[r >> 1] = half2_t{(_Float16)g0, (_Float16)g1};
    }
    __device__ __forceinline__ float get(int r) const { return (float)v[r >> 1][r & 1]; }
};

// Value of lane l ^ 32 (same lane of the other half-wave), via v_permlane32_swap.
__device__ __forceinline__ float other_half(float v) {
    const unsigned u = __float_as_uint(v);
    const auto r = __builtin_amdgcn_permlane32_swap(u, u, false, false);
    return __uint_as_float((threadIdx.x & 32) ? r[0] : r[1]);
}

template <bool BF16>
__device__ __forceinline__ void mma(floatx16 &acc, const void *w, int idx, const float *x) {
    // idx: chunk index; BF16: 8 values per chunk, FP32: 1 value per chunk
    if constexpr (BF16) {
        const bf16x8 a = reinterpret_cast<const bf16x8 *>(w)[idx];
        bf16x8 b;
#pragma unroll
        for (int q = 0; q < 8; q++) b[q] = (__bf16)x[q];
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    } else {
        const float a = reinterpret_cast<const float *>(w)[idx];
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, x[0], acc, 0, 0, 0);
    }
}

// out[ob] (+)= W^T x for a 64 -> 64 layer given the packed [2 ob][64][32] array.
template <bool BF16>
__device__ __forceinline__ void layer64(const void *w, int lane, const float x[2][16], floatx16 out[2]) {
    constexpr int CH = BF16 ? 4 : 32;          // chunks per 64-wide input
    constexpr int PER = BF16 ? 8 : 1;
#pragma unroll
    for (int ob = 0; ob < 2; ob++) {
        const char *base = reinterpret_cast<const char *>(w) + ((size_t)ob * 64 + lane) * 32 * (BF16 ? 2 : 4);
#pragma unroll
        for (int c = 0; c < CH; c++) {
            const int v = c * PER;                // first input register of the chunk
            mma<BF16>(out[ob], base, c, &x[v / 16][v % 16]);
        }
    }
}

__device__ __forceinline__ const float *vptr(const float *vec, int layer, int which, int blk, int h) {
    return vec + ((((size_t)layer * 3 + which) * 2 + blk) * 2 + h) * 16;
}

}  // namespace
}  // namespace matpl
