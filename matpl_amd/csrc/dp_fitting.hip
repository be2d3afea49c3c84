// DP fitting net on the matrix cores: A [N][4][E] -> Ei [N] and dEi/dA.
//
// Reference: DP.calculate_Ei / calculate_xyz_scater (src/model/dp_dp.py:304-415)
// and FittingNet.forward (src/model/dp_embedding.py:140-163):
//   DR[p*M2 + q] = sum_w A[w][p] A[w][q]        (p < E, q < M2)
//   x1 = tanh(DR W0 + b0);  x_{l+1} = x_l + rdt_l * tanh(x_l W_l + b_l)  (l >= 1)
//   Ei = x_L . w_out + b_out
// and the backward dEi/dA of that chain (the reference gets it from autograd).
//
// One wavefront owns a tile of 32 atoms of one centre type. Activations are
// [feature x atom] tiles in the 32x32 accumulator layout (lane l: atom
// n = l % 32, register r of block b: feature 32 b + f(r, h), f(r,h) =
// 8 (r/4) + 4 h + r%4, h = l / 32), so a layer's output registers are the next
// layer's B operand with no data movement; hidden widths are padded to 64
// (two blocks). DR is formed on the fly in the B-operand layout of the first
// layer (K-step s of 16 = DR row p = s when M2 = 16), and dDR is contracted into
// dA block by block without being stored.
#include "dp_fitting_common.h"

namespace matpl {
namespace {

template <bool BF16, int L>
__global__ void __launch_bounds__(WAVE, 2)
dp_fitting_kernel(FitParams p) {
    const int lane = threadIdx.x;
    const int n = lane & 31, h = lane >> 5;
    // A of the tile's 32 atoms, one row of 4E values per atom, row stride 4E + 1
    // (odd) so that lanes reading the same element of different atoms hit
    // different banks. The layer-0 backward overwrites column p of A with dA
    // once the DR block that reads column p is done (blocks go in increasing p),
    // so dA needs no buffer of its own: 12.9 KB of LDS per wave for E = 25.
    extern __shared__ float fit_smem[];
    const int E = p.E, RS = 4 * E + 1;
    float *As = fit_smem, *dAs = fit_smem;
#define AS(a, w, c) As[(a) * RS + (w) * E + (c)]
#define DAS(a, w, c) dAs[(a) * RS + (w) * E + (c)]
    for (int tile = blockIdx.x; tile < p.ntiles; tile += gridDim.x) {
        const int t = p.tiles[3 * tile], a0 = p.tiles[3 * tile + 1], cnt = p.tiles[3 * tile + 2];
        // stage A
        // (atom, offset) advance by 64 elements per step without a division (4E >= 64)
        for (int x = lane, an = 0, wc = lane; x < 32 * 4 * E; x += WAVE) {
            const int64_t src = p.perm ? (int64_t)(an < cnt ? p.perm[a0 + an] : 0) * 4 * E + wc : (int64_t)a0 * 4 * E + x;
            As[an * RS + wc] = an < cnt ? p.A[src] : 0.0f;
            wc += WAVE;
            if (wc >= 4 * E) { wc -= 4 * E; an++; }
        }
        __syncthreads();
        const void *w0f = reinterpret_cast<const char *>(p.w0f) + (size_t)t * p.type_stride_w0f * (BF16 ? 2 : 4);
        const void *w0b = reinterpret_cast<const char *>(p.w0b) + (size_t)t * p.type_stride_w0b * (BF16 ? 2 : 4);
        const char *whf = reinterpret_cast<const char *>(p.whf) + (size_t)t * p.type_stride_wh * (BF16 ? 2 : 4);
        const char *whb = reinterpret_cast<const char *>(p.whb) + (size_t)t * p.type_stride_wh * (BF16 ? 2 : 4);
        const float *vec = p.vec + (size_t)t * p.type_stride_vec;
        const size_t lstride = (size_t)2 * 64 * 32 * (BF16 ? 2 : 4);

        // ---- layer 0: x1 = tanh(DR W0 + b0)
        floatx16 acc[2];
#pragma unroll
        for (int ob = 0; ob < 2; ob++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[ob][r] = vptr(vec, 0, 0, ob, h)[r];
        if constexpr (BF16) {
            // K-step s covers DR rows p = s (M2 = 16), this lane's 8 columns q = 8h..8h+7
            float aq[4][8];
#pragma unroll
            for (int w = 0; w < 4; w++)
#pragma unroll
                for (int q = 0; q < 8; q++) aq[w][q] = AS(n, w, 8 * h + q);
            // K-steps go in pairs (s, s + 1) with two register sets of weight chunks; each set is
            // reloaded right after its MFMA has read it, one pair ahead of its next use, so the
            // global-memory latency overlaps a pair of steps (no moves of in-flight registers).
            // A step beyond E (odd E) multiplies zeros.
            const bf16x8 *wr0 = reinterpret_cast<const bf16x8 *>(reinterpret_cast<const char *>(w0f) +
                                                                 (size_t)lane * p.Dh * 2);
            const bf16x8 *wr1 = reinterpret_cast<const bf16x8 *>(reinterpret_cast<const char *>(w0f) +
                                                                 ((size_t)64 + lane) * p.Dh * 2);
            const int Em = E - 1;
            bf16x8 a0 = wr0[0], a1 = wr1[0];
            bf16x8 b0 = wr0[1 < Em ? 1 : Em], b1 = wr1[1 < Em ? 1 : Em];
            // odd K-steps accumulate separately (two independent MFMA chains per output block)
            floatx16 accb[2];
#pragma unroll
            for (int ob = 0; ob < 2; ob++)
#pragma unroll
                for (int r = 0; r < 16; r++) accb[ob][r] = 0.0f;
            for (int s = 0; s < E; s += 2) {
#pragma unroll
                for (int half = 0; half < 2; half++) {
                    const int st = s + half;
                    const int sc = st < Em ? st : Em;
                    const float live = st < E ? 1.0f : 0.0f;
                    const float ap0 = AS(n, 0, sc) * live, ap1 = AS(n, 1, sc) * live, ap2 = AS(n, 2, sc) * live,
                                ap3 = AS(n, 3, sc) * live;
                    bf16x8 b;
#pragma unroll
                    for (int q = 0; q < 8; q++)
                        b[q] = (__bf16)(ap0 * aq[0][q] + ap1 * aq[1][q] + ap2 * aq[2][q] + ap3 * aq[3][q]);
                    const int sn = st + 2 < Em ? st + 2 : Em;
                    if (half == 0) {
                        acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b, acc[0], 0, 0, 0);
                        acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b, acc[1], 0, 0, 0);
                        a0 = wr0[sn];
                        a1 = wr1[sn];
                        __builtin_amdgcn_sched_barrier(0);   // keep the prefetch here
                    } else {
                        accb[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b0, b, accb[0], 0, 0, 0);
                        accb[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b1, b, accb[1], 0, 0, 0);
                        b0 = wr0[sn];
                        b1 = wr1[sn];
                        __builtin_amdgcn_sched_barrier(0);
                    }
                }
            }
#pragma unroll
            for (int ob = 0; ob < 2; ob++)
#pragma unroll
                for (int r = 0; r < 16; r++) acc[ob][r] += accb[ob][r];
        } else {
            // K-step s covers DR index k = 2 s + h
            for (int s = 0; s < p.D / 2; s++) {
                const int k = 2 * s + h;
                const int pp = k / M2, qq = k % M2;
                float drv = AS(n, 0, pp) * AS(n, 0, qq) + AS(n, 1, pp) * AS(n, 1, qq) +
                            AS(n, 2, pp) * AS(n, 2, qq) + AS(n, 3, pp) * AS(n, 3, qq);
#pragma unroll
                for (int ob = 0; ob < 2; ob++)
                    mma<false>(acc[ob], reinterpret_cast<const char *>(w0f) + ((size_t)ob * 64 + lane) * p.Dh * 4, s,
                               &drv);
            }
        }
        // gates per layer: (1 - t^2) for layer 0, rdt (1 - t^2) after
        typename std::conditional<BF16, GatesH, Gates16<false>>::type gk[L][2];
        float x[2][16];
#pragma unroll
        for (int ob = 0; ob < 2; ob++)
#pragma unroll
            for (int r = 0; r < 16; r += 2) {
                const float ta = act<BF16>(acc[ob][r]), tb = act<BF16>(acc[ob][r + 1]);
                x[ob][r] = ta;
                x[ob][r + 1] = tb;
                gk[0][ob].set(r, 1.0f - ta * ta, 1.0f - tb * tb);
            }
        // ---- hidden residual layers
#pragma unroll
        for (int l = 1; l < L; l++) {
#pragma unroll
            for (int ob = 0; ob < 2; ob++)
#pragma unroll
                for (int r = 0; r < 16; r++) acc[ob][r] = vptr(vec, l, 0, ob, h)[r];
            layer64<BF16>(whf + (l - 1) * lstride, lane, x, acc);
#pragma unroll
            for (int ob = 0; ob < 2; ob++) {
                const float *rdt = vptr(vec, l, 1, ob, h);
#pragma unroll
                for (int r = 0; r < 16; r += 2) {
                    const float ta = act<BF16>(acc[ob][r]), tb = act<BF16>(acc[ob][r + 1]);
                    x[ob][r] += rdt[r] * ta;
                    x[ob][r + 1] += rdt[r + 1] * tb;
                    gk[l][ob].set(r, rdt[r] * (1.0f - ta * ta), rdt[r + 1] * (1.0f - tb * tb));
                }
            }
        }
        // ---- output
        const float *wout = vec + (size_t)L * 3 * 2 * 2 * 16;        // [2 blk][2 h][16]
        const float bout = wout[2 * 2 * 16];
        float e = 0.0f;
#pragma unroll
        for (int ob = 0; ob < 2; ob++)
#pragma unroll
            for (int r = 0; r < 16; r++) e += x[ob][r] * wout[(ob * 2 + h) * 16 + r];
        e += __shfl_xor(e, 32, 64);
        if (h == 0 && n < cnt) p.Ei[p.perm ? p.perm[a0 + n] : a0 + n] = e + bout;
        // ---- backward: dx = w_out
        float dx[2][16];
#pragma unroll
        for (int ob = 0; ob < 2; ob++)
#pragma unroll
            for (int r = 0; r < 16; r++) dx[ob][r] = wout[(ob * 2 + h) * 16 + r];
#pragma unroll
        for (int l = L - 1; l >= 1; l--) {
            float dz[2][16];
#pragma unroll
            for (int ob = 0; ob < 2; ob++) {
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    dz[ob][r] = dx[ob][r] * gk[l][ob].get(r);
                    acc[ob][r] = dx[ob][r];
                }
            }
            layer64<BF16>(whb + (l - 1) * lstride, lane, dz, acc);
#pragma unroll
            for (int ob = 0; ob < 2; ob++)
#pragma unroll
                for (int r = 0; r < 16; r++) dx[ob][r] = acc[ob][r];
        }
        float dz0[2][16];
#pragma unroll
        for (int ob = 0; ob < 2; ob++)
#pragma unroll
            for (int r = 0; r < 16; r++) dz0[ob][r] = dx[ob][r] * gk[0][ob].get(r);
        // ---- layer 0 backward, block by block: dDR rows 32 blk + f(r,h)
        //   k = p*16 + q: r < 8 -> p = 2 blk, q = f(r,h); r >= 8 -> p = 2 blk + 1, q = f(r-8,h)
        float aq2[4][8];            // A[w][f(r,h)] for r < 8
#pragma unroll
        for (int w = 0; w < 4; w++)
#pragma unroll
            for (int r = 0; r < 8; r++) aq2[w][r] = AS(n, w, feat(r, h));
        float d2[4][8];             // dA[w][q] of the q-side (q = f(r,h)), over all blocks
#pragma unroll
        for (int w = 0; w < 4; w++)
#pragma unroll
            for (int r = 0; r < 8; r++) d2[w][r] = 0.0f;
        // VALU part of one DR block: its rows p0 = 2 blk, p1 = 2 blk + 1 (dA of the p side, summed over
        // both halves of the wave) and the q-side terms (kept in d2)
        auto block_dA = [&](int blk, const floatx16 &g) {
            const int p0 = 2 * blk, p1 = 2 * blk + 1;
            float s0[4] = {0.f, 0.f, 0.f, 0.f}, s1[4] = {0.f, 0.f, 0.f, 0.f};
            const float ap0[4] = {p0 < E ? AS(n, 0, p0) : 0.f, p0 < E ? AS(n, 1, p0) : 0.f,
                                  p0 < E ? AS(n, 2, p0) : 0.f, p0 < E ? AS(n, 3, p0) : 0.f};
            const float ap1[4] = {p1 < E ? AS(n, 0, p1) : 0.f, p1 < E ? AS(n, 1, p1) : 0.f,
                                  p1 < E ? AS(n, 2, p1) : 0.f, p1 < E ? AS(n, 3, p1) : 0.f};
#pragma unroll
            for (int r = 0; r < 8; r++) {
#pragma unroll
                for (int w = 0; w < 4; w++) {
                    s0[w] += g[r] * aq2[w][r];
                    s1[w] += g[r + 8] * aq2[w][r];
                    d2[w][r] += g[r] * ap0[w] + g[r + 8] * ap1[w];
                }
            }
#pragma unroll
            for (int w = 0; w < 4; w++) {
                s0[w] += other_half(s0[w]);
                s1[w] += other_half(s1[w]);
            }
            if (h == 0 && p0 < E) {
#pragma unroll
                for (int w = 0; w < 4; w++) DAS(n, w, p0) = s0[w];
            }
            if (h == 1 && p1 < E) {
#pragma unroll
                for (int w = 0; w < 4; w++) DAS(n, w, p1) = s1[w];
            }
        };
        if constexpr (BF16) {
            // dz0 as bf16 B operands once; blocks in pairs (two independent MFMA chains), the weight
            // chunks of the next pair loaded while this pair computes
            bf16x8 dzb[4];
#pragma unroll
            for (int c = 0; c < 4; c++)
#pragma unroll
                for (int q = 0; q < 8; q++) dzb[c][q] = (__bf16)dz0[(c * 8) / 16][(c * 8) % 16 + q];
            const int nbm = p.nb0 - 1;
            auto wrow = [&](int b) {
                return reinterpret_cast<const bf16x8 *>(reinterpret_cast<const char *>(w0b) +
                                                        ((size_t)b * 64 + lane) * 32 * 2);
            };
            bf16x8 nA[4], nB[4];
            {
                const bf16x8 *ra = wrow(0), *rb = wrow(1 < nbm ? 1 : nbm);
#pragma unroll
                for (int c = 0; c < 4; c++) { nA[c] = ra[c]; nB[c] = rb[c]; }
            }
            for (int blk = 0; blk < p.nb0; blk += 2) {
                bf16x8 wA[4], wB[4];
#pragma unroll
                for (int c = 0; c < 4; c++) { wA[c] = nA[c]; wB[c] = nB[c]; }
                {
                    const int ba = blk + 2 < nbm ? blk + 2 : nbm, bb = blk + 3 < nbm ? blk + 3 : nbm;
                    const bf16x8 *ra = wrow(ba), *rb = wrow(bb);
#pragma unroll
                    for (int c = 0; c < 4; c++) { nA[c] = ra[c]; nB[c] = rb[c]; }
                }
                floatx16 gA, gB;
#pragma unroll
                for (int r = 0; r < 16; r++) { gA[r] = 0.0f; gB[r] = 0.0f; }
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    gA = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wA[c], dzb[c], gA, 0, 0, 0);
                    gB = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wB[c], dzb[c], gB, 0, 0, 0);
                }
                block_dA(blk, gA);
                if (blk + 1 < p.nb0) block_dA(blk + 1, gB);
            }
        } else {
            for (int blk = 0; blk < p.nb0; blk++) {
                floatx16 g;
#pragma unroll
                for (int r = 0; r < 16; r++) g[r] = 0.0f;
                const char *wb = reinterpret_cast<const char *>(w0b) + ((size_t)blk * 64 + lane) * 32 * 4;
#pragma unroll
                for (int c = 0; c < 32; c++) mma<false>(g, wb, c, &dz0[c / 16][c % 16]);
                block_dA(blk, g);
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 8; r++) {
            const int q = feat(r, h);
#pragma unroll
            for (int w = 0; w < 4; w++) DAS(n, w, q) += d2[w][r];
        }
        __syncthreads();
        for (int x2 = lane, an = 0, wc = lane; x2 < 32 * 4 * E; x2 += WAVE) {
            if (an < cnt)
                p.dA[p.perm ? (int64_t)p.perm[a0 + an] * 4 * E + wc : (int64_t)a0 * 4 * E + x2] = dAs[an * RS + wc];
            wc += WAVE;
            if (wc >= 4 * E) { wc -= 4 * E; an++; }
        }
        __syncthreads();
    }
#undef AS
#undef DAS
}

}  // namespace
}  // namespace matpl

using namespace matpl;

static int fitting_launch(const float *A, const int32_t *tiles, const int32_t *perm, int ntiles, const void *w0f,
                          const void *w0b, const void *whf, const void *whb, const float *vec,
                          int64_t type_stride_w0f, int64_t type_stride_w0b, int64_t type_stride_wh,
                          int64_t type_stride_vec, int E, int nlayers, int bf16, float *Ei, float *dA,
                          void *stream) {
    MATPL_CHECK_ARG(E >= M2 && E <= E_MAX, "dp_fitting: E must be in [%d,%d]", M2, E_MAX);
    MATPL_CHECK_ARG(nlayers >= 1 && nlayers <= 4, "dp_fitting: hidden layers must be in [1,4]");
    if (ntiles <= 0) return OK;
    MATPL_CHECK_ARG(A && tiles && w0f && w0b && vec && Ei && dA && (nlayers == 1 || (whf && whb)),
                    "dp_fitting: null pointer");
    FitParams p{};
    p.A = A; p.dA = dA; p.Ei = Ei; p.tiles = tiles; p.perm = perm;
    p.w0f = w0f; p.w0b = w0b; p.whf = whf; p.whb = whb; p.vec = vec;
    p.type_stride_w0f = type_stride_w0f; p.type_stride_w0b = type_stride_w0b;
    p.type_stride_wh = type_stride_wh; p.type_stride_vec = type_stride_vec;
    p.ntiles = ntiles; p.E = E; p.D = E * M2; p.Dh = (E * M2) / 2; p.nb0 = (E * M2 + 31) / 32; p.L = nlayers;
    const int grid = ntiles < 8192 ? ntiles : 8192;
    const size_t lds = sizeof(float) * 32 * (4 * E + 1);
    hipStream_t st = (hipStream_t)stream;
#define MATPL_FIT_CASE(LL)                                                                              \
    case LL:                                                                                            \
        if (bf16) hipLaunchKernelGGL((dp_fitting_kernel<true, LL>), dim3(grid), dim3(WAVE), lds, st, p);\
        else hipLaunchKernelGGL((dp_fitting_kernel<false, LL>), dim3(grid), dim3(WAVE), lds, st, p);\
        break;
    switch (nlayers) {
        MATPL_FIT_CASE(1)
        MATPL_FIT_CASE(2)
        MATPL_FIT_CASE(3)
        MATPL_FIT_CASE(4)
        default:
            set_error("dp_fitting: %d hidden layers not compiled (1..4)", nlayers);
            return ERR_ARGUMENT;
    }
#undef MATPL_FIT_CASE
    return check_launch("dp_fitting_kernel");
}

MATPL_EXPORT int matpl_dp_fitting(const float *A, const int32_t *tiles, int ntiles, const void *w0f, const void *w0b,
                                  const void *whf, const void *whb, const float *vec, int64_t type_stride_w0f,
                                  int64_t type_stride_w0b, int64_t type_stride_wh, int64_t type_stride_vec, int E,
                                  int nlayers, int bf16, float *Ei, float *dA, void *stream) {
    return fitting_launch(A, tiles, nullptr, ntiles, w0f, w0b, whf, whb, vec, type_stride_w0f, type_stride_w0b,
                          type_stride_wh, type_stride_vec, E, nlayers, bf16, Ei, dA, stream);
}

MATPL_EXPORT int matpl_dp_fitting_perm(const float *A, const int32_t *tiles, const int32_t *perm, int ntiles,
                                       const void *w0f, const void *w0b, const void *whf, const void *whb,
                                       const float *vec, int64_t type_stride_w0f, int64_t type_stride_w0b,
                                       int64_t type_stride_wh, int64_t type_stride_vec, int E, int nlayers, int bf16,
                                       float *Ei, float *dA, void *stream) {
    MATPL_CHECK_ARG(perm != nullptr || ntiles <= 0, "dp_fitting_perm: null perm");
    return fitting_launch(A, tiles, perm, ntiles, w0f, w0b, whf, whb, vec, type_stride_w0f, type_stride_w0b,
                          type_stride_wh, type_stride_vec, E, nlayers, bf16, Ei, dA, stream);
}
