// DP fitting nets in the training step: the second derivatives the force loss needs, on the matrix
// cores.
//
// The reference trains through dE = autograd.grad(Etot, Ri, create_graph=True) (src/model/
// dp_dp.py:429) and loss.backward() (src/PWMLFF/dp_mods/dp_trainer.py:274), which for the fitting
// nets (src/model/dp_embedding.py:96-163; DR = A^T A', dp_dp.py:304-342) is a double backward. The
// force loss reaches the fitting side only through <Abar, dA_t>, Abar = dE/dA and dA_t the tangent
// of the descriptor along the force residual (computed by the embedding kernels), and
// <Abar_i, dA_t_i> = Edot_i, the forward-mode derivative of E_i along dA_t_i. So per atom this
// kernel runs the net forward with tangents (values x, tangents xdot) and then the reverse pass of
//   L_i = gE_i E_i - Edot_i        (gE = dL/dEi of the energy loss)
// giving Ahat_i = dL_i/dA_i and, per layer, the rows whose sums over atoms are the parameter
// gradients:
//   layer z = x_in W + b, zdot = xdot_in W, t = tanh z, x = x_in + r t, xdot = xdot_in + r (1-t^2) zdot
//   tbar = r xbar - 2 r t zdot xdotbar,  zdotbar = r (1-t^2) xdotbar,  zbar = (1-t^2) tbar
//   rbar += t xbar + (1-t^2) zdot xdotbar,  Wbar += x_in^T zbar + xdot_in^T zdotbar,  bbar += zbar
//   xbar_in = xbar + W zbar,  xdotbar_in = xdotbar + W zdotbar
// (layer 0: x = t, no r; output: xbar = gE w_out, xdotbar = -w_out, w_outbar += gE x - xdot), and
// at DR: Ahat[w][p] = sum_q DRbar[p][q] A[w][q] + sum_p' DRbar[p'][p] A[w][p'] + the same with
// DRdotbar and dA_t. The rows (x_in, xdot_in, zbar, zdotbar, ...) are written per atom in perm
// order, so that each type's atoms are one contiguous block; the caller sums them (products over
// atoms for the weights).
//
// Layout as the energy kernel (dp_fitting.hip): one wavefront per tile of 32 atoms of one type,
// activations in the 32x32 accumulator layout, hidden widths padded to 64, the weight packing of
// matpl_amd/dp_fitting.py. Per layer the tanh output t and the tangent zdot go to a scratch row of
// the atom for the reverse pass (the lane that wrote them reads them back).
#include "dp_fitting_common.h"

namespace matpl {
namespace {

struct FitTrainParams {
    const float *A;          // [N][4][E]
    const float *At;         // [N][4][E] tangent of A
    const float *gE;         // [N] dL/dEi of the energy loss (may be null: 0)
    const int32_t *perm;     // [Npos] atom of each position; tiles index positions
    const int32_t *tiles;    // [ntiles][3] type, first position, count (<= 32)
    const void *w0f, *w0b, *whf, *whb;
    const float *vec;
    int64_t type_stride_w0f, type_stride_w0b, type_stride_wh, type_stride_vec;
    float *Ahat;             // [N][4][E]
    float *xw0;              // [Npos][2][D]: DR, DRdot
    float *zw0;              // [Npos][2][64]: zbar, zdotbar of layer 0
    float *xh;               // [Npos][L-1][2][64]: x_in, xdot_in of hidden layer l = 1..L-1
    float *zh;               // [Npos][L-1][2][64]: zbar, zdotbar of hidden layer l
    float *rb;               // [Npos][L-1][64]: rbar terms of hidden layer l
    float *xo;               // [Npos][64]: gE x - xdot (output weights)
    float *tz;               // [Npos][L][2 (t, zdot)][2 (h)][32] scratch in the lanes' register order
    int ntiles, E, D, Dh, nb0, L;
};

// row[32 ob + f(r, h)] = v[ob][r] for the 64 features of this lane's half (float4 stores)
__device__ __forceinline__ void store_feat(float *row, const float v[2][16], int h) {
#pragma unroll
    for (int ob = 0; ob < 2; ob++)
#pragma unroll
        for (int q = 0; q < 4; q++)
            *reinterpret_cast<float4 *>(row + 32 * ob + 8 * q + 4 * h) =
                make_float4(v[ob][4 * q], v[ob][4 * q + 1], v[ob][4 * q + 2], v[ob][4 * q + 3]);
}

// the lane's 32 values of one (layer, which) scratch slot: [tile-local atom][h] chunks of 32 floats
__device__ __forceinline__ float *tz_slot(const FitTrainParams &p, int64_t pos, int layer, int which, int h) {
    return p.tz + (((pos * p.L + layer) * 2 + which) * 2 + h) * 32;
}
__device__ __forceinline__ void tz_put(float *s, const float v[2][16]) {
#pragma unroll
    for (int ob = 0; ob < 2; ob++)
#pragma unroll
        for (int q = 0; q < 4; q++)
            reinterpret_cast<float4 *>(s)[ob * 4 + q] =
                make_float4(v[ob][4 * q], v[ob][4 * q + 1], v[ob][4 * q + 2], v[ob][4 * q + 3]);
}
// dead lanes (positions past the tile's count) neither write nor read the scratch: they read zeros
__device__ __forceinline__ void tz_get(const float *s, float v[2][16], bool live) {
#pragma unroll
    for (int ob = 0; ob < 2; ob++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const float4 x = live ? reinterpret_cast<const float4 *>(s)[ob * 4 + q] : make_float4(0.f, 0.f, 0.f, 0.f);
            v[ob][4 * q] = x.x; v[ob][4 * q + 1] = x.y; v[ob][4 * q + 2] = x.z; v[ob][4 * q + 3] = x.w;
        }
}

template <bool BF16, int L>
__global__ void __launch_bounds__(WAVE)
dp_fit_train_kernel(FitTrainParams p) {
    const int lane = threadIdx.x;
    const int n = lane & 31, h = lane >> 5;
    // A and its tangent of the tile's 32 atoms (row stride 4E + 1, odd); the layer-0 reverse pass
    // accumulates Ahat in a third array
    extern __shared__ float fit_smem[];
    const int E = p.E, RS = 4 * E + 1;
    float *As = fit_smem, *Ts = fit_smem + 32 * RS, *Hs = fit_smem + 64 * RS;
#define AS(a, w, c) As[(a) * RS + (w) * E + (c)]
#define TS(a, w, c) Ts[(a) * RS + (w) * E + (c)]
#define HS(a, w, c) Hs[(a) * RS + (w) * E + (c)]
    for (int tile = blockIdx.x; tile < p.ntiles; tile += gridDim.x) {
        const int t = p.tiles[3 * tile], s0 = p.tiles[3 * tile + 1], cnt = p.tiles[3 * tile + 2];
        for (int x = lane; x < 32 * 4 * E; x += WAVE) {
            const int an = x / (4 * E), wc = x - an * 4 * E;
            float a = 0.0f, b = 0.0f;
            if (an < cnt) {
                const int64_t atom = p.perm[s0 + an];
                a = p.A[atom * 4 * E + wc];
                b = p.At[atom * 4 * E + wc];
            }
            As[an * RS + wc] = a;
            Ts[an * RS + wc] = b;
            Hs[an * RS + wc] = 0.0f;
        }
        __syncthreads();
        const bool live = n < cnt;
        const int64_t pos = s0 + (live ? n : 0);
        const int64_t atom = p.perm[pos];
        const float ge = (live && p.gE) ? p.gE[atom] : 0.0f;
        const size_t es = BF16 ? 2 : 4;
        const void *w0f = reinterpret_cast<const char *>(p.w0f) + (size_t)t * p.type_stride_w0f * es;
        const void *w0b = reinterpret_cast<const char *>(p.w0b) + (size_t)t * p.type_stride_w0b * es;
        const char *whf = reinterpret_cast<const char *>(p.whf) + (size_t)t * p.type_stride_wh * es;
        const char *whb = reinterpret_cast<const char *>(p.whb) + (size_t)t * p.type_stride_wh * es;
        const float *vec = p.vec + (size_t)t * p.type_stride_vec;
        const size_t lstride = (size_t)2 * 64 * 32 * es;
        float *xw0 = p.xw0 + pos * 2 * p.D;

        // ---- layer 0: z = DR W0 + b0, zdot = DRdot W0 (DR = A^T A', DRdot = At^T A' + A^T At')
        floatx16 acc[2], tac[2];
#pragma unroll
        for (int ob = 0; ob < 2; ob++)
#pragma unroll
            for (int r = 0; r < 16; r++) { acc[ob][r] = vptr(vec, 0, 0, ob, h)[r]; tac[ob][r] = 0.0f; }
        if constexpr (BF16) {
            // K-step s: DR row p = s, this lane's columns q = 8h..8h+7
            float aq[4][8], tq[4][8];
#pragma unroll
            for (int w = 0; w < 4; w++)
#pragma unroll
                for (int q = 0; q < 8; q++) { aq[w][q] = AS(n, w, 8 * h + q); tq[w][q] = TS(n, w, 8 * h + q); }
            const bf16x8 *wr0 = reinterpret_cast<const bf16x8 *>(reinterpret_cast<const char *>(w0f) +
                                                                 (size_t)lane * p.Dh * 2);
            const bf16x8 *wr1 = reinterpret_cast<const bf16x8 *>(reinterpret_cast<const char *>(w0f) +
                                                                 ((size_t)64 + lane) * p.Dh * 2);
            for (int s = 0; s < E; s++) {
                const float ap[4] = {AS(n, 0, s), AS(n, 1, s), AS(n, 2, s), AS(n, 3, s)};
                const float tp[4] = {TS(n, 0, s), TS(n, 1, s), TS(n, 2, s), TS(n, 3, s)};
                float dr[8], drt[8];
                bf16x8 b, bt;
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    dr[q] = ap[0] * aq[0][q] + ap[1] * aq[1][q] + ap[2] * aq[2][q] + ap[3] * aq[3][q];
                    drt[q] = tp[0] * aq[0][q] + tp[1] * aq[1][q] + tp[2] * aq[2][q] + tp[3] * aq[3][q] +
                             ap[0] * tq[0][q] + ap[1] * tq[1][q] + ap[2] * tq[2][q] + ap[3] * tq[3][q];
                    b[q] = (__bf16)dr[q];
                    bt[q] = (__bf16)drt[q];
                }
                if (live) {
                    float *o = xw0 + s * M2 + 8 * h;
                    reinterpret_cast<float4 *>(o)[0] = make_float4(dr[0], dr[1], dr[2], dr[3]);
                    reinterpret_cast<float4 *>(o)[1] = make_float4(dr[4], dr[5], dr[6], dr[7]);
                    o += p.D;
                    reinterpret_cast<float4 *>(o)[0] = make_float4(drt[0], drt[1], drt[2], drt[3]);
                    reinterpret_cast<float4 *>(o)[1] = make_float4(drt[4], drt[5], drt[6], drt[7]);
                }
                const bf16x8 a0 = wr0[s], a1 = wr1[s];
                acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b, acc[0], 0, 0, 0);
                acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b, acc[1], 0, 0, 0);
                tac[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, bt, tac[0], 0, 0, 0);
                tac[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, bt, tac[1], 0, 0, 0);
            }
        } else {
            // K-step s: DR index k = 2 s + h
            for (int s = 0; s < p.D / 2; s++) {
                const int k = 2 * s + h;
                const int pp = k / M2, qq = k % M2;
                float drv = 0.0f, drt = 0.0f;
#pragma unroll
                for (int w = 0; w < 4; w++) {
                    drv += AS(n, w, pp) * AS(n, w, qq);
                    drt += TS(n, w, pp) * AS(n, w, qq) + AS(n, w, pp) * TS(n, w, qq);
                }
                if (live) {
                    xw0[k] = drv;
                    xw0[p.D + k] = drt;
                }
#pragma unroll
                for (int ob = 0; ob < 2; ob++) {
                    const char *wb = reinterpret_cast<const char *>(w0f) + ((size_t)ob * 64 + lane) * p.Dh * 4;
                    mma<false>(acc[ob], wb, s, &drv);
                    mma<false>(tac[ob], wb, s, &drt);
                }
            }
        }
        float x[2][16], xd[2][16];
        {
            float tv[2][16], zd[2][16];
#pragma unroll
            for (int ob = 0; ob < 2; ob++)
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const float ta = act<BF16>(acc[ob][r]);
                    tv[ob][r] = ta;
                    zd[ob][r] = tac[ob][r];
                    x[ob][r] = ta;
                    xd[ob][r] = (1.0f - ta * ta) * tac[ob][r];
                }
            if (live) {
                tz_put(tz_slot(p, pos, 0, 0, h), tv);
                tz_put(tz_slot(p, pos, 0, 1, h), zd);
            }
        }
        // ---- hidden residual layers
#pragma unroll
        for (int l = 1; l < L; l++) {
            if (live) {
                float *row = p.xh + (pos * (L - 1) + (l - 1)) * 2 * 64;
                store_feat(row, x, h);
                store_feat(row + 64, xd, h);
            }
#pragma unroll
            for (int ob = 0; ob < 2; ob++)
#pragma unroll
                for (int r = 0; r < 16; r++) { acc[ob][r] = vptr(vec, l, 0, ob, h)[r]; tac[ob][r] = 0.0f; }
            layer64<BF16>(whf + (l - 1) * lstride, lane, x, acc);
            layer64<BF16>(whf + (l - 1) * lstride, lane, xd, tac);
            float tv[2][16], zd[2][16];
#pragma unroll
            for (int ob = 0; ob < 2; ob++) {
                const float *rdt = vptr(vec, l, 1, ob, h);
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const float ta = act<BF16>(acc[ob][r]);
                    tv[ob][r] = ta;
                    zd[ob][r] = tac[ob][r];
                    x[ob][r] += rdt[r] * ta;
                    xd[ob][r] += rdt[r] * (1.0f - ta * ta) * tac[ob][r];
                }
            }
            if (live) {
                tz_put(tz_slot(p, pos, l, 0, h), tv);
                tz_put(tz_slot(p, pos, l, 1, h), zd);
            }
        }
        // ---- output: E = x . w_out + b_out; the rows of w_out's gradient, gE x - xdot
        const float *wout = vec + (size_t)L * 3 * 2 * 2 * 16;        // [2 blk][2 h][16]
        {
            float xo[2][16];
#pragma unroll
            for (int ob = 0; ob < 2; ob++)
#pragma unroll
                for (int r = 0; r < 16; r++) xo[ob][r] = ge * x[ob][r] - xd[ob][r];
            if (live) store_feat(p.xo + pos * 64, xo, h);
        }
        // ---- reverse pass of L = gE E - Edot
        float xb[2][16], xdb[2][16];
#pragma unroll
        for (int ob = 0; ob < 2; ob++)
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const float wo = wout[(ob * 2 + h) * 16 + r];
                xb[ob][r] = live ? ge * wo : 0.0f;
                xdb[ob][r] = live ? -wo : 0.0f;
            }
#pragma unroll
        for (int l = L - 1; l >= 1; l--) {
            float tv[2][16], zd[2][16], zb[2][16], zdb[2][16], rbv[2][16];
            tz_get(tz_slot(p, pos, l, 0, h), tv, live);
            tz_get(tz_slot(p, pos, l, 1, h), zd, live);
#pragma unroll
            for (int ob = 0; ob < 2; ob++) {
                const float *rdt = vptr(vec, l, 1, ob, h);
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const float q = 1.0f - tv[ob][r] * tv[ob][r];
                    const float tb = rdt[r] * (xb[ob][r] - 2.0f * tv[ob][r] * zd[ob][r] * xdb[ob][r]);
                    zdb[ob][r] = rdt[r] * q * xdb[ob][r];
                    zb[ob][r] = q * tb;
                    rbv[ob][r] = tv[ob][r] * xb[ob][r] + q * zd[ob][r] * xdb[ob][r];
                    acc[ob][r] = xb[ob][r];
                    tac[ob][r] = xdb[ob][r];
                }
            }
            if (live) {
                float *row = p.zh + (pos * (L - 1) + (l - 1)) * 2 * 64;
                store_feat(row, zb, h);
                store_feat(row + 64, zdb, h);
                store_feat(p.rb + (pos * (L - 1) + (l - 1)) * 64, rbv, h);
            }
            layer64<BF16>(whb + (l - 1) * lstride, lane, zb, acc);
            layer64<BF16>(whb + (l - 1) * lstride, lane, zdb, tac);
#pragma unroll
            for (int ob = 0; ob < 2; ob++)
#pragma unroll
                for (int r = 0; r < 16; r++) { xb[ob][r] = acc[ob][r]; xdb[ob][r] = tac[ob][r]; }
        }
        // layer 0 (x = t): tbar = xbar - 2 t zdot xdotbar
        float z0[2][16], zd0[2][16];
        {
            float tv[2][16], zd[2][16];
            tz_get(tz_slot(p, pos, 0, 0, h), tv, live);
            tz_get(tz_slot(p, pos, 0, 1, h), zd, live);
#pragma unroll
            for (int ob = 0; ob < 2; ob++)
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const float q = 1.0f - tv[ob][r] * tv[ob][r];
                    zd0[ob][r] = q * xdb[ob][r];
                    z0[ob][r] = q * (xb[ob][r] - 2.0f * tv[ob][r] * zd[ob][r] * xdb[ob][r]);
                }
            if (live) {
                store_feat(p.zw0 + pos * 2 * 64, z0, h);
                store_feat(p.zw0 + pos * 2 * 64 + 64, zd0, h);
            }
        }
        // ---- DR reverse, block by block: DRbar rows 32 blk + f(r,h) = W0 z0, DRdotbar = W0 zd0;
        //   k = p*16 + q: r < 8 -> p = 2 blk, q = f(r,h); r >= 8 -> p = 2 blk + 1, q = f(r-8,h).
        // Ahat[w][p] += sum_q G[p][q] S[w][q] and Ahat[w][q] += sum_p G[p][q] S[w][p] for (G, S) =
        // (DRbar, A) and (DRdotbar, At).
        float aq2[4][8], tq2[4][8];
#pragma unroll
        for (int w = 0; w < 4; w++)
#pragma unroll
            for (int r = 0; r < 8; r++) { aq2[w][r] = AS(n, w, feat(r, h)); tq2[w][r] = TS(n, w, feat(r, h)); }
        float d2[4][8];             // Ahat[w][q] of the q side, over all blocks
#pragma unroll
        for (int w = 0; w < 4; w++)
#pragma unroll
            for (int r = 0; r < 8; r++) d2[w][r] = 0.0f;
        for (int blk = 0; blk < p.nb0; blk++) {
            floatx16 g, gt;
#pragma unroll
            for (int r = 0; r < 16; r++) { g[r] = 0.0f; gt[r] = 0.0f; }
            if constexpr (BF16) {
                const bf16x8 *wb = reinterpret_cast<const bf16x8 *>(reinterpret_cast<const char *>(w0b) +
                                                                    ((size_t)blk * 64 + lane) * 32 * 2);
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    bf16x8 b, bt;
#pragma unroll
                    for (int q = 0; q < 8; q++) {
                        b[q] = (__bf16)z0[(c * 8) / 16][(c * 8) % 16 + q];
                        bt[q] = (__bf16)zd0[(c * 8) / 16][(c * 8) % 16 + q];
                    }
                    const bf16x8 a = wb[c];
                    g = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, g, 0, 0, 0);
                    gt = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bt, gt, 0, 0, 0);
                }
            } else {
                const char *wb = reinterpret_cast<const char *>(w0b) + ((size_t)blk * 64 + lane) * 32 * 4;
#pragma unroll
                for (int c = 0; c < 32; c++) {
                    mma<false>(g, wb, c, &z0[c / 16][c % 16]);
                    mma<false>(gt, wb, c, &zd0[c / 16][c % 16]);
                }
            }
            const int p0 = 2 * blk, p1 = 2 * blk + 1;
            float s0[4], s1[4];
#pragma unroll
            for (int w = 0; w < 4; w++) {
                const float ap0 = p0 < E ? AS(n, w, p0) : 0.f, ap1 = p1 < E ? AS(n, w, p1) : 0.f;
                const float tp0 = p0 < E ? TS(n, w, p0) : 0.f, tp1 = p1 < E ? TS(n, w, p1) : 0.f;
                float a0 = 0.f, a1 = 0.f;
#pragma unroll
                for (int r = 0; r < 8; r++) {
                    a0 += g[r] * aq2[w][r] + gt[r] * tq2[w][r];
                    a1 += g[r + 8] * aq2[w][r] + gt[r + 8] * tq2[w][r];
                    d2[w][r] += g[r] * ap0 + g[r + 8] * ap1 + gt[r] * tp0 + gt[r + 8] * tp1;
                }
                s0[w] = a0 + other_half(a0);
                s1[w] = a1 + other_half(a1);
            }
            // rows p0 (written by half 0) and p1 (half 1): one writer per entry, no other lane adds
            // to them in this block; the q-side terms are added after the loop
            if (h == 0 && p0 < E) {
#pragma unroll
                for (int w = 0; w < 4; w++) HS(n, w, p0) += s0[w];
            }
            if (h == 1 && p1 < E) {
#pragma unroll
                for (int w = 0; w < 4; w++) HS(n, w, p1) += s1[w];
            }
        }
        __syncthreads();
        // q-side terms (q < 16): columns f(r, h), disjoint between the halves
#pragma unroll
        for (int r = 0; r < 8; r++) {
            const int q = feat(r, h);
#pragma unroll
            for (int w = 0; w < 4; w++) HS(n, w, q) += d2[w][r];
        }
        __syncthreads();
        for (int x2 = lane; x2 < 32 * 4 * E; x2 += WAVE) {
            const int an = x2 / (4 * E), wc = x2 - an * 4 * E;
            if (an < cnt) p.Ahat[(int64_t)p.perm[s0 + an] * 4 * E + wc] = Hs[an * RS + wc];
        }
        __syncthreads();
    }
#undef AS
#undef TS
#undef HS
}

}  // namespace
}  // namespace matpl

using namespace matpl;

MATPL_EXPORT int matpl_dp_fit_train(const float *A, const float *At, const float *gE, const int32_t *perm,
                                    const int32_t *tiles, int ntiles, const void *w0f, const void *w0b,
                                    const void *whf, const void *whb, const float *vec, int64_t type_stride_w0f,
                                    int64_t type_stride_w0b, int64_t type_stride_wh, int64_t type_stride_vec, int E,
                                    int nlayers, int bf16, float *Ahat, float *xw0, float *zw0, float *xh, float *zh,
                                    float *rb, float *xo, float *tz, void *stream) {
    MATPL_CHECK_ARG(E >= M2 && E <= E_MAX, "dp_fit_train: E must be in [%d,%d]", M2, E_MAX);
    MATPL_CHECK_ARG(nlayers >= 1 && nlayers <= 4, "dp_fit_train: hidden layers must be in [1,4]");
    if (ntiles <= 0) return OK;
    MATPL_CHECK_ARG(A && At && perm && tiles && w0f && w0b && vec && Ahat && xw0 && zw0 && xo && tz &&
                        (nlayers == 1 || (whf && whb && xh && zh && rb)),
                    "dp_fit_train: null pointer");
    MATPL_CHECK_ARG(aligned16(xw0) && aligned16(zw0) && aligned16(xo) && aligned16(tz) &&
                        (nlayers == 1 || (aligned16(xh) && aligned16(zh) && aligned16(rb))),
                    "dp_fit_train: row buffers must be 16-byte aligned");
    FitTrainParams p{};
    p.A = A; p.At = At; p.gE = gE; p.perm = perm; p.tiles = tiles;
    p.w0f = w0f; p.w0b = w0b; p.whf = whf; p.whb = whb; p.vec = vec;
    p.type_stride_w0f = type_stride_w0f; p.type_stride_w0b = type_stride_w0b;
    p.type_stride_wh = type_stride_wh; p.type_stride_vec = type_stride_vec;
    p.Ahat = Ahat; p.xw0 = xw0; p.zw0 = zw0; p.xh = xh; p.zh = zh; p.rb = rb; p.xo = xo; p.tz = tz;
    p.ntiles = ntiles; p.E = E; p.D = E * M2; p.Dh = (E * M2) / 2; p.nb0 = (E * M2 + 31) / 32; p.L = nlayers;
    const int grid = ntiles < 8192 ? ntiles : 8192;
    const size_t lds = sizeof(float) * 3 * 32 * (4 * E + 1);
    hipStream_t st = (hipStream_t)stream;
#define MATPL_FIT_TRAIN_CASE(LL)                                                                         \
    case LL:                                                                                             \
        if (bf16) hipLaunchKernelGGL((dp_fit_train_kernel<true, LL>), dim3(grid), dim3(WAVE), lds, st, p); \
        else hipLaunchKernelGGL((dp_fit_train_kernel<false, LL>), dim3(grid), dim3(WAVE), lds, st, p);   \
        break;
    switch (nlayers) {
        MATPL_FIT_TRAIN_CASE(1)
        MATPL_FIT_TRAIN_CASE(2)
        MATPL_FIT_TRAIN_CASE(3)
        MATPL_FIT_TRAIN_CASE(4)
        default:
            set_error("dp_fit_train: %d hidden layers not compiled (1..4)", nlayers);
            return ERR_ARGUMENT;
    }
#undef MATPL_FIT_TRAIN_CASE
    return check_launch("dp_fit_train_kernel");
}
