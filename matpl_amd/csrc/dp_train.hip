// Hand-written force-loss double backward of the DP se_e2_a embedding nets
// for the batched training step (the reference differentiates
// dE = autograd.grad(Etot, Ri, create_graph=True), src/model/dp_dp.py:429,
// and loss.backward() through it, src/PWMLFF/dp_mods/dp_trainer.py:274).
//
// With A_i = c_i sum_j Ri_ij (x) G(s_ij) (c_i = 1 / (m * type_nums of i's
// frame)), Abar_i = dE/dA_i from the fitting net and the force residual
// r_k = dL/dF_k, the force part of the loss gradient is -dQ/dtheta with
// Q = sum_i <Abar_i, dA_i> and dA_i = c_i sum_j (v_ij (x) G_ij + Ri_ij (x)
// v_ij0 G'_ij): the tangent of A_i along the displacement field r, where
// v_ij = dRi_ij/d(dR_ij) (r_j - r_i) and G' = dG/ds. So:
//
//   dp_train_tangent_kernel (one wavefront per atom, like the fused forward):
//       recomputes Ri, dRi/d(dR), G and G' (forward-mode through the net on
//       the matrix cores) and forms dA [N][4][E].
//   The caller differentiates the fitting net twice in torch:
//       Ahat = d(L_E - <Abar(A), dA>)/dA and the fitting-net gradients.
//   dp_train_embgrad_kernel (one wavefront per (bucket t, chunk of atoms)):
//       for every in-cutoff neighbor the embedding-net parameter gradient of
//       a . G + b . G' with a = c (Ahat^T Ri - Abar^T v) and
//       b = -c v_0 (Abar^T Ri): the net's forward (primal and tangent) is
//       recomputed, back-propagated through both on the matrix cores, and
//       the weight gradients sum_n x_n (x) zbar_n are themselves matrix-core
//       products over the tile's neighbors. Those need the activations with
//       the feature index on the lanes; the transposes are matrix-core
//       products with the identity (no LDS). Each wavefront keeps the
//       gradient of one type pair in registers while its atoms keep that
//       pair and adds it to the global accumulators when the pair changes.
//       Consecutive atoms of one frame and type whose bucket-t entries fit
//       one 64-lane pass are compacted together and share tiles (a bucket
//       holds ~10 in-cutoff neighbors in a 4-type system: one atom per tile
//       would leave two thirds of the lanes idle); a and b are then formed
//       per lane from its own atom's Ahat and Abar rows.
//
// The empty-slot term of A (constant in positions) is differentiated by the
// caller. The kernels assume the reference's default embedding: no
// resnet_dt, E < 32 (row E of the layer-2/3 input operands carries the bias
// gradient).
#include "dp_fused_common.h"

namespace matpl {
namespace {

// Atoms whose bucket entries the gradient kernel takes together (one compaction pass, shared tiles):
// 8 in BF16; 4 in FP32, whose transposes take a per-wavefront LDS tile instead.
template <bool BF16>
constexpr int eg_atoms() { return BF16 ? 8 : 4; }
constexpr int EG_TSTRIDE = 33;                        // FP32 transpose tile [32][33] floats
constexpr int EG_TFLOAT4 = 32 * EG_TSTRIDE / 4;       // its size in float4

// Per-pair gradient accumulator layout (floats): W3 [E][E], b3 [E], W2 [E][E],
// b2 [E], w1 [E], b1 [E]; W[in][out] as the net applies x @ W.
template <int E>
struct GradLayout {
    static constexpr int W3 = 0, B3 = E * E, W2 = B3 + E, B2 = W2 + E * E, W1 = B2 + E, B1 = W1 + E;
    static constexpr int SIZE = B1 + E;
};

struct TrainParams {
    FusedParams f;
    const float *resid;   // [N][3] force residual r = dL/dF
    const float *abar;    // [N][4][E] dE/dA (not scaled by c)
    const float *ahat;    // [N][4][E] dL/dA from the fitting side (not scaled by c)
    float *dA;            // [N][4][E] tangent output (scaled by c)
    int32_t *nin;         // [N][T] in-cutoff neighbors per bucket (tangent kernel output, optional)
    float *grad;          // [T*T][GradLayout::SIZE], accumulated
    int chunk;            // atoms per wavefront (embgrad kernel)
};

// Tangent of the normalized Ri along u = r_j - r_i (v = dRi/d(dR) u), and the
// values the backward needs, for a compacted neighbor e = (dx, dy, dz, slot).
__device__ __forceinline__ void ri_tangent(const float4 e, const float *ti, float r, float s, float ds, const float u[3],
                                           float v[4]) {
    const float ir = __builtin_amdgcn_rcpf(r);
    const float ux = e.x * u[0] + e.y * u[1] + e.z * u[2];
    const float c1 = (ds - s * ir) * ux * ir * ir;       // d(s/r)/dr (x . u) / r
    v[0] = ds * ux * ir * ti[TI_ISTD];
    v[1] = (e.x * c1 + s * u[0] * ir) * ti[TI_ISTD + 1];
    v[2] = (e.y * c1 + s * u[1] * ir) * ti[TI_ISTD + 2];
    v[3] = (e.z * c1 + s * u[2] * ir) * ti[TI_ISTD + 3];
}

// Forward of the embedding net with its tangent in s (the reference's
// EmbeddingNet, dp_embedding.py:73-94, without resnet_dt). Per lane (its
// neighbor, its half's features): h1 = tanh(z1), ḣ1 = (1-h1^2) w1,
// t2 = tanh(z2), ż2 = W2^T ḣ1, t3, ż3 likewise; G = h1 + t2 + t3 and
// G' = ḣ1 + (1-t2^2) ż2 + (1-t3^2) ż3. Pre-activations are true (unscaled)
// values: in BF16 mode the stored forward weights carry TANH_K.
template <int E, bool BF16>
struct EmbTan {
    float h1[16], hd1[16], t2[16], zd2[16], t3[16], zd3[16];

    __device__ __forceinline__ void run(const float *vecb, const unsigned char *matb, int pair, int lane, float s) {
        constexpr int NR = NReg<E>::value;
        constexpr float IK = BF16 ? 1.0f / TANH_K : 1.0f;
        const int h = lane >> 5;
        const float *w1 = vec_ptr(vecb, pair, V_W1, h), *b1 = vec_ptr(vecb, pair, V_B1, h);
        const float *b2 = vec_ptr(vecb, pair, V_B2, h), *b3 = vec_ptr(vecb, pair, V_B3, h);
#pragma unroll
        for (int r = 0; r < 16; r++) {
            h1[r] = (r < NR) ? act<BF16>(s * w1[r] + b1[r]) : 0.0f;
            hd1[r] = (r < NR) ? (1.0f - h1[r] * h1[r]) * (w1[r] * IK) : 0.0f;
        }
        floatx16 acc, tac;
#pragma unroll
        for (int r = 0; r < 16; r++) { acc[r] = b2[r]; tac[r] = 0.0f; }
        mfma_layer<BF16, false>(mat_ptr<BF16>(matb, pair, M_W2F), lane, h1, acc);
        mfma_layer<BF16, false>(mat_ptr<BF16>(matb, pair, M_W2F), lane, hd1, tac);
        float h2[16], hd2[16];
#pragma unroll
        for (int r = 0; r < 16; r++) {
            t2[r] = (r < NR) ? act<BF16>(acc[r]) : 0.0f;
            zd2[r] = (r < NR) ? tac[r] * IK : 0.0f;
            h2[r] = h1[r] + t2[r];
            hd2[r] = hd1[r] + (1.0f - t2[r] * t2[r]) * zd2[r];
        }
#pragma unroll
        for (int r = 0; r < 16; r++) { acc[r] = b3[r]; tac[r] = 0.0f; }
        mfma_layer<BF16, false>(mat_ptr<BF16>(matb, pair, M_W3F), lane, h2, acc);
        mfma_layer<BF16, false>(mat_ptr<BF16>(matb, pair, M_W3F), lane, hd2, tac);
#pragma unroll
        for (int r = 0; r < 16; r++) {
            t3[r] = (r < NR) ? act<BF16>(acc[r]) : 0.0f;
            zd3[r] = (r < NR) ? tac[r] * IK : 0.0f;
        }
    }
};

// ---------------------------------------------------------------- tangent of A
template <int E, bool BF16>
__global__ void __launch_bounds__(FUSED_WAVES * WAVE)
dp_train_tangent_kernel(TrainParams tp) {
    constexpr int NR = NReg<E>::value;
    const FusedParams &p = tp.f;
    extern __shared__ float4 smem[];
    const int lane = threadIdx.x & (WAVE - 1), wid = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const BlockLds L = block_lds<BF16, false, false>(p, smem, wid);
    float4 *nb = L.nb;
    int *nvs = L.nvs;
    const int h = lane >> 5, n = lane & 31;
    for (int64_t i = logical_block(p.xcd) * FUSED_WAVES + wid; i < p.natoms; i += (int64_t)gridDim.x * FUSED_WAVES) {
        const int ti = p.type_map[i];
        const float *tinfo = p.tinfo + ti * NTI;
        const AtomGeo g = atom_geo(p, i);
        const float ri[3] = {tp.resid[3 * i], tp.resid[3 * i + 1], tp.resid[3 * i + 2]};
        if (!compact(p, g, i, lane, nb, nvs, nullptr) && lane == 0) atomicCAS(p.status, 0, (int)(i + 1));
        __builtin_amdgcn_wave_barrier();
        if (tp.nin && lane < p.T) tp.nin[i * p.T + lane] = nvs[lane];
        float acc[4][16];
#pragma unroll
        for (int w = 0; w < 4; w++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[w][r] = 0.0f;
        for (int t = 0; t < p.T; t++) {
            const int nv = nvs[t];
            const int pair = ti * p.T + t;
            for (int k0 = 0; k0 < nv; k0 += TILE) {
                const int k = k0 + n;
                const bool live = k < nv;
                const float4 e = live ? nb[t * p.cap + k] : make_float4(1.f, 0.f, 0.f, 0.f);
                float Ri[4] = {0.f, 0.f, 0.f, 0.f}, v[4] = {0.f, 0.f, 0.f, 0.f};
                float r = 1.f, s = 0.f, ds = 0.f;
                if (live) {
                    ri_of(e, tinfo, p.rmin, p.rmax, p.iw, Ri, r, s, ds);
                    const int64_t slot = (i * p.T + t) * (int64_t)p.mnl + __float_as_int(e.w);
                    const int64_t j = g.fbase + p.list[slot] - 1;
                    const float u[3] = {tp.resid[3 * j] - ri[0], tp.resid[3 * j + 1] - ri[1],
                                        tp.resid[3 * j + 2] - ri[2]};
                    ri_tangent(e, tinfo, r, s, ds, u, v);
                }
                int pr = pair;
                asm volatile("" : "+v"(pr));
                EmbTan<E, BF16> et;
                et.run(p.vec, reinterpret_cast<const unsigned char *>(p.mat), pr, lane, Ri[0]);
                // dA += v (x) G + Ri (x) v0 G' (dead lanes: v = Ri = 0)
#pragma unroll
                for (int rr = 0; rr < NR; rr++) {
                    const float G = et.h1[rr] + et.t2[rr] + et.t3[rr];
                    const float Gd = et.hd1[rr] + (1.0f - et.t2[rr] * et.t2[rr]) * et.zd2[rr] +
                                     (1.0f - et.t3[rr] * et.t3[rr]) * et.zd3[rr];
                    const float vg = v[0] * Gd;
#pragma unroll
                    for (int w = 0; w < 4; w++) acc[w][rr] += v[w] * G + Ri[w] * vg;
                }
            }
        }
        // sum over the neighbors (lanes of each half); lane n of half h writes register n's feature
        float out[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int rr = 0; rr < NR; rr++) {
#pragma unroll
            for (int w = 0; w < 4; w++) {
                const float x = half_sum(acc[w][rr]);
                if (n == rr) out[w] = x;
            }
        }
        if (n < NR && feat(n, h) < E) {
            float scale;
            frame_type_mask(p, i, scale);
            float *o = tp.dA + (size_t)i * 4 * E;
#pragma unroll
            for (int w = 0; w < 4; w++) o[w * E + feat(n, h)] = out[w] * scale;
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// ---------------------------------------------------------------- embedding gradients
// Matrix-core transpose: X in the accumulator layout (lane = neighbor n,
// register r = feature f(r, h)) -> X^T with the neighbor index on the
// registers (lane = feature l % 32, register r = neighbor f(r, l / 32)),
// as X^T I with X as the A operand (rows = neighbors, K = features in the
// register order) and the identity as the B operand.
template <bool BF16>
struct Transposer {
    bf16x8 id0, id1;      // BF16: identity B operand, K-chunks 0 and 1 (registers 0-7, 8-15)
    float *scr;           // FP32: this wavefront's [32][33] LDS tile
    int h, n;

    __device__ __forceinline__ void init(int lane, float *scratch) {
        h = lane >> 5;
        n = lane & 31;
        scr = scratch;
        if constexpr (BF16) {
#pragma unroll
            for (int q = 0; q < 8; q++) {
                id0[q] = (__bf16)(feat(q, h) == n ? 1.0f : 0.0f);
                id1[q] = (__bf16)(feat(8 + q, h) == n ? 1.0f : 0.0f);
            }
        }
    }

    __device__ __forceinline__ floatx16 apply(const float x[16]) const {
        floatx16 d;
        if constexpr (BF16) {
#pragma unroll
            for (int r = 0; r < 16; r++) d[r] = 0.0f;
            bf16x8 a0, a1;
#pragma unroll
            for (int q = 0; q < 8; q++) { a0[q] = (__bf16)x[q]; a1[q] = (__bf16)x[8 + q]; }
            d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, id0, d, 0, 0, 0);
            d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, id1, d, 0, 0, 0);
        } else {
            // FP32: through LDS (sixteen K = 2 matrix-core products against the identity cost more
            // than the tile's outer products): lane (n, h) writes row feat(r, h), column n; lane
            // (f, h) reads row f, columns feat(r, h). The odd stride keeps both phases conflict-free.
#pragma unroll
            for (int r = 0; r < 16; r++) scr[feat(r, h) * EG_TSTRIDE + n] = x[r];
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int r = 0; r < 16; r++) d[r] = scr[n * EG_TSTRIDE + feat(r, h)];
            __builtin_amdgcn_wave_barrier();
        }
        return d;
    }
};

// acc[in][out] += sum over the tile's neighbors of X[in][n] Y[out][n], X and Y
// transposed (neighbor index on the registers).
template <bool BF16>
__device__ __forceinline__ void outer_acc(const floatx16 &xt, const floatx16 &yt, floatx16 &acc) {
    if constexpr (BF16) {
        bf16x8 a0, a1, b0, b1;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            a0[q] = (__bf16)xt[q]; a1[q] = (__bf16)xt[8 + q];
            b0[q] = (__bf16)yt[q]; b1[q] = (__bf16)yt[8 + q];
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b0, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1, acc, 0, 0, 0);
    } else {
#pragma unroll
        for (int r = 0; r < 16; r++) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(xt[r], yt[r], acc, 0, 0, 0);
    }
}

// Adds one wavefront's gradient of pair `pair` to the global accumulators and clears it.
template <int E>
__device__ __forceinline__ void flush_pair(float *grad, int pair, int lane, floatx16 &g3, floatx16 &g2, float w1s[16],
                                           float b1s[16]) {
    using GL = GradLayout<E>;
    constexpr int NR = NReg<E>::value;
    float *gp = grad + (size_t)pair * GL::SIZE;
    const int h = lane >> 5, out = lane & 31;
    // g3, g2: lane holds [in = f(r, h)][out]; row E is the bias gradient
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int in = feat(r, h);
        if (out < E && in <= E) {
            if (g3[r] != 0.0f) atomic_add(gp + (in < E ? GL::W3 + in * E + out : GL::B3 + out), g3[r]);
            if (g2[r] != 0.0f) atomic_add(gp + (in < E ? GL::W2 + in * E + out : GL::B2 + out), g2[r]);
        }
        g3[r] = 0.0f;
        g2[r] = 0.0f;
    }
    // w1, b1: per-lane sums over this lane's neighbors -> sums over the half; lane n writes register n
    float ow = 0.0f, ob = 0.0f;
#pragma unroll
    for (int r = 0; r < NR; r++) {
        const float xw = half_sum(w1s[r]), xb = half_sum(b1s[r]);
        if (out == r) { ow = xw; ob = xb; }
        w1s[r] = 0.0f;
        b1s[r] = 0.0f;
    }
    if (out < NR && feat(out, h) < E) {
        if (ow != 0.0f) atomic_add(gp + GL::W1 + feat(out, h), ow);
        if (ob != 0.0f) atomic_add(gp + GL::B1 + feat(out, h), ob);
    }
}

template <int E, bool BF16>
__global__ void __launch_bounds__(FUSED_WAVES * WAVE)
dp_train_embgrad_kernel(TrainParams tp) {
    constexpr int NR = NReg<E>::value;
    static_assert(E < 32, "row E of the layer inputs carries the bias gradient");
    constexpr float IK = BF16 ? 1.0f / TANH_K : 1.0f;
    const FusedParams &p = tp.f;
    extern __shared__ float4 smem[];
    const int lane = threadIdx.x & (WAVE - 1), wid = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const int h = lane >> 5, n = lane & 31;
    // per wave: the batch's compacted neighbors [NBW] (dx, dy, dz, atom of the batch), their
    // residual differences r_j - r_i [NBW], Ahat and Abar of the batch's atoms [EGA][32 features][4]
    constexpr int EGA = eg_atoms<BF16>();
    const int NBW = p.cap > WAVE ? p.cap : WAVE;
    float4 *nb = smem + (size_t)wid * (2 * NBW + 2 * EGA * 32 + (BF16 ? 0 : EG_TFLOAT4));
    float4 *us = nb + NBW;
    float4 *ahs = us + NBW, *abs = ahs + EGA * 32;
    float *tscr = reinterpret_cast<float *>(abs + EGA * 32);     // FP32 transpose tile
    const int64_t wg = (int64_t)blockIdx.x * FUSED_WAVES + wid;
    const int t = (int)(wg % p.T);
    const int64_t i0 = (wg / p.T) * tp.chunk;
    int64_t i1 = i0 + tp.chunk;
    if (i1 > p.natoms) i1 = p.natoms;
    Transposer<BF16> tr;
    tr.init(lane, tscr);
    floatx16 g3, g2;
    float w1s[16], b1s[16];
#pragma unroll
    for (int r = 0; r < 16; r++) { g3[r] = 0.0f; g2[r] = 0.0f; w1s[r] = 0.0f; b1s[r] = 0.0f; }
    int cur = -1;       // pair held in the accumulators
    const uint64_t below = (1ull << lane) - 1ull;
    for (int64_t i = i0; i < i1;) {
        // a batch: atoms i .. i + na - 1 of one frame and one type, at most EGA of them, whose
        // bucket-t entries fit one pass of the wavefront together (or a single atom with more), so
        // that short buckets share tiles
        const int64_t ia_l = i + (lane < EGA ? lane : 0);
        const bool in_l = lane < EGA && ia_l < i1;
        const int c_l = in_l ? p.counts[ia_l * p.T + t] : 0;
        const int ty_l = in_l ? p.type_map[ia_l] : -1;
        const int fr_l = in_l ? (p.atom_frame ? p.atom_frame[ia_l] : 0) : -1;
        const int ti = __builtin_amdgcn_readfirstlane(ty_l);
        const int fr = __builtin_amdgcn_readfirstlane(fr_l);
        int off[EGA + 1];
        off[0] = 0;
        int na = 0;
        bool open = true;
#pragma unroll
        for (int q = 0; q < EGA; q++) {
            const int cq = __builtin_amdgcn_readlane(c_l, q);
            const int tq = __builtin_amdgcn_readlane(ty_l, q);
            const int fq = __builtin_amdgcn_readlane(fr_l, q);
            open = open && (q == 0 || (tq == ti && fq == fr && off[q] + cq <= WAVE));
            if (open) na = q + 1;
            off[q + 1] = off[q] + (open ? cq : 0);
        }
        const int tot = off[EGA];
        const int pair = ti * p.T + t;
        if (pair != cur) {
            if (cur >= 0) flush_pair<E>(tp.grad, cur, lane, g3, g2, w1s, b1s);
            cur = pair;
        }
        if (tot == 0) {
            i += na;
            continue;
        }
        float scale;
        frame_type_mask(p, i, scale);
        const float *tinfo = p.tinfo + ti * NTI;
        const AtomGeo g0 = atom_geo(p, i);
        // the batch's Ahat and Abar, scaled by c, feature-major
        {
            float *ahf = reinterpret_cast<float *>(ahs), *abf = reinterpret_cast<float *>(abs);
            for (int x = lane; x < na * 128; x += WAVE) {
                const int q = x >> 7, f = (x & 127) >> 2, w = x & 3;
                const size_t src = (size_t)(i + q) * 4 * E + w * E + f;
                ahf[x] = f < E ? tp.ahat[src] * scale : 0.0f;
                abf[x] = f < E ? tp.abar[src] * scale : 0.0f;
            }
        }
        // in-cutoff neighbors of bucket t, atom by atom in order, with r_j - r_i (loaded with the
        // neighbor's position, so that the tiles below read nothing from global memory)
        int nv = 0;
        for (int e0 = 0; e0 < tot; e0 += WAVE) {
            const int e = e0 + lane;
            int q = 0;
#pragma unroll
            for (int qq = 1; qq < EGA; qq++) q += (e >= off[qq]) ? 1 : 0;
            double dx = 0.0, dy = 0.0, dz = 0.0;
            float u0 = 0.0f, u1 = 0.0f, u2 = 0.0f;
            if (e < tot) {
                const int64_t ia = i + q;
                AtomGeo gl = g0;
                gl.xi = p.pos[3 * ia]; gl.yi = p.pos[3 * ia + 1]; gl.zi = p.pos[3 * ia + 2];
                const int64_t slot = (ia * p.T + t) * (int64_t)p.mnl + (e - off[q]);
                const int64_t j = neighbor_disp(p, gl, slot, dx, dy, dz);
                u0 = tp.resid[3 * j] - tp.resid[3 * ia];
                u1 = tp.resid[3 * j + 1] - tp.resid[3 * ia + 1];
                u2 = tp.resid[3 * j + 2] - tp.resid[3 * ia + 2];
            }
            const double r2 = dx * dx + dy * dy + dz * dz;
            const bool ok = e < tot && r2 < p.rc2 && r2 > 1e-10;
            const uint64_t bal = __ballot(ok);
            const int idx = nv + __popcll(bal & below);
            if (ok && idx < NBW) {
                nb[idx] = make_float4((float)dx, (float)dy, (float)dz, __int_as_float(q));
                us[idx] = make_float4(u0, u1, u2, 0.0f);
            }
            nv += __popcll(bal);
        }
        if (nv > NBW) nv = NBW;
        __builtin_amdgcn_wave_barrier();
        for (int k0 = 0; k0 < nv; k0 += TILE) {
            const int k = k0 + n;
            const bool live = k < nv;
            const float4 e = live ? nb[k] : make_float4(1.f, 0.f, 0.f, 0.f);
            float Ri[4] = {0.f, 0.f, 0.f, 0.f}, v[4] = {0.f, 0.f, 0.f, 0.f};
            float r = 1.f, s = 0.f, ds = 0.f;
            int qa = 0;     // the neighbor's atom in the batch
            if (live) {
                ri_of(e, tinfo, p.rmin, p.rmax, p.iw, Ri, r, s, ds);
                const float4 uu = us[k];
                const float u[3] = {uu.x, uu.y, uu.z};
                ri_tangent(e, tinfo, r, s, ds, u, v);
                qa = __float_as_int(e.w);
            }
            const int pr = pair;
            EmbTan<E, BF16> et;
            et.run(p.vec, reinterpret_cast<const unsigned char *>(p.mat), pr, lane, Ri[0]);
            // upstream of G: a = Ahat Ri - Abar v; of G': b = -v0 Abar Ri, with the rows of the lane's
            // own atom of the batch (FP32)
            floatx16 a, b;
            {
                const float4 *ahq = ahs + qa * 32, *abq = abs + qa * 32;
#pragma unroll
                for (int rr = 0; rr < 16; rr++) {
                    a[rr] = 0.0f;
                    b[rr] = 0.0f;
                    if (rr < NR) {
                        const int f = feat(rr, h);
                        const float4 A = ahq[f], B = abq[f];
                        const float bri = B.x * Ri[0] + B.y * Ri[1] + B.z * Ri[2] + B.w * Ri[3];
                        a[rr] = (A.x * Ri[0] + A.y * Ri[1] + A.z * Ri[2] + A.w * Ri[3]) -
                                (B.x * v[0] + B.y * v[1] + B.z * v[2] + B.w * v[3]);
                        b[rr] = -v[0] * bri;
                    }
                }
            }
            // layer 3 (G = h2 + t3, G' = ḣ2 + (1 - t3^2) ż3): adjoints of z3 and ż3
            float zb[16], zbd[16], x[16], xd[16];
#pragma unroll
            for (int rr = 0; rr < 16; rr++) {
                const float q = 1.0f - et.t3[rr] * et.t3[rr];
                zbd[rr] = rr < NR ? q * b[rr] : 0.0f;
                zb[rr] = rr < NR ? q * (a[rr] - 2.0f * et.t3[rr] * et.zd3[rr] * b[rr]) : 0.0f;
                // layer-3 inputs h2, ḣ2; row E (register 13, half 0 for E = 25) is 1: the bias row
                const float q2 = 1.0f - et.t2[rr] * et.t2[rr];
                x[rr] = rr < NR ? et.h1[rr] + et.t2[rr] : 0.0f;
                xd[rr] = rr < NR ? et.hd1[rr] + q2 * et.zd2[rr] : 0.0f;
                if (feat(rr, h) == E) x[rr] = 1.0f;
            }
            {
                const floatx16 xt = tr.apply(x), yt = tr.apply(zb), xdt = tr.apply(xd), ydt = tr.apply(zbd);
                outer_acc<BF16>(xt, yt, g3);
                outer_acc<BF16>(xdt, ydt, g3);
            }
            // h̄2 = a + W3 z̄3, ḣ̄2 = b + W3 ż̄3 (W3B: true weights)
            mfma_layer<BF16, false>(mat_ptr<BF16>(reinterpret_cast<const unsigned char *>(p.mat), pr, M_W3B), lane,
                                    zb, a);
            mfma_layer<BF16, false>(mat_ptr<BF16>(reinterpret_cast<const unsigned char *>(p.mat), pr, M_W3B), lane,
                                    zbd, b);
            // layer 2 (h2 = h1 + t2, ḣ2 = ḣ1 + (1 - t2^2) ż2)
#pragma unroll
            for (int rr = 0; rr < 16; rr++) {
                const float q = 1.0f - et.t2[rr] * et.t2[rr];
                zbd[rr] = rr < NR ? q * b[rr] : 0.0f;
                zb[rr] = rr < NR ? q * (a[rr] - 2.0f * et.t2[rr] * et.zd2[rr] * b[rr]) : 0.0f;
                x[rr] = rr < NR ? et.h1[rr] : 0.0f;
                xd[rr] = rr < NR ? et.hd1[rr] : 0.0f;
                if (feat(rr, h) == E) x[rr] = 1.0f;
            }
            {
                const floatx16 xt = tr.apply(x), yt = tr.apply(zb), xdt = tr.apply(xd), ydt = tr.apply(zbd);
                outer_acc<BF16>(xt, yt, g2);
                outer_acc<BF16>(xdt, ydt, g2);
            }
            mfma_layer<BF16, false>(mat_ptr<BF16>(reinterpret_cast<const unsigned char *>(p.mat), pr, M_W2B), lane,
                                    zb, a);
            mfma_layer<BF16, false>(mat_ptr<BF16>(reinterpret_cast<const unsigned char *>(p.mat), pr, M_W2B), lane,
                                    zbd, b);
            // layer 1 (h1 = tanh(x w1 + b1) with x = Ri[0] the normalized input, ḣ1 = (1 - h1^2) w1):
            // w̄1 += x z̄1 + ż̄1, b̄1 += z̄1
            const float *w1 = vec_ptr(p.vec, pr, V_W1, h);
#pragma unroll
            for (int rr = 0; rr < NR; rr++) {
                const float q = 1.0f - et.h1[rr] * et.h1[rr];
                const float zd1 = q * b[rr];
                const float z1 = q * (a[rr] - 2.0f * et.h1[rr] * (w1[rr] * IK) * b[rr]);
                w1s[rr] += Ri[0] * z1 + zd1;
                b1s[rr] += z1;
            }
        }
        __builtin_amdgcn_wave_barrier();
        i += na;
    }
    if (cur >= 0) flush_pair<E>(tp.grad, cur, lane, g3, g2, w1s, b1s);
}

template <int E, bool BF16>
int launch_train(TrainParams &tp, int which, hipStream_t st) {
    FusedParams &p = tp.f;
    if (which == 0) {
        const size_t lds = fused_lds_bytes<BF16, false, false>(p.T, p.cap);
        MATPL_CHECK_ARG(lds <= 64 * 1024, "dp_train_tangent: T*mnl too large for LDS (%zu bytes)", lds);
        int64_t blocks = (p.natoms + FUSED_WAVES - 1) / FUSED_WAVES;
        if (blocks > 16384) blocks = 16384;
        p.xcd = 0;
        hipLaunchKernelGGL((dp_train_tangent_kernel<E, BF16>), dim3((unsigned)blocks), dim3(FUSED_WAVES * WAVE), lds, st,
                           tp);
        return check_launch("dp_train_tangent_kernel");
    }
    const size_t nbw = p.cap > WAVE ? p.cap : WAVE;
    const size_t lds = sizeof(float4) * FUSED_WAVES * (2 * nbw + 2 * eg_atoms<BF16>() * 32 + (BF16 ? 0 : EG_TFLOAT4));
    MATPL_CHECK_ARG(lds <= 64 * 1024, "dp_train_embgrad: list width too large for LDS (%zu bytes)", lds);
    const int64_t chunks = (p.natoms + tp.chunk - 1) / tp.chunk;
    const int64_t waves = chunks * p.T;
    const int64_t blocks = (waves + FUSED_WAVES - 1) / FUSED_WAVES;
    MATPL_CHECK_ARG(blocks <= (1 << 30), "dp_train_embgrad: too many chunks");
    // waves beyond chunks * T find an empty atom range
    hipLaunchKernelGGL((dp_train_embgrad_kernel<E, BF16>), dim3((unsigned)blocks), dim3(FUSED_WAVES * WAVE), lds, st, tp);
    return check_launch("dp_train_embgrad_kernel");
}

int train_dispatch(TrainParams &tp, int E, int bf16, int which, hipStream_t st) {
    FusedParams &p = tp.f;
    MATPL_CHECK_ARG(p.T >= 1 && p.T <= MAXT, "dp_train: ntypes must be in [1,%d]", MAXT);
    MATPL_CHECK_ARG(p.mnl >= 1 && p.m >= 1, "dp_train: bad bucket sizes");
    p.cap = p.m < p.mnl ? p.m : p.mnl;
    if (p.natoms == 0) return OK;
    MATPL_CHECK_ARG(p.pos && p.lattice && p.list && p.shift && p.counts && p.type_map && p.vec && p.mat && p.tinfo &&
                        p.status && tp.resid,
                    "dp_train: null pointer");
    if (E == 25) return bf16 ? launch_train<25, true>(tp, which, st) : launch_train<25, false>(tp, which, st);
    if (E == 16) return bf16 ? launch_train<16, true>(tp, which, st) : launch_train<16, false>(tp, which, st);
    set_error("dp_train: embedding width %d not compiled (16, 25)", E);
    return ERR_ARGUMENT;
}

void fill_common(FusedParams &p, const double *pos, const double *lattice, const int32_t *atom_frame,
                 const int64_t *frame_start, const uint32_t *frame_types, const int32_t *list, const int32_t *shift,
                 const int32_t *counts, const int32_t *type_map, const float *vec, const void *mat, const float *tinfo,
                 int64_t natoms, int ntypes, int mnl, int m, double rcut, double rcut_smooth, double scale,
                 int32_t *status) {
    p.pos = pos; p.lattice = lattice; p.atom_frame = atom_frame; p.frame_start = frame_start;
    p.frame_types = frame_types;
    p.list = list; p.shift = shift; p.counts = counts; p.type_map = type_map;
    p.vec = vec; p.mat = mat; p.tinfo = tinfo; p.status = status;
    p.natoms = natoms; p.T = ntypes; p.mnl = mnl; p.m = m;
    p.rc2 = rcut * rcut; p.rmin = (float)rcut_smooth; p.rmax = (float)rcut; p.scale = (float)scale;
    p.iw = (float)(1.0 / (rcut - rcut_smooth));
}

}  // namespace
}  // namespace matpl

using namespace matpl;

MATPL_EXPORT int matpl_dp_train_tangent(const double *pos, const double *lattice, const int32_t *atom_frame,
                                        const int64_t *frame_start, const uint32_t *frame_types, const int32_t *list,
                                        const int32_t *shift, const int32_t *counts, const int32_t *type_map,
                                        const float *vec, const void *mat, const float *tinfo, int64_t natoms,
                                        int ntypes, int mnl, int m, int E, int bf16, double rcut, double rcut_smooth,
                                        double scale, const float *resid, float *dA, int32_t *nin, int32_t *status,
                                        void *stream) {
    MATPL_CHECK_ARG(rcut > rcut_smooth, "dp_train: rcut must exceed rcut_smooth");
    TrainParams tp{};
    fill_common(tp.f, pos, lattice, atom_frame, frame_start, frame_types, list, shift, counts, type_map, vec, mat, tinfo,
                natoms, ntypes, mnl, m, rcut, rcut_smooth, scale, status);
    tp.resid = resid;
    tp.dA = dA;
    tp.nin = nin;
    MATPL_CHECK_ARG(natoms == 0 || dA, "dp_train_tangent: null dA");
    return train_dispatch(tp, E, bf16, 0, (hipStream_t)stream);
}

MATPL_EXPORT int matpl_dp_train_embgrad(const double *pos, const double *lattice, const int32_t *atom_frame,
                                        const int64_t *frame_start, const uint32_t *frame_types, const int32_t *list,
                                        const int32_t *shift, const int32_t *counts, const int32_t *type_map,
                                        const float *vec, const void *mat, const float *tinfo, int64_t natoms,
                                        int ntypes, int mnl, int m, int E, int bf16, double rcut, double rcut_smooth,
                                        double scale, const float *resid, const float *abar, const float *ahat,
                                        float *grad, int chunk, int32_t *status, void *stream) {
    MATPL_CHECK_ARG(rcut > rcut_smooth, "dp_train: rcut must exceed rcut_smooth");
    MATPL_CHECK_ARG(chunk >= 1, "dp_train_embgrad: chunk must be positive");
    TrainParams tp{};
    fill_common(tp.f, pos, lattice, atom_frame, frame_start, frame_types, list, shift, counts, type_map, vec, mat, tinfo,
                natoms, ntypes, mnl, m, rcut, rcut_smooth, scale, status);
    tp.resid = resid;
    tp.abar = abar;
    tp.ahat = ahat;
    tp.grad = grad;
    tp.chunk = chunk;
    MATPL_CHECK_ARG(natoms == 0 || (abar && ahat && grad), "dp_train_embgrad: null pointer");
    return train_dispatch(tp, E, bf16, 1, (hipStream_t)stream);
}

MATPL_EXPORT int matpl_dp_train_grad_size(int E) {
    if (E == 25) return GradLayout<25>::SIZE;
    if (E == 16) return GradLayout<16>::SIZE;
    return -1;
}
