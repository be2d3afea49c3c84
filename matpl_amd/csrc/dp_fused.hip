// Fused DP se_e2_a descriptor + embedding-net kernels for gfx950.
//
// The reference evaluates DP.forward (src/model/dp_dp.py:167-530) as ~40
// torch kernels over every slot of a [B, N, T*m] padded neighbor tensor.
// Here one wavefront owns one centre atom i at a time:
//
//   fwd : neighbors of each type are read from the neighbor list, their
//         displacement is formed in FP64 (minimum image + cutoff test), the
//         in-cutoff ones are compacted into LDS as the normalized
//         Ri = ((s, s x/r, s y/r, s z/r) - davg) / dstd in FP32, and the
//         embedding net G = emb_{ti,tj}(Ri_0) runs on 32-neighbor tiles on the
//         matrix cores. The contraction A = sum Ri^T G runs on the matrix
//         cores too in BF16 mode (each tile's Ri^T and G^T are staged in LDS,
//         neighbors as the reduction index), in FP32 mode per lane with wave
//         sums. With a compression table G comes from the interval's
//         polynomial instead of the net. The empty slots of each type
//         bucket (m - n_t of them, plus neighbors beyond rcut) add the closed
//         form (m - n_t) * [s_pad,0,0,0]^T G_{ti,tj}(s_pad), s_pad =
//         -davg_0/dstd_0 (dp_dp.py:293-297 normalizes empty slots too).
//         Out: A [N][4][E] scaled by 1/(m * type_nums) (dp_dp.py:413).
//         type_nums and the buckets that take part are per frame: the
//         reference forms the (ti, tj) pairs only over the types of the
//         frame's atom_type list (dp_dp.py:150-165, 197), so a type absent
//         from that list adds no empty-slot term and does not count in
//         type_nums (frame_types; without it every model type takes part).
//   bwd : given dA = dE/dA, recomputes the embedding forward for each tile,
//         back-propagates through the net (same matrix-core layout), forms
//         dE/dRi per neighbor, contracts it with the analytic dRi/dx
//         (dp_dp.py:266-299) and produces the virial V[a][c] += f_a dR_c
//         (FP64 per frame; calculateVirial.cu:80) and the forces. With a
//         reverse-slot map (neighbor list symmetric) the force of slot
//         (i -> j) is stored in j's entry naming i, and gather_kernel sums,
//         for every atom, its centre term and its own entries, contiguously,
//         in a fixed order and without atomics (device-scope atomics leave
//         the per-XCD L2 on this part and throttle the scatter); without one,
//         FP32 atomics at j are used.
//
// Matrix-core layout. Activations live as [feature x neighbor] tiles in the
// 32x32 accumulator layout of v_mfma_f32_32x32x16_bf16: lane l holds
// neighbor n = l % 32 and, in register r (0..15), feature
// f(r, h) = 8*(r/4) + 4*h + r%4 with h = l / 32. A layer D = W^T X takes the
// previous accumulators directly as its B operand (registers 8c..8c+7 are
// K-chunk c), so the reduction index is permuted; the weights are stored
// pre-permuted per lane (fwd[l][r] = W[f(r,h)][l%32], bwd[l][r] =
// W[l%32][f(r,h)]) so no data moves between layers. The FP32 variant uses
// v_mfma_f32_32x32x2_f32 with K-chunk r = registers r of both halves.
#include "dp_fused_common.h"

namespace matpl {
namespace {

// Launch shape of each fused kernel (0 = forward, 1 = backward), process-wide,
// set by matpl_dp_fused_set_launch (the autotuner): the persistent grid's block
// cap, where the weights live (-1 auto: LDS when they fit, 0 global, 1 LDS) and
// the XCD-aware block order.
struct LaunchCfg {
    int max_blocks;
    int lds_weights;
    int xcd;
    int waves;          // wavefronts per block: 2, 4 or 8 (net mode; table mode always 4)
};
LaunchCfg g_launch[2] = {{4096, -1, 0, FUSED_WAVES}, {4096, -1, 0, FUSED_WAVES}};

template <int E, bool BF16, bool LW, int CK, bool GST, int W>
__global__ void __launch_bounds__(W * WAVE)
dp_fused_fwd_kernel(FusedParams p) {
    constexpr int NR = NReg<E>::value;
    extern __shared__ float4 smem[];
    // wid is the same on every lane: readfirstlane lets the compiler keep the atom index and the
    // per-atom bucket counts in scalar registers
    const int lane = threadIdx.x & (WAVE - 1), wid = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const BlockLds L = block_lds<BF16, LW, false, W>(p, smem, wid);
    float4 *nb = L.nb;
    int *nvs = L.nvs;
    const int h = lane >> 5, n = lane & 31;

    for (int64_t i = logical_block(p.xcd) * W + wid; i < p.natoms; i += (int64_t)gridDim.x * W) {
        const int ti = p.type_map[i];
        const float *tinfo = p.tinfo + ti * NTI;
        const AtomGeo g = atom_geo(p, i);
        if (!compact(p, g, i, lane, nb, nvs, nullptr) && lane == 0) atomicCAS(p.status, 0, (int)(i + 1));
        __builtin_amdgcn_wave_barrier();
        // BF16: A = sum over tiles of Ri^T G on the matrix cores. Each tile stages Ri^T [4][32]
        // and G^T [32][32] (bf16) in the wave's LDS; two 32x32x16 MFMAs (neighbors 0-15, 16-31)
        // accumulate D[w][e] with the neighbors as the reduction index, so lane e of half 0
        // ends the atom holding A[0..3][e] in D[0..3]. FP32: per-lane FMAs and wave sums.
        floatx16 D;
        float Aacc[4][16];
        if constexpr (BF16) {
#pragma unroll
            for (int r = 0; r < 16; r++) D[r] = 0.0f;
        } else {
#pragma unroll
            for (int w = 0; w < 4; w++)
#pragma unroll
                for (int r = 0; r < 16; r++) Aacc[w][r] = 0.0f;
        }
        for (int t = 0; t < p.T; t++) {
            const int nv = nvs[t];
            const int pair = ti * p.T + t;
            for (int k0 = 0; k0 < nv; k0 += TILE) {
                const int k = k0 + n;
                // re-read the weights every tile instead of keeping them live across tiles
                int tp = pair;
                asm volatile("" : "+v"(tp));
                float Ri[4] = {0.f, 0.f, 0.f, 0.f};
                float r, s, ds;
                if (k < nv) ri_of(nb[t * p.cap + k], tinfo, p.rmin, p.rmax, p.iw, Ri, r, s, ds);
                float G[16];
                if constexpr (CK > 0 && !BF16) {
                    float Gd[16];
                    table_eval<E, CK>(p, tp, Ri[0], h, G, Gd);
                } else if constexpr (CK == 0 && GST) {
                    // store mode: G and dG/ds of every in-cutoff neighbor for the backward
                    float Gd[16];
                    embed_fwd_tan<E, BF16, LW>(L.vec, L.mat, tp, lane, Ri[0], G, Gd);
                    if (k < nv) store_g<E, BF16>(p, i, t, k, h, G, Gd);
                } else if constexpr (CK == 0) {
                    float h1[16], t2[16], t3[16];
                    embed_fwd<E, BF16, LW>(L.vec, L.mat, tp, lane, Ri[0], h1, t2, t3, G);
                }
                const float live = (k < nv) ? 1.0f : 0.0f;
                if constexpr (BF16) {
                    constexpr int GS = CK > 0 ? 40 : 32;      // G^T row stride
                    __bf16 *rit = L.xt, *gt = L.xt + 4 * 32;
                    if (h == 0) {
#pragma unroll
                        for (int w = 0; w < 4; w++) rit[w * 32 + n] = (__bf16)(Ri[w] * live);
                    }
                    if constexpr (CK > 0) {
                        // table mode: half h evaluates neighbor 2 it + h, lane n its feature n, reading the
                        // interval's row in one coalesced load per half-wave; every neighbor of the tile is
                        // evaluated (a dead one at row 0, x = 0) so that G^T holds finite values only
                        float2 *tb = reinterpret_cast<float2 *>(L.xt + 4 * 32 + 32 * 40);
                        if (h == 0) {
                            int rowoff = 0;
                            float x = 0.0f;
                            if (k < nv) table_row<CK>(p, tp, Ri[0], rowoff, x);
                            tb[n] = make_float2(__int_as_float(rowoff), x);
                        }
                        __builtin_amdgcn_wave_barrier();
#pragma unroll 4
                        for (int it = 0; it < TILE / 2; it++) {
                            const int nn = 2 * it + h;
                            const float2 q = tb[nn];
                            float v, d;
                            horner<CK>(p.ctab + __float_as_int(q.x) + n * CK, q.y, v, d);
                            if (n < E) gt[n * GS + nn] = (__bf16)v;
                        }
                    } else {
#pragma unroll
                        for (int rr = 0; rr < NR; rr++) {
                            const int f = feat(rr, h);
                            if (f < E) gt[f * GS + n] = (__bf16)G[rr];
                        }
                    }
                    __builtin_amdgcn_wave_barrier();
                    // rows w >= 4 and columns e >= E of D are never read: their lanes read row 0
                    const int ia = n < 4 ? n : 0, jb = n < E ? n : 0;
                    const bf16x8 a0 = *reinterpret_cast<const bf16x8 *>(rit + ia * 32 + 8 * h);
                    const bf16x8 a1 = *reinterpret_cast<const bf16x8 *>(rit + ia * 32 + 16 + 8 * h);
                    const bf16x8 b0 = *reinterpret_cast<const bf16x8 *>(gt + jb * GS + 8 * h);
                    const bf16x8 b1 = *reinterpret_cast<const bf16x8 *>(gt + jb * GS + 16 + 8 * h);
                    D = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b0, D, 0, 0, 0);
                    D = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1, D, 0, 0, 0);
                    __builtin_amdgcn_wave_barrier();
                } else {
#pragma unroll
                    for (int w = 0; w < 4; w++)
#pragma unroll
                        for (int rr = 0; rr < NR; rr++) Aacc[w][rr] += (Ri[w] * live) * G[rr];
                }
            }
        }
        float out[4] = {0.f, 0.f, 0.f, 0.f};
        int f = -1, rr_f = 0, h_f = 0;        // feature this lane writes, and its (register, half)
        if constexpr (BF16) {
            if (h == 0 && n < E) {
                f = n;
                rr_f = 4 * (n >> 3) + (n & 3);
                h_f = (n >> 2) & 1;
#pragma unroll
                for (int w = 0; w < 4; w++) out[w] = D[w];
            }
        } else {
            // reduce over the neighbors (lanes of each half)
#pragma unroll
            for (int w = 0; w < 4; w++)
#pragma unroll
                for (int rr = 0; rr < NR; rr++) Aacc[w][rr] = half_sum(Aacc[w][rr]);
            // lane n of half h writes register n's feature (n < NR)
#pragma unroll
            for (int rr = 0; rr < NR; rr++)
                if (n == rr) {
#pragma unroll
                    for (int w = 0; w < 4; w++) out[w] = Aacc[w][rr];
                }
            if (n < NR && feat(n, h) < E) {
                f = feat(n, h);
                rr_f = n;
                h_f = h;
            }
        }
        if (f >= 0) {
            float scale;
            const uint32_t tmask = frame_type_mask(p, i, scale);
            float pad = 0.0f;
            for (int t = 0; t < p.T; t++) {
                if (!((tmask >> t) & 1u)) continue;     // the bucket of a type outside the frame's list
                const float npad = (float)(p.m - nvs[t]);
                pad += npad * vec_ptr(L.vec, ti * p.T + t, V_GPAD, h_f)[rr_f];
            }
#pragma unroll
            for (int w = 0; w < 4; w++) out[w] += tinfo[TI_SPAD + w] * pad;
            float *Ao = p.A + (size_t)i * 4 * E;
#pragma unroll
            for (int w = 0; w < 4; w++) Ao[w * E + f] = out[w] * scale;
        }
        __builtin_amdgcn_wave_barrier();
    }
}

template <int E, bool BF16, bool LW, int CK, bool GST, int W>
__global__ void __launch_bounds__(W * WAVE)
dp_fused_bwd_kernel(FusedParams p) {
    constexpr int NR = NReg<E>::value;
    extern __shared__ float4 smem[];
    // wid is the same on every lane: readfirstlane lets the compiler keep the atom index and the
    // per-atom bucket counts in scalar registers
    const int lane = threadIdx.x & (WAVE - 1), wid = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const BlockLds L = block_lds<BF16, LW, true, W>(p, smem, wid);
    float4 *nb = L.nb;
    int *nvs = L.nvs;
    float4 *dAs = L.dAs;        // [32 features] x (dA_0..dA_3) of the current atom, scale folded in
    double *vacc = L.vacc;      // this wave's virial of frame vframe (written by lane 0)
    const int h = lane >> 5, n = lane & 31;
    if (lane < 9) vacc[lane] = 0.0;
    int vframe = -1;
    // virial of half-0 lanes: an FP32 partial per atom (vf, the lane's share of the atom's slots),
    // added to an FP64 lane sum over this wave's atoms of the current frame (vd); the lane sums are
    // reduced across the wave into the frame sum only when the frame changes and at the end, and
    // they do not depend on which atoms share a wave beyond FP64 rounding
    float vf[9];
    double vd[9];
#pragma unroll
    for (int v = 0; v < 9; v++) { vf[v] = 0.0f; vd[v] = 0.0; }

    for (int64_t i = logical_block(p.xcd) * W + wid; i < p.natoms; i += (int64_t)gridDim.x * W) {
        const int ti = p.type_map[i];
        const float *tinfo = p.tinfo + ti * NTI;
        const AtomGeo g = atom_geo(p, i);
        const int frame = p.atom_frame ? p.atom_frame[i] : 0;
        if (frame != vframe) {
            // flush this wave's virial of the previous frame: lane sums -> frame sum
#pragma unroll
            for (int v = 0; v < 9; v++) {
                const double x = wave_sum(vd[v]);
                if (lane == 0) vacc[v] += x;
                vd[v] = 0.0;
            }
            __builtin_amdgcn_wave_barrier();
            if (vframe >= 0 && lane < 9 && vacc[lane] != 0.0) atomic_add(p.virial + 9 * vframe + lane, vacc[lane]);
            if (lane < 9) vacc[lane] = 0.0;
            vframe = frame;
        }
        // the atom's dA (two values per lane) is loaded before the compaction so that its latency
        // overlaps the neighbor loads
        float dav[2];
        {
            const float *dAi = p.dA + (size_t)i * 4 * E;
#pragma unroll
            for (int u = 0; u < 2; u++) {
                const int x = lane + u * WAVE, f = x >> 2, w = x & 3;
                dav[u] = f < E ? dAi[w * E + f] : 0.0f;
            }
        }
        compact(p, g, i, lane, nb, nvs, p.fslot);
        float scale;
        frame_type_mask(p, i, scale);
#pragma unroll
        for (int u = 0; u < 2; u++) reinterpret_cast<float *>(dAs)[lane + u * WAVE] = dav[u] * scale;
        __builtin_amdgcn_wave_barrier();
        // A operands of dG = dA Ri^T (32 features x 4, K split in two steps of 2): lane (n, h)
        // holds dA[n][h] and dA[n][2 + h]
        const float da0 = reinterpret_cast<const float *>(dAs)[4 * n + h];
        const float da1 = reinterpret_cast<const float *>(dAs)[4 * n + 2 + h];
        float fc[3] = {0.f, 0.f, 0.f};     // centre-force accumulator (half 0 lanes)
        for (int t = 0; t < p.T; t++) {
            const int nv = nvs[t];
            const int pair = ti * p.T + t;
            for (int k0 = 0; k0 < nv; k0 += TILE) {
                const int k = k0 + n;
                // re-read the weights every tile instead of keeping them live across tiles
                int tp = pair;
                asm volatile("" : "+v"(tp));
                const bool live = k < nv;
                float4 e = live ? nb[t * p.cap + k] : make_float4(1.f, 0.f, 0.f, 0.f);
                const int64_t slot = (i * p.T + t) * (int64_t)p.mnl + __float_as_int(e.w);
                // gather mode: the partner slot, loaded here so that its latency hides behind the tile
                const int rs = (p.fslot && live && h == 0) ? p.rev[slot] : -1;
                float Ri[4] = {0.f, 0.f, 0.f, 0.f};
                float r = 1.f, s = 0.f, ds = 0.f;
                if (live) ri_of(e, tinfo, p.rmin, p.rmax, p.iw, Ri, r, s, ds);
                float dRi[4] = {0.f, 0.f, 0.f, 0.f};
                float dsum = 0.0f;
                // dG = dA Ri^T on the matrix cores in FP32: rows = features (accumulator layout),
                // columns = the tile's neighbors; dead lanes have Ri = 0 and so dG = 0
                floatx16 acc;
#pragma unroll
                for (int rr = 0; rr < 16; rr++) acc[rr] = 0.0f;
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(da0, h ? Ri[1] : Ri[0], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(da1, h ? Ri[3] : Ri[2], acc, 0, 0, 0);
                if constexpr (CK > 0 || GST) {
                    // table mode (G from the interval's polynomial) or store mode (G and G' written by
                    // the forward): dE/ds = sum_f dG_f G'_f(s), no back-propagation through the net
                    float G[16], Gd[16];
                    if constexpr (CK > 0) {
                        table_eval<E, CK>(p, tp, Ri[0], h, G, Gd);
                    } else {
                        if (live) {
                            load_g<E, BF16>(p, i, t, k, h, G, Gd);
                        } else {
#pragma unroll
                            for (int rr = 0; rr < 16; rr++) { G[rr] = 0.0f; Gd[rr] = 0.0f; }
                        }
                    }
#pragma unroll
                    for (int rr = 0; rr < NR; rr++) {
                        const float4 d = dAs[feat(rr, h)];
                        dRi[0] += d.x * G[rr];
                        dRi[1] += d.y * G[rr];
                        dRi[2] += d.z * G[rr];
                        dRi[3] += d.w * G[rr];
                        dsum += acc[rr] * Gd[rr];
                    }
                } else {
                const float *w1 = vec_ptr(L.vec, tp, V_W1, h);
                float G[16];
                Gate16<false> g1, g2, g3;
                {
                    float h1[16], t2[16], t3[16];
                    embed_fwd<E, BF16, LW>(L.vec, L.mat, tp, lane, Ri[0], h1, t2, t3, G);
                    // the resnet scales r3, r2 are folded into the backward matrices
                    float q[16];
#pragma unroll
                    for (int rr = 0; rr < 16; rr++) q[rr] = rr < NR ? 1.0f - t3[rr] * t3[rr] : 0.0f;
                    g3.set(q);
#pragma unroll
                    for (int rr = 0; rr < 16; rr++) q[rr] = rr < NR ? 1.0f - t2[rr] * t2[rr] : 0.0f;
                    g2.set(q);
#pragma unroll
                    for (int rr = 0; rr < 16; rr++) q[rr] = rr < NR ? w1[rr] * (1.0f - h1[rr] * h1[rr]) : 0.0f;
                    g1.set(q);
                }
                // dRi = dA G (partial over this half's features)
#pragma unroll
                for (int rr = 0; rr < NR; rr++) {
                    const float4 d = dAs[feat(rr, h)];
                    dRi[0] += d.x * G[rr];
                    dRi[1] += d.y * G[rr];
                    dRi[2] += d.z * G[rr];
                    dRi[3] += d.w * G[rr];
                }
                // layer 3: G = h2 + r3 t3 -> dh2 = dG + (W3 diag r3) (dG (1-t3^2))
                float dz[16];
#pragma unroll
                for (int rr = 0; rr < 16; rr++) dz[rr] = rr < NR ? acc[rr] * g3.get(rr) : 0.0f;
                mfma_layer<BF16, LW>(mat_ptr<BF16>(L.mat, tp, M_W3B), lane, dz, acc);
                // layer 2: h2 = h1 + r2 t2 -> dh1 = dh2 + (W2 diag r2) (dh2 (1-t2^2))
#pragma unroll
                for (int rr = 0; rr < 16; rr++) dz[rr] = rr < NR ? acc[rr] * g2.get(rr) : 0.0f;
                mfma_layer<BF16, LW>(mat_ptr<BF16>(L.mat, tp, M_W2B), lane, dz, acc);
                // layer 1: h1 = tanh(s w1 + b1) -> ds = sum_f dh1 (1-h1^2) w1
#pragma unroll
                for (int rr = 0; rr < NR; rr++) dsum += acc[rr] * g1.get(rr);
                if constexpr (BF16) dsum *= 1.0f / TANH_K;     // w1 is stored pre-scaled
                }
                dRi[0] += dsum;
                // sums over the two halves' features, on the half-0 lanes that form the force
                fold_halves2(dRi[0], dRi[1]);
                fold_halves2(dRi[2], dRi[3]);
                // force of this slot: f_a = sum_w dE/dRi_w * Ri_d[w][a], Ri_d = -dfeat/dstd
                if (live && h == 0) {
                    const float ir = __builtin_amdgcn_rcpf(r);
                    const float x[3] = {e.x, e.y, e.z};
                    float gw[4];
#pragma unroll
                    for (int w = 0; w < 4; w++) gw[w] = dRi[w] * tinfo[TI_ISTD + w];
                    const float common = (ds - s * ir) * ir * ir;
                    const float gx = gw[1] * x[0] + gw[2] * x[1] + gw[3] * x[2];
                    const float cf = gw[0] * ds * ir + common * gx;
                    float f[3];
#pragma unroll
                    for (int a = 0; a < 3; a++) f[a] = -(x[a] * cf + gw[a + 1] * s * ir);
                    if (p.fslot) {
                        // gather mode: the force on j goes to the partner slot (j's entry naming i),
                        // so that j's gather reads its own entries contiguously
                        if (rs >= 0) p.fslot[rs] = make_float4(f[0], f[1], f[2], 0.f);
                    } else {
                        const int64_t j = g.fbase + p.list[slot] - 1;
                        atomic_add(p.force + 3 * j, f[0]);
                        atomic_add(p.force + 3 * j + 1, f[1]);
                        atomic_add(p.force + 3 * j + 2, f[2]);
                    }
#pragma unroll
                    for (int a = 0; a < 3; a++) {
                        fc[a] -= f[a];
#pragma unroll
                        for (int c = 0; c < 3; c++) vf[3 * a + c] += f[a] * x[c];
                    }
                }
            }
        }
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float x = half_sum(fc[a]);      // only half-0 lanes hold centre-force terms
            if (lane == 0) {
                if (p.fcenter) p.fcenter[3 * i + a] = x;
                else atomic_add(p.force + 3 * i + a, x);
            }
        }
#pragma unroll
        for (int v = 0; v < 9; v++) {
            vd[v] += (double)vf[v];
            vf[v] = 0.0f;
        }
        __builtin_amdgcn_wave_barrier();
    }
    // the last frame's lane sums
#pragma unroll
    for (int v = 0; v < 9; v++) {
        const double x = wave_sum(vd[v]);
        if (lane == 0) vacc[v] += x;
    }
    __builtin_amdgcn_wave_barrier();
    // virial: wave sums -> one set of atomics per block and frame
    __shared__ int vfr[W];
    if (lane == 0) vfr[wid] = vframe;
    __syncthreads();
    if (threadIdx.x < 9) {
        const int v = threadIdx.x;
        const double *vall = vacc - wid * 9;     // wave 0's row
        double acc = 0.0;
        int fr = -1;
        for (int w = 0; w < W; w++) {
            if (vfr[w] < 0) continue;
            if (vfr[w] != fr) {
                if (fr >= 0 && acc != 0.0) atomic_add(p.virial + 9 * fr + v, acc);
                fr = vfr[w];
                acc = 0.0;
            }
            acc += vall[w * 9 + v];
        }
        if (fr >= 0 && acc != 0.0) atomic_add(p.virial + 9 * fr + v, acc);
    }
}

// force[i] = fcenter[i] + sum over the used entries s of i's list of fslot[s]
// (each written by the partner atom through the reverse slot map); one
// wavefront per atom reading its own entries contiguously, fixed reduction
// order (deterministic).
__global__ void __launch_bounds__(256)
gather_kernel(const int32_t *__restrict__ counts, const float4 *__restrict__ fslot, const float *__restrict__ fcenter,
              float *__restrict__ force, int T, int mnl, int64_t n) {
    // the wave index is uniform: the atom index and its bucket offsets stay in scalar registers
    const int lane = threadIdx.x & (WAVE - 1), wid = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    for (int64_t i = (int64_t)blockIdx.x * 4 + wid; i < n; i += (int64_t)gridDim.x * 4) {
        // all buckets as one sequence, two entries per lane per pass, the pass's loads issued together
        int off[MAXT + 1];
        off[0] = 0;
#pragma unroll
        for (int t = 0; t < MAXT; t++) off[t + 1] = off[t] + (t < T ? counts[i * T + t] : 0);
        const int total = off[MAXT];
        float a0 = 0.f, a1 = 0.f, a2 = 0.f;
        for (int e0 = 0; e0 < total; e0 += 2 * WAVE) {
            float4 f[2];
#pragma unroll
            for (int u = 0; u < 2; u++) {
                const int e = e0 + u * WAVE + lane;
                int t = 0, base = 0;
#pragma unroll
                for (int q = 1; q < MAXT; q++) {
                    if (e >= off[q]) { t = q; base = off[q]; }
                }
                f[u] = e < total ? fslot[(i * T + t) * (int64_t)mnl + (e - base)] : make_float4(0.f, 0.f, 0.f, 0.f);
            }
#pragma unroll
            for (int u = 0; u < 2; u++) { a0 += f[u].x; a1 += f[u].y; a2 += f[u].z; }
        }
        a0 = wave_sum_dpp(a0); a1 = wave_sum_dpp(a1); a2 = wave_sum_dpp(a2);
        if (lane == 0) {
            force[3 * i] = fcenter[3 * i] + a0;
            force[3 * i + 1] = fcenter[3 * i + 1] + a1;
            force[3 * i + 2] = fcenter[3 * i + 2] + a2;
        }
    }
}

template <int E, bool BF16, bool LW, int CK = 0, bool GST = false, int W = FUSED_WAVES>
int launch_lw(FusedParams &p, bool fwd, hipStream_t st) {
    const size_t lds =
        fwd ? fused_lds_bytes<BF16, LW, false>(p.T, p.cap, W) : fused_lds_bytes<BF16, LW, true>(p.T, p.cap, W);
    MATPL_CHECK_ARG(lds <= 64 * 1024, "dp_fused: T*mnl too large for LDS (%zu bytes, %d waves per block)", lds, W);
    // persistent-style grid: a few blocks per CU, each strides over atoms
    const LaunchCfg &cfg = g_launch[fwd ? 0 : 1];
    int64_t blocks = (p.natoms + W - 1) / W;
    if (blocks > cfg.max_blocks) blocks = cfg.max_blocks;
    p.xcd = 0;
    if (cfg.xcd && blocks >= NXCD) {
        blocks -= blocks % NXCD;
        p.xcd = 1;
    }
    if (blocks < 1) blocks = 1;
    if (fwd) {
        hipLaunchKernelGGL((dp_fused_fwd_kernel<E, BF16, LW, CK, GST, W>), dim3((unsigned)blocks), dim3(W * WAVE), lds,
                           st, p);
        return check_launch("dp_fused_fwd_kernel");
    }
    hipLaunchKernelGGL((dp_fused_bwd_kernel<E, BF16, LW, CK, GST, W>), dim3((unsigned)blocks), dim3(W * WAVE), lds, st,
                       p);
    return check_launch("dp_fused_bwd_kernel");
}

// Net mode: weights in LDS or global memory, 2, 4 or 8 wavefronts per block.
template <int E, bool BF16, bool GST>
int launch_net(FusedParams &p, bool fwd, hipStream_t st, bool lw, int w) {
    if (lw) {
        if (w == 2) return launch_lw<E, BF16, true, 0, GST, 2>(p, fwd, st);
        if (w == 8) return launch_lw<E, BF16, true, 0, GST, 8>(p, fwd, st);
        return launch_lw<E, BF16, true, 0, GST, 4>(p, fwd, st);
    }
    if (w == 2) return launch_lw<E, BF16, false, 0, GST, 2>(p, fwd, st);
    if (w == 8) return launch_lw<E, BF16, false, 0, GST, 8>(p, fwd, st);
    return launch_lw<E, BF16, false, 0, GST, 4>(p, fwd, st);
}

// Weights go to LDS when they and the neighbor tiles fit in 64 KB per block
// (2 blocks = 8 waves per CU); otherwise they are read from global memory.
// Table mode (ck = 4 or 6 coefficients) reads no net weights.
template <int E, bool BF16>
int launch(FusedParams &p, bool fwd, hipStream_t st, int ck) {
    if (ck == 4) return launch_lw<E, BF16, false, 4>(p, fwd, st);
    if (ck == 6) return launch_lw<E, BF16, false, 6>(p, fwd, st);
    const int mode = g_launch[fwd ? 0 : 1].lds_weights;
    // a wavefront count set for another workload whose neighbor tiles would exceed the LDS budget
    // here falls back to fewer wavefronts per block
    int w = g_launch[fwd ? 0 : 1].waves;
    while (w > 2 && (fwd ? fused_lds_bytes<BF16, false, false>(p.T, p.cap, w)
                         : fused_lds_bytes<BF16, false, true>(p.T, p.cap, w)) > 64 * 1024)
        w /= 2;
    const bool fits = (fwd ? fused_lds_bytes<BF16, true, false>(p.T, p.cap, w)
                           : fused_lds_bytes<BF16, true, true>(p.T, p.cap, w)) <= 64 * 1024;
    MATPL_CHECK_ARG(mode != 1 || fits, "dp_fused: weights do not fit in LDS for this T and list width");
    const bool lw = mode == 1 || (mode < 0 && fits);
    if (p.gbuf) return launch_net<E, BF16, true>(p, fwd, st, lw, w);
    return launch_net<E, BF16, false>(p, fwd, st, lw, w);
}

int dispatch(FusedParams &p, int E, int bf16, bool fwd, hipStream_t st, int ck = 0) {
    MATPL_CHECK_ARG(p.T >= 1 && p.T <= MAXT, "dp_fused: ntypes must be in [1,%d]", MAXT);
    MATPL_CHECK_ARG(p.mnl >= 1 && p.m >= 1, "dp_fused: bad bucket sizes");
    p.cap = p.m < p.mnl ? p.m : p.mnl;
    if (p.natoms == 0) return OK;
    MATPL_CHECK_ARG(p.pos && p.lattice && p.list && p.shift && p.counts && p.type_map && p.vec &&
                        (ck ? p.ctab != nullptr : p.mat != nullptr) && p.tinfo && p.status, "dp_fused: null pointer");
    if (E == 25) return bf16 ? launch<25, true>(p, fwd, st, ck) : launch<25, false>(p, fwd, st, ck);
    if (E == 32) return bf16 ? launch<32, true>(p, fwd, st, ck) : launch<32, false>(p, fwd, st, ck);
    if (E == 16) return bf16 ? launch<16, true>(p, fwd, st, ck) : launch<16, false>(p, fwd, st, ck);
    set_error("dp_fused: embedding width %d not compiled (16, 25, 32)", E);
    return ERR_ARGUMENT;
}

}  // namespace
}  // namespace matpl

using namespace matpl;

MATPL_EXPORT int matpl_dp_fused_forward(const double *pos, const double *lattice, const int32_t *atom_frame,
                                        const int64_t *frame_start, const uint32_t *frame_types,
                                        const int32_t *list, const int32_t *shift,
                                        const int32_t *counts, const int32_t *type_map, const float *vec,
                                        const void *mat, const float *tinfo, int64_t natoms, int ntypes, int mnl,
                                        int m, int E, int bf16, double rcut, double rcut_smooth, double scale,
                                        float *A, int32_t *status, void *stream) {
    FusedParams p{};
    p.pos = pos; p.lattice = lattice; p.atom_frame = atom_frame; p.frame_start = frame_start;
    p.frame_types = frame_types;
    p.list = list; p.shift = shift; p.counts = counts; p.type_map = type_map;
    p.vec = vec; p.mat = mat; p.tinfo = tinfo; p.A = A; p.status = status;
    p.natoms = natoms; p.T = ntypes; p.mnl = mnl; p.m = m;
    p.rc2 = rcut * rcut; p.rmin = (float)rcut_smooth; p.rmax = (float)rcut; p.scale = (float)scale;
    MATPL_CHECK_ARG(rcut > rcut_smooth, "dp_fused: rcut must exceed rcut_smooth");
    p.iw = (float)(1.0 / (rcut - rcut_smooth));
    MATPL_CHECK_ARG(natoms == 0 || A, "dp_fused_forward: null A");
    return dispatch(p, E, bf16, true, (hipStream_t)stream);
}

MATPL_EXPORT int matpl_dp_fused_backward(const double *pos, const double *lattice, const int32_t *atom_frame,
                                         const int64_t *frame_start, const uint32_t *frame_types,
                                         const int32_t *list, const int32_t *shift,
                                         const int32_t *counts, const int32_t *type_map, const float *vec,
                                         const void *mat, const float *tinfo, int64_t natoms, int ntypes, int mnl,
                                         int m, int E, int bf16, double rcut, double rcut_smooth, double scale,
                                         const float *dA, const int32_t *rev, float *fslot, float *fcenter,
                                         float *force, double *virial, int32_t *status, void *stream) {
    FusedParams p{};
    p.pos = pos; p.lattice = lattice; p.atom_frame = atom_frame; p.frame_start = frame_start;
    p.frame_types = frame_types;
    p.list = list; p.shift = shift; p.counts = counts; p.type_map = type_map;
    p.vec = vec; p.mat = mat; p.tinfo = tinfo; p.dA = dA; p.force = force; p.virial = virial; p.status = status;
    p.natoms = natoms; p.T = ntypes; p.mnl = mnl; p.m = m;
    p.rc2 = rcut * rcut; p.rmin = (float)rcut_smooth; p.rmax = (float)rcut; p.scale = (float)scale;
    MATPL_CHECK_ARG(rcut > rcut_smooth, "dp_fused: rcut must exceed rcut_smooth");
    p.iw = (float)(1.0 / (rcut - rcut_smooth));
    MATPL_CHECK_ARG(natoms == 0 || (dA && force && virial), "dp_fused_backward: null output");
    if (rev) {
        MATPL_CHECK_ARG(fslot && fcenter && aligned16(fslot), "dp_fused_backward: gather mode needs fslot/fcenter");
        p.fslot = reinterpret_cast<float4 *>(fslot);
        p.rev = rev;
        p.fcenter = fcenter;
    }
    return dispatch(p, E, bf16, false, (hipStream_t)stream);
}

// Store mode: the forward also writes G and dG/ds of every in-cutoff neighbor (fp16 pairs) into
// gbuf, and the backward reads them instead of re-running the net and back-propagating through it.
MATPL_EXPORT int64_t matpl_dp_fused_gstore_words(int64_t natoms, int ntypes, int mnl, int m, int E, int bf16) {
    const int cap = m < mnl ? m : mnl;
    int w = -1;
    if (E == 25) w = bf16 ? GStore<25, true>::WORDS : GStore<25, false>::WORDS;
    if (E == 32) w = bf16 ? GStore<32, true>::WORDS : GStore<32, false>::WORDS;
    if (E == 16) w = bf16 ? GStore<16, true>::WORDS : GStore<16, false>::WORDS;
    if (w < 0 || natoms < 0 || ntypes < 1 || cap < 1) return -1;
    return natoms * ntypes * (int64_t)cap * 2 * w;
}

MATPL_EXPORT int matpl_dp_fused_forward_store(const double *pos, const double *lattice, const int32_t *atom_frame,
                                              const int64_t *frame_start, const uint32_t *frame_types,
                                              const int32_t *list, const int32_t *shift, const int32_t *counts,
                                              const int32_t *type_map, const float *vec, const void *mat,
                                              const float *tinfo, int64_t natoms, int ntypes, int mnl, int m, int E,
                                              int bf16, double rcut, double rcut_smooth, double scale, float *A,
                                              uint32_t *gbuf, int32_t *status, void *stream) {
    MATPL_CHECK_ARG(natoms == 0 || (gbuf && aligned16(gbuf)), "dp_fused_forward_store: gbuf must be 16-byte aligned");
    FusedParams p{};
    p.pos = pos; p.lattice = lattice; p.atom_frame = atom_frame; p.frame_start = frame_start;
    p.frame_types = frame_types;
    p.list = list; p.shift = shift; p.counts = counts; p.type_map = type_map;
    p.vec = vec; p.mat = mat; p.tinfo = tinfo; p.A = A; p.status = status; p.gbuf = gbuf;
    p.natoms = natoms; p.T = ntypes; p.mnl = mnl; p.m = m;
    p.rc2 = rcut * rcut; p.rmin = (float)rcut_smooth; p.rmax = (float)rcut; p.scale = (float)scale;
    MATPL_CHECK_ARG(rcut > rcut_smooth, "dp_fused: rcut must exceed rcut_smooth");
    p.iw = (float)(1.0 / (rcut - rcut_smooth));
    MATPL_CHECK_ARG(natoms == 0 || A, "dp_fused_forward: null A");
    return dispatch(p, E, bf16, true, (hipStream_t)stream);
}

MATPL_EXPORT int matpl_dp_fused_backward_stored(const double *pos, const double *lattice, const int32_t *atom_frame,
                                                const int64_t *frame_start, const uint32_t *frame_types,
                                                const int32_t *list, const int32_t *shift, const int32_t *counts,
                                                const int32_t *type_map, const float *vec, const void *mat,
                                                const float *tinfo, int64_t natoms, int ntypes, int mnl, int m, int E,
                                                int bf16, double rcut, double rcut_smooth, double scale,
                                                const float *dA, const uint32_t *gbuf, const int32_t *rev, float *fslot,
                                                float *fcenter, float *force, double *virial, int32_t *status,
                                                void *stream) {
    MATPL_CHECK_ARG(natoms == 0 || (gbuf && aligned16(gbuf)), "dp_fused_backward_stored: gbuf must be 16-byte aligned");
    FusedParams p{};
    p.pos = pos; p.lattice = lattice; p.atom_frame = atom_frame; p.frame_start = frame_start;
    p.frame_types = frame_types;
    p.list = list; p.shift = shift; p.counts = counts; p.type_map = type_map;
    p.vec = vec; p.mat = mat; p.tinfo = tinfo; p.dA = dA; p.force = force; p.virial = virial; p.status = status;
    p.gbuf = const_cast<uint32_t *>(gbuf);
    p.natoms = natoms; p.T = ntypes; p.mnl = mnl; p.m = m;
    p.rc2 = rcut * rcut; p.rmin = (float)rcut_smooth; p.rmax = (float)rcut; p.scale = (float)scale;
    MATPL_CHECK_ARG(rcut > rcut_smooth, "dp_fused: rcut must exceed rcut_smooth");
    p.iw = (float)(1.0 / (rcut - rcut_smooth));
    MATPL_CHECK_ARG(natoms == 0 || (dA && force && virial), "dp_fused_backward: null output");
    if (rev) {
        MATPL_CHECK_ARG(fslot && fcenter && aligned16(fslot), "dp_fused_backward: gather mode needs fslot/fcenter");
        p.fslot = reinterpret_cast<float4 *>(fslot);
        p.rev = rev;
        p.fcenter = fcenter;
    }
    return dispatch(p, E, bf16, false, (hipStream_t)stream);
}

// Table-mode descriptor shared by the two compressed entry points.
static int set_table(FusedParams &p, const float *ctab, int order, int rows, int n_fine, double x0, double dx,
                     int E) {
    MATPL_CHECK_ARG(order == 3 || order == 5, "dp_fused_compressed: order must be 3 or 5");
    MATPL_CHECK_ARG(rows >= 1 && n_fine >= 0 && n_fine <= rows && dx > 0.0, "dp_fused_compressed: bad table sizes");
    MATPL_CHECK_ARG(E <= 32, "dp_fused_compressed: embedding width above 32");
    MATPL_CHECK_ARG(ctab && aligned16(ctab), "dp_fused_compressed: table must be 16-byte aligned");
    p.ctab = ctab;
    p.crows = rows;
    p.cnf = n_fine;
    p.cx0 = (float)x0;
    p.cdx = (float)dx;
    p.cidx = (float)(1.0 / dx);
    p.cxfe = (float)(x0 + n_fine * dx);
    return OK;
}

MATPL_EXPORT int matpl_dp_fused_forward_compressed(
    const double *pos, const double *lattice, const int32_t *atom_frame, const int64_t *frame_start,
    const uint32_t *frame_types, const int32_t *list, const int32_t *shift, const int32_t *counts,
    const int32_t *type_map, const float *vec,
    const float *ctab, int order, int rows, int n_fine, double x0, double dx, const float *tinfo, int64_t natoms,
    int ntypes, int mnl, int m, int E, int bf16, double rcut, double rcut_smooth, double scale, float *A,
    int32_t *status, void *stream) {
    FusedParams p{};
    int rc = set_table(p, ctab, order, rows, n_fine, x0, dx, E);
    if (rc) return rc;
    p.pos = pos; p.lattice = lattice; p.atom_frame = atom_frame; p.frame_start = frame_start;
    p.frame_types = frame_types;
    p.list = list; p.shift = shift; p.counts = counts; p.type_map = type_map;
    p.vec = vec; p.tinfo = tinfo; p.A = A; p.status = status;
    p.natoms = natoms; p.T = ntypes; p.mnl = mnl; p.m = m;
    p.rc2 = rcut * rcut; p.rmin = (float)rcut_smooth; p.rmax = (float)rcut; p.scale = (float)scale;
    MATPL_CHECK_ARG(rcut > rcut_smooth, "dp_fused: rcut must exceed rcut_smooth");
    p.iw = (float)(1.0 / (rcut - rcut_smooth));
    MATPL_CHECK_ARG(natoms == 0 || A, "dp_fused_forward: null A");
    return dispatch(p, E, bf16, true, (hipStream_t)stream, order + 1);
}

MATPL_EXPORT int matpl_dp_fused_backward_compressed(
    const double *pos, const double *lattice, const int32_t *atom_frame, const int64_t *frame_start,
    const uint32_t *frame_types, const int32_t *list, const int32_t *shift, const int32_t *counts,
    const int32_t *type_map, const float *vec,
    const float *ctab, int order, int rows, int n_fine, double x0, double dx, const float *tinfo, int64_t natoms,
    int ntypes, int mnl, int m, int E, int bf16, double rcut, double rcut_smooth, double scale, const float *dA,
    const int32_t *rev, float *fslot, float *fcenter, float *force, double *virial, int32_t *status, void *stream) {
    FusedParams p{};
    int rc = set_table(p, ctab, order, rows, n_fine, x0, dx, E);
    if (rc) return rc;
    p.pos = pos; p.lattice = lattice; p.atom_frame = atom_frame; p.frame_start = frame_start;
    p.frame_types = frame_types;
    p.list = list; p.shift = shift; p.counts = counts; p.type_map = type_map;
    p.vec = vec; p.tinfo = tinfo; p.dA = dA; p.force = force; p.virial = virial; p.status = status;
    p.natoms = natoms; p.T = ntypes; p.mnl = mnl; p.m = m;
    p.rc2 = rcut * rcut; p.rmin = (float)rcut_smooth; p.rmax = (float)rcut; p.scale = (float)scale;
    MATPL_CHECK_ARG(rcut > rcut_smooth, "dp_fused: rcut must exceed rcut_smooth");
    p.iw = (float)(1.0 / (rcut - rcut_smooth));
    MATPL_CHECK_ARG(natoms == 0 || (dA && force && virial), "dp_fused_backward: null output");
    if (rev) {
        MATPL_CHECK_ARG(fslot && fcenter && aligned16(fslot), "dp_fused_backward: gather mode needs fslot/fcenter");
        p.fslot = reinterpret_cast<float4 *>(fslot);
        p.rev = rev;
        p.fcenter = fcenter;
    }
    return dispatch(p, E, bf16, false, (hipStream_t)stream, order + 1);
}

MATPL_EXPORT int matpl_dp_fused_gather(const int32_t *counts, const float *fslot, const float *fcenter, float *force,
                                       int ntypes, int mnl, int64_t natoms, void *stream) {
    if (natoms == 0) return OK;
    MATPL_CHECK_ARG(counts && fslot && fcenter && force && aligned16(fslot), "dp_fused_gather: bad pointer");
    MATPL_CHECK_ARG(ntypes >= 1 && mnl >= 1, "dp_fused_gather: bad sizes");
    hipLaunchKernelGGL(gather_kernel, dim3(grid_cap(natoms, 4)), dim3(256), 0, (hipStream_t)stream, counts,
                       reinterpret_cast<const float4 *>(fslot), fcenter, force, ntypes, mnl, natoms);
    return check_launch("gather_kernel");
}

MATPL_EXPORT int matpl_dp_fused_set_launch(int kernel, int max_blocks, int lds_weights, int xcd) {
    MATPL_CHECK_ARG(kernel == 0 || kernel == 1, "dp_fused_set_launch: kernel is 0 (forward) or 1 (backward)");
    MATPL_CHECK_ARG(max_blocks >= 1 && max_blocks <= (1 << 20), "dp_fused_set_launch: max_blocks out of range");
    MATPL_CHECK_ARG(lds_weights >= -1 && lds_weights <= 1 && (xcd == 0 || xcd == 1),
                    "dp_fused_set_launch: lds_weights in {-1, 0, 1}, xcd in {0, 1}");
    g_launch[kernel] = LaunchCfg{max_blocks, lds_weights, xcd, g_launch[kernel].waves};
    return OK;
}

// Wavefronts per block (2, 4 or 8) of the net-mode fused forward (kernel 0) or backward (1) for
// later calls in this process; results do not depend on it beyond the virial's FP64 summation order.
MATPL_EXPORT int matpl_dp_fused_set_waves(int kernel, int waves) {
    MATPL_CHECK_ARG(kernel == 0 || kernel == 1, "dp_fused_set_waves: kernel is 0 (forward) or 1 (backward)");
    MATPL_CHECK_ARG(waves == 2 || waves == 4 || waves == 8, "dp_fused_set_waves: waves in {2, 4, 8}");
    g_launch[kernel].waves = waves;
    return OK;
}

