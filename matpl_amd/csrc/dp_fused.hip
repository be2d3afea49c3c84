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
//         matrix cores. The contraction A = sum Ri^T G is accumulated in FP32
//         per lane and reduced across the wave; the empty slots of each type
//         bucket (m - n_t of them, plus neighbors beyond rcut) add the closed
//         form (m - n_t) * [s_pad,0,0,0]^T G_{ti,tj}(s_pad), s_pad =
//         -davg_0/dstd_0 (dp_dp.py:293-297 normalizes empty slots too).
//         Out: A [N][4][E] scaled by 1/(m * type_nums) (dp_dp.py:413).
//   bwd : given dA = dE/dA, recomputes the embedding forward for each tile,
//         back-propagates through the net (same matrix-core layout), forms
//         dE/dRi per neighbor, contracts it with the analytic dRi/dx
//         (dp_dp.py:266-299) and produces the virial V[a][c] += f_a dR_c
//         (FP64 per frame; calculateVirial.cu:80) and the forces. With a
//         reverse-slot map (neighbor list symmetric) each slot's neighbor
//         force is stored in its list slot and gather_kernel sums, for every
//         atom, its centre term and the slots that point back at it, in a
//         fixed order and without atomics (device-scope atomics leave the
//         per-XCD L2 on this part and throttle the scatter); without one,
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
#include "common.h"

namespace matpl {
namespace {

constexpr int WAVE = 64;
constexpr int TILE = 32;
constexpr int FUSED_WAVES = 4;     // wavefronts per block
constexpr int MAXT = 8;

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float floatx16;

__device__ __forceinline__ constexpr int feat(int r, int h) { return 8 * (r >> 2) + 4 * h + (r & 3); }

// number of accumulator registers that hold at least one real feature
// (register r holds features f(r,0) <= f(r,1); E = 25 -> 13 registers)
constexpr int nreg(int E) {
    int n = 0;
    for (int r = 0; r < 16; r++)
        if (feat(r, 0) < E) n = r + 1;
    return n;
}
template <int E>
struct NReg {
    static constexpr int value = nreg(E);
};
static_assert(nreg(25) == 13 && nreg(32) == 16 && nreg(16) == 8, "register count");

// Per-pair vector parameters, each [2 halves][16 registers] floats.
enum VecId { V_W1 = 0, V_B1, V_B2, V_B3, V_R2, V_R3, V_GPAD, NVEC };
// Per-pair matrices, each [64 lanes][16] values (bf16 or f32).
enum MatId { M_W2F = 0, M_W3F, M_W2B, M_W3B, NMAT };
// Per-centre-type info, floats.
// davg[4], 1/dstd[4], dstd[4], Ri of an empty slot (-davg/dstd)[4]
enum TypeInfo { TI_DAVG = 0, TI_ISTD = 4, TI_STD = 8, TI_SPAD = 12, NTI = 16 };

struct FusedParams {
    const double *pos;          // [N][3] Cartesian
    const double *lattice;      // [F][9], row d = lattice vector d
    const int32_t *atom_frame;  // [N] frame of each atom (nullptr: all frame 0)
    const int64_t *frame_start; // [F+1] first atom of each frame (nullptr: 0)
    const int32_t *list;        // [N][T][mnl] 1-based within frame, 0 = empty
    const int32_t *shift;       // [N][T][mnl] packed image
    const int32_t *counts;      // [N][T] entries used per bucket
    const int32_t *type_map;    // [N]
    const float *vec;           // [T*T][NVEC][2][16]
    const void *mat;            // [T*T][NMAT][64][16] bf16 or f32
    const float *tinfo;         // [T][NTI]
    const float *dA;            // [N][4][E] (bwd)
    float *A;                   // [N][4][E] (fwd)
    float *force;               // [N][3] (bwd, accumulated)
    double *virial;             // [F][9] (bwd, accumulated)
    float4 *fslot;              // [N][T][mnl] per-slot neighbor force (bwd, gather mode)
    float *fcenter;             // [N][3] centre term (bwd, gather mode)
    int32_t *status;            // 1 + first atom whose in-cutoff count exceeds m
    int64_t natoms;
    int T, mnl, m;
    double rc2;                 // model cutoff^2 (FP64 test)
    float rmin, rmax, scale;    // scale = 1 / (m * type_nums)
};

__device__ __forceinline__ float fast_tanh(float x) {
    // tanh x = 1 - 2 / (2^(2x log2 e) + 1); saturates correctly at +-inf
    const float e = __builtin_amdgcn_exp2f(x * 2.8853900817779268f);
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
}

template <bool BF16>
__device__ __forceinline__ float act(float x) {
    if constexpr (BF16) return fast_tanh(x);
    else return tanhf(x);
}

// s(r) and ds/dr, dp_dp.py:244-249 (FP32)
__device__ __forceinline__ void switch_fn(float r, float rmin, float rmax, float &s, float &ds) {
    if (r < rmin) {
        s = 1.0f / r;
        ds = -s * s;
    } else if (r < rmax) {
        const float inv = 1.0f / (rmax - rmin);
        const float u = (r - rmin) * inv;
        const float vv = u * u * u * (-6.0f * u * u + 15.0f * u - 10.0f) + 1.0f;
        const float dvv = -30.0f * u * u * (u - 1.0f) * (u - 1.0f) * inv;
        const float ir = 1.0f / r;
        s = vv * ir;
        ds = (dvv - s) * ir;
    } else {
        s = 0.0f;
        ds = 0.0f;
    }
}

// One matrix-core layer: acc (C-init) += W^T x, x given as 16 fp32 registers.
template <bool BF16>
__device__ __forceinline__ void mfma_layer(const void *mat, int lane, const float x[16], floatx16 &acc) {
    if constexpr (BF16) {
        const bf16x8 *w = reinterpret_cast<const bf16x8 *>(mat) + lane * 2;
        bf16x8 b0, b1;
#pragma unroll
        for (int q = 0; q < 8; q++) { b0[q] = (__bf16)x[q]; b1[q] = (__bf16)x[8 + q]; }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w[0], b0, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w[1], b1, acc, 0, 0, 0);
    } else {
        const float *w = reinterpret_cast<const float *>(mat) + lane * 16;
#pragma unroll
        for (int r = 0; r < 16; r++) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(w[r], x[r], acc, 0, 0, 0);
    }
}

template <bool BF16>
__device__ __forceinline__ const void *mat_ptr(const void *base, int pair, int which) {
    const size_t per = (size_t)64 * 16 * (BF16 ? sizeof(__bf16) : sizeof(float));
    return reinterpret_cast<const char *>(base) + ((size_t)pair * NMAT + which) * per;
}

__device__ __forceinline__ const float *vec_ptr(const float *base, int pair, int which, int h) {
    return base + (((size_t)pair * NVEC + which) * 2 + h) * 16;
}

// Embedding forward for one tile. s = embedding input of this lane's neighbor.
// Out: h1 = tanh(layer1), t2, t3 = tanh of layers 2/3, G = output.
template <int E, bool BF16>
__device__ __forceinline__ void embed_fwd(const FusedParams &p, int pair, int lane, float s, float h1[16],
                                          float t2[16], float t3[16], float G[16]) {
    constexpr int NR = NReg<E>::value;
    const int h = lane >> 5;
    const float *w1 = vec_ptr(p.vec, pair, V_W1, h), *b1 = vec_ptr(p.vec, pair, V_B1, h);
    const float *b2 = vec_ptr(p.vec, pair, V_B2, h), *b3 = vec_ptr(p.vec, pair, V_B3, h);
    const float *r2 = vec_ptr(p.vec, pair, V_R2, h), *r3 = vec_ptr(p.vec, pair, V_R3, h);
#pragma unroll
    for (int r = 0; r < 16; r++) h1[r] = (r < NR) ? act<BF16>(s * w1[r] + b1[r]) : 0.0f;
    floatx16 acc;
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = b2[r];
    mfma_layer<BF16>(mat_ptr<BF16>(p.mat, pair, M_W2F), lane, h1, acc);
    float h2[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        t2[r] = (r < NR) ? act<BF16>(acc[r]) : 0.0f;
        h2[r] = h1[r] + r2[r] * t2[r];
    }
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = b3[r];
    mfma_layer<BF16>(mat_ptr<BF16>(p.mat, pair, M_W3F), lane, h2, acc);
#pragma unroll
    for (int r = 0; r < 16; r++) {
        t3[r] = (r < NR) ? act<BF16>(acc[r]) : 0.0f;
        G[r] = h2[r] + r3[r] * t3[r];
    }
}

struct AtomGeo {
    double xi, yi, zi;
    const double *L;
    int64_t fbase;
};

__device__ __forceinline__ AtomGeo atom_geo(const FusedParams &p, int64_t i) {
    AtomGeo g;
    const int f = p.atom_frame ? p.atom_frame[i] : 0;
    g.L = p.lattice + 9 * (int64_t)f;
    g.fbase = p.frame_start ? p.frame_start[f] : 0;
    g.xi = p.pos[3 * i]; g.yi = p.pos[3 * i + 1]; g.zi = p.pos[3 * i + 2];
    return g;
}

// Displacement of bucket entry (i,t,k) in FP64; returns j (global) and r^2.
__device__ __forceinline__ int64_t neighbor_disp(const FusedParams &p, const AtomGeo &g, int64_t slot, double &dx,
                                                 double &dy, double &dz) {
    const int64_t j = g.fbase + p.list[slot] - 1;
    const int sh = p.shift[slot];
    const double s1 = (double)((sh & 255) - 128), s2 = (double)(((sh >> 8) & 255) - 128),
                 s3 = (double)(((sh >> 16) & 255) - 128);
    const double *L = g.L;
    dx = (p.pos[3 * j] - g.xi) + (s1 * L[0] + s2 * L[3] + s3 * L[6]);
    dy = (p.pos[3 * j + 1] - g.yi) + (s1 * L[1] + s2 * L[4] + s3 * L[7]);
    dz = (p.pos[3 * j + 2] - g.zi) + (s1 * L[2] + s2 * L[5] + s3 * L[8]);
    return j;
}

// Compacts the in-cutoff neighbors of atom i into LDS: nb[t*mnl + c] =
// (dx, dy, dz, bucket slot k as float bits); nv[t] = count. Used bucket
// entries beyond the cutoff get a zero per-slot force when fslot is given.
// Returns false on overflow.
__device__ __forceinline__ bool compact(const FusedParams &p, const AtomGeo &g, int64_t i, int lane, float4 *nb,
                                        int *nv, float4 *fslot) {
    bool ok_all = true;
    for (int t = 0; t < p.T; t++) {
        const int cnt = p.counts[i * p.T + t];
        int n = 0;
        for (int k0 = 0; k0 < cnt; k0 += WAVE) {
            const int k = k0 + lane;
            bool ok = false;
            float4 e = make_float4(0.f, 0.f, 0.f, 0.f);
            if (k < cnt) {
                double dx, dy, dz;
                const int64_t slot = (i * p.T + t) * (int64_t)p.mnl + k;
                neighbor_disp(p, g, slot, dx, dy, dz);
                const double r2 = dx * dx + dy * dy + dz * dz;
                ok = r2 < p.rc2 && r2 > 1e-10;
                e = make_float4((float)dx, (float)dy, (float)dz, __int_as_float(k));
                if (!ok && fslot) fslot[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
            const uint64_t bal = __ballot(ok);
            const int idx = n + __popcll(bal & ((1ull << lane) - 1ull));
            if (ok) nb[t * p.mnl + idx] = e;
            n += __popcll(bal);
        }
        if (lane == 0) nv[t] = n;
        if (n > p.m) ok_all = false;
    }
    return ok_all;
}

// Normalized Ri of a compacted neighbor (FP32).
__device__ __forceinline__ void ri_of(const float4 e, const float *ti, float rmin, float rmax, float Ri[4],
                                      float &r, float &s, float &ds) {
    r = sqrtf(e.x * e.x + e.y * e.y + e.z * e.z);
    switch_fn(r, rmin, rmax, s, ds);
    const float ir = 1.0f / r;
    Ri[0] = (s - ti[TI_DAVG]) * ti[TI_ISTD];
    Ri[1] = (s * e.x * ir - ti[TI_DAVG + 1]) * ti[TI_ISTD + 1];
    Ri[2] = (s * e.y * ir - ti[TI_DAVG + 2]) * ti[TI_ISTD + 2];
    Ri[3] = (s * e.z * ir - ti[TI_DAVG + 3]) * ti[TI_ISTD + 3];
}

// Sum over the 32 lanes of each half-wave.
__device__ __forceinline__ float half_sum(float v) {
#pragma unroll
    for (int off = 1; off < 32; off <<= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <int E, bool BF16>
__global__ void __launch_bounds__(FUSED_WAVES * WAVE)
dp_fused_fwd_kernel(FusedParams p) {
    constexpr int NR = NReg<E>::value;
    extern __shared__ float4 smem[];
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    float4 *nb = smem + (size_t)wid * p.T * p.mnl;
    int *nvs = reinterpret_cast<int *>(smem + (size_t)FUSED_WAVES * p.T * p.mnl) + wid * MAXT;
    const int h = lane >> 5, n = lane & 31;

    for (int64_t i = (int64_t)blockIdx.x * FUSED_WAVES + wid; i < p.natoms; i += (int64_t)gridDim.x * FUSED_WAVES) {
        const int ti = p.type_map[i];
        const float *tinfo = p.tinfo + ti * NTI;
        const AtomGeo g = atom_geo(p, i);
        if (!compact(p, g, i, lane, nb, nvs, nullptr) && lane == 0) atomicCAS(p.status, 0, (int)(i + 1));
        __builtin_amdgcn_wave_barrier();
        float Aacc[4][16];
#pragma unroll
        for (int w = 0; w < 4; w++)
#pragma unroll
            for (int r = 0; r < 16; r++) Aacc[w][r] = 0.0f;
        for (int t = 0; t < p.T; t++) {
            const int nv = nvs[t];
            const int pair = ti * p.T + t;
            for (int k0 = 0; k0 < nv; k0 += TILE) {
                const int k = k0 + n;
                float Ri[4] = {0.f, 0.f, 0.f, 0.f};
                float r, s, ds;
                if (k < nv) ri_of(nb[t * p.mnl + k], tinfo, p.rmin, p.rmax, Ri, r, s, ds);
                float h1[16], t2[16], t3[16], G[16];
                embed_fwd<E, BF16>(p, pair, lane, Ri[0], h1, t2, t3, G);
                const float live = (k < nv) ? 1.0f : 0.0f;
#pragma unroll
                for (int w = 0; w < 4; w++)
#pragma unroll
                    for (int rr = 0; rr < NR; rr++) Aacc[w][rr] += (Ri[w] * live) * G[rr];
            }
        }
        // reduce over the neighbors (lanes of each half)
#pragma unroll
        for (int w = 0; w < 4; w++)
#pragma unroll
            for (int rr = 0; rr < NR; rr++) Aacc[w][rr] = half_sum(Aacc[w][rr]);
        // lane n of half h writes register n's feature (n < NR)
        float out[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int rr = 0; rr < NR; rr++)
            if (n == rr) {
#pragma unroll
                for (int w = 0; w < 4; w++) out[w] = Aacc[w][rr];
            }
        if (n < NR) {
            const int f = feat(n, h);
            if (f < E) {
                float pad = 0.0f;
                for (int t = 0; t < p.T; t++) {
                    const float npad = (float)(p.m - nvs[t]);
                    pad += npad * vec_ptr(p.vec, ti * p.T + t, V_GPAD, h)[n];
                }
#pragma unroll
                for (int w = 0; w < 4; w++) out[w] += tinfo[TI_SPAD + w] * pad;
                float *Ao = p.A + (size_t)i * 4 * E;
#pragma unroll
                for (int w = 0; w < 4; w++) Ao[w * E + f] = out[w] * p.scale;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

template <int E, bool BF16>
__global__ void __launch_bounds__(FUSED_WAVES * WAVE)
dp_fused_bwd_kernel(FusedParams p) {
    constexpr int NR = NReg<E>::value;
    extern __shared__ float4 smem[];
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    float4 *nb = smem + (size_t)wid * p.T * p.mnl;
    int *nvs = reinterpret_cast<int *>(smem + (size_t)FUSED_WAVES * p.T * p.mnl) + wid * MAXT;
    const int h = lane >> 5, n = lane & 31;
    double vir[9];
#pragma unroll
    for (int v = 0; v < 9; v++) vir[v] = 0.0;
    int vframe = -1;

    for (int64_t i = (int64_t)blockIdx.x * FUSED_WAVES + wid; i < p.natoms; i += (int64_t)gridDim.x * FUSED_WAVES) {
        const int ti = p.type_map[i];
        const float *tinfo = p.tinfo + ti * NTI;
        const AtomGeo g = atom_geo(p, i);
        const int frame = p.atom_frame ? p.atom_frame[i] : 0;
        if (frame != vframe) {
            // flush the per-lane virial of the previous frame
            if (vframe >= 0) {
#pragma unroll
                for (int v = 0; v < 9; v++) {
                    double x = wave_sum(vir[v]);
                    if (lane == 0 && x != 0.0) atomic_add(p.virial + 9 * vframe + v, x);
                    vir[v] = 0.0;
                }
            }
            vframe = frame;
        }
        compact(p, g, i, lane, nb, nvs, p.fslot);
        __builtin_amdgcn_wave_barrier();
        // dA for this lane's features, with the 1/(m*type_nums) scale folded in
        float dAl[4][16];
        const float *dAi = p.dA + (size_t)i * 4 * E;
#pragma unroll
        for (int rr = 0; rr < 16; rr++) {
            const int f = feat(rr, h);
#pragma unroll
            for (int w = 0; w < 4; w++) dAl[w][rr] = (rr < NR && f < E) ? dAi[w * E + f] * p.scale : 0.0f;
        }
        float fc[3] = {0.f, 0.f, 0.f};     // centre-force accumulator (half 0 lanes)
        for (int t = 0; t < p.T; t++) {
            const int nv = nvs[t];
            const int pair = ti * p.T + t;
            const float *w1 = vec_ptr(p.vec, pair, V_W1, h);
            const float *r2v = vec_ptr(p.vec, pair, V_R2, h), *r3v = vec_ptr(p.vec, pair, V_R3, h);
            for (int k0 = 0; k0 < nv; k0 += TILE) {
                const int k = k0 + n;
                const bool live = k < nv;
                float4 e = live ? nb[t * p.mnl + k] : make_float4(1.f, 0.f, 0.f, 0.f);
                float Ri[4] = {0.f, 0.f, 0.f, 0.f};
                float r = 1.f, s = 0.f, ds = 0.f;
                if (live) ri_of(e, tinfo, p.rmin, p.rmax, Ri, r, s, ds);
                float h1[16], t2[16], t3[16], G[16];
                embed_fwd<E, BF16>(p, pair, lane, Ri[0], h1, t2, t3, G);
                // dG = dA^T Ri, dRi = dA G (partial over this half's features)
                float dG[16], dRi[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int rr = 0; rr < 16; rr++) {
                    dG[rr] = dAl[0][rr] * Ri[0] + dAl[1][rr] * Ri[1] + dAl[2][rr] * Ri[2] + dAl[3][rr] * Ri[3];
#pragma unroll
                    for (int w = 0; w < 4; w++) dRi[w] += dAl[w][rr] * G[rr];
                }
                // layer 3: G = h2 + r3 t3 -> dh2 = dG + W3 (dG r3 (1-t3^2))
                float dz[16];
                floatx16 acc;
#pragma unroll
                for (int rr = 0; rr < 16; rr++) {
                    dz[rr] = dG[rr] * r3v[rr] * (1.0f - t3[rr] * t3[rr]);
                    acc[rr] = dG[rr];
                }
                mfma_layer<BF16>(mat_ptr<BF16>(p.mat, pair, M_W3B), lane, dz, acc);
                // layer 2: h2 = h1 + r2 t2 -> dh1 = dh2 + W2 (dh2 r2 (1-t2^2))
#pragma unroll
                for (int rr = 0; rr < 16; rr++) dz[rr] = acc[rr] * r2v[rr] * (1.0f - t2[rr] * t2[rr]);
                mfma_layer<BF16>(mat_ptr<BF16>(p.mat, pair, M_W2B), lane, dz, acc);
                // layer 1: h1 = tanh(s w1 + b1) -> ds = sum_f dh1 (1-h1^2) w1
                float dsum = 0.0f;
#pragma unroll
                for (int rr = 0; rr < NR; rr++) dsum += acc[rr] * (1.0f - h1[rr] * h1[rr]) * w1[rr];
                dRi[0] += dsum;
#pragma unroll
                for (int w = 0; w < 4; w++) dRi[w] += __shfl_xor(dRi[w], 32, 64);
                // force of this slot: f_a = sum_w dE/dRi_w * Ri_d[w][a], Ri_d = -dfeat/dstd
                if (live && h == 0) {
                    const float ir = 1.0f / r;
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
                    const int64_t slot = (i * p.T + t) * (int64_t)p.mnl + __float_as_int(e.w);
                    if (p.fslot) {
                        // gather mode: the neighbor term is summed later through the reverse slot map
                        p.fslot[slot] = make_float4(f[0], f[1], f[2], 0.f);
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
                        for (int c = 0; c < 3; c++) vir[3 * a + c] += (double)(f[a] * x[c]);
                    }
                }
            }
        }
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float x = wave_sum(fc[a]);
            if (lane == 0) {
                if (p.fcenter) p.fcenter[3 * i + a] = x;
                else atomic_add(p.force + 3 * i + a, x);
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    // virial: wave sums -> one set of atomics per block and frame
    __shared__ double vred[FUSED_WAVES][9];
    __shared__ int vfr[FUSED_WAVES];
#pragma unroll
    for (int v = 0; v < 9; v++) {
        const double x = wave_sum(vir[v]);
        if (lane == 0) vred[wid][v] = x;
    }
    if (lane == 0) vfr[wid] = vframe;
    __syncthreads();
    if (threadIdx.x < 9) {
        const int v = threadIdx.x;
        double acc = 0.0;
        int fr = -1;
        for (int w = 0; w < FUSED_WAVES; w++) {
            if (vfr[w] < 0) continue;
            if (vfr[w] != fr) {
                if (fr >= 0 && acc != 0.0) atomic_add(p.virial + 9 * fr + v, acc);
                fr = vfr[w];
                acc = 0.0;
            }
            acc += vred[w][v];
        }
        if (fr >= 0 && acc != 0.0) atomic_add(p.virial + 9 * fr + v, acc);
    }
}

// force[i] = fcenter[i] + sum over used slots s of i of fslot[rev[s]]; one
// wavefront per atom, fixed reduction order (deterministic).
__global__ void __launch_bounds__(256)
gather_kernel(const int32_t *__restrict__ rev, const int32_t *__restrict__ counts, const float4 *__restrict__ fslot,
              const float *__restrict__ fcenter, float *__restrict__ force, int T, int mnl, int64_t n) {
    const int lane = threadIdx.x & (WAVE - 1);
    for (int64_t i = (int64_t)blockIdx.x * 4 + threadIdx.x / WAVE; i < n; i += (int64_t)gridDim.x * 4) {
        float a0 = 0.f, a1 = 0.f, a2 = 0.f;
        for (int t = 0; t < T; t++) {
            const int cnt = counts[i * T + t];
            const int32_t *rv = rev + (i * T + t) * (int64_t)mnl;
            for (int k = lane; k < cnt; k += WAVE) {
                const int r = rv[k];
                if (r >= 0) {
                    const float4 f = fslot[r];
                    a0 += f.x; a1 += f.y; a2 += f.z;
                }
            }
        }
        a0 = wave_sum(a0); a1 = wave_sum(a1); a2 = wave_sum(a2);
        if (lane == 0) {
            force[3 * i] = fcenter[3 * i] + a0;
            force[3 * i + 1] = fcenter[3 * i + 1] + a1;
            force[3 * i + 2] = fcenter[3 * i + 2] + a2;
        }
    }
}

template <int E, bool BF16>
int launch(FusedParams &p, bool fwd, hipStream_t st) {
    const size_t lds = sizeof(float4) * FUSED_WAVES * (size_t)p.T * p.mnl + sizeof(int) * FUSED_WAVES * MAXT;
    MATPL_CHECK_ARG(lds <= 64 * 1024, "dp_fused: T*mnl too large for LDS (%zu bytes)", lds);
    // persistent-style grid: a few blocks per CU, each strides over atoms
    int64_t blocks = (p.natoms + FUSED_WAVES - 1) / FUSED_WAVES;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    if (fwd) {
        hipLaunchKernelGGL((dp_fused_fwd_kernel<E, BF16>), dim3((unsigned)blocks), dim3(FUSED_WAVES * WAVE), lds, st, p);
        return check_launch("dp_fused_fwd_kernel");
    }
    hipLaunchKernelGGL((dp_fused_bwd_kernel<E, BF16>), dim3((unsigned)blocks), dim3(FUSED_WAVES * WAVE), lds, st, p);
    return check_launch("dp_fused_bwd_kernel");
}

int dispatch(FusedParams &p, int E, int bf16, bool fwd, hipStream_t st) {
    MATPL_CHECK_ARG(p.T >= 1 && p.T <= MAXT, "dp_fused: ntypes must be in [1,%d]", MAXT);
    MATPL_CHECK_ARG(p.mnl >= 1 && p.m >= 1, "dp_fused: bad bucket sizes");
    if (p.natoms == 0) return OK;
    MATPL_CHECK_ARG(p.pos && p.lattice && p.list && p.shift && p.counts && p.type_map && p.vec && p.mat && p.tinfo &&
                        p.status, "dp_fused: null pointer");
    if (E == 25) return bf16 ? launch<25, true>(p, fwd, st) : launch<25, false>(p, fwd, st);
    if (E == 32) return bf16 ? launch<32, true>(p, fwd, st) : launch<32, false>(p, fwd, st);
    if (E == 16) return bf16 ? launch<16, true>(p, fwd, st) : launch<16, false>(p, fwd, st);
    set_error("dp_fused: embedding width %d not compiled (16, 25, 32)", E);
    return ERR_ARGUMENT;
}

}  // namespace
}  // namespace matpl

using namespace matpl;

MATPL_EXPORT int matpl_dp_fused_forward(const double *pos, const double *lattice, const int32_t *atom_frame,
                                        const int64_t *frame_start, const int32_t *list, const int32_t *shift,
                                        const int32_t *counts, const int32_t *type_map, const float *vec,
                                        const void *mat, const float *tinfo, int64_t natoms, int ntypes, int mnl,
                                        int m, int E, int bf16, double rcut, double rcut_smooth, double scale,
                                        float *A, int32_t *status, void *stream) {
    FusedParams p{};
    p.pos = pos; p.lattice = lattice; p.atom_frame = atom_frame; p.frame_start = frame_start;
    p.list = list; p.shift = shift; p.counts = counts; p.type_map = type_map;
    p.vec = vec; p.mat = mat; p.tinfo = tinfo; p.A = A; p.status = status;
    p.natoms = natoms; p.T = ntypes; p.mnl = mnl; p.m = m;
    p.rc2 = rcut * rcut; p.rmin = (float)rcut_smooth; p.rmax = (float)rcut; p.scale = (float)scale;
    MATPL_CHECK_ARG(natoms == 0 || A, "dp_fused_forward: null A");
    return dispatch(p, E, bf16, true, (hipStream_t)stream);
}

MATPL_EXPORT int matpl_dp_fused_backward(const double *pos, const double *lattice, const int32_t *atom_frame,
                                         const int64_t *frame_start, const int32_t *list, const int32_t *shift,
                                         const int32_t *counts, const int32_t *type_map, const float *vec,
                                         const void *mat, const float *tinfo, int64_t natoms, int ntypes, int mnl,
                                         int m, int E, int bf16, double rcut, double rcut_smooth, double scale,
                                         const float *dA, const int32_t *rev, float *fslot, float *fcenter,
                                         float *force, double *virial, int32_t *status, void *stream) {
    FusedParams p{};
    p.pos = pos; p.lattice = lattice; p.atom_frame = atom_frame; p.frame_start = frame_start;
    p.list = list; p.shift = shift; p.counts = counts; p.type_map = type_map;
    p.vec = vec; p.mat = mat; p.tinfo = tinfo; p.dA = dA; p.force = force; p.virial = virial; p.status = status;
    p.natoms = natoms; p.T = ntypes; p.mnl = mnl; p.m = m;
    p.rc2 = rcut * rcut; p.rmin = (float)rcut_smooth; p.rmax = (float)rcut; p.scale = (float)scale;
    MATPL_CHECK_ARG(natoms == 0 || (dA && force && virial), "dp_fused_backward: null output");
    if (rev) {
        MATPL_CHECK_ARG(fslot && fcenter && aligned16(fslot), "dp_fused_backward: gather mode needs fslot/fcenter");
        p.fslot = reinterpret_cast<float4 *>(fslot);
        p.fcenter = fcenter;
    }
    int rc = dispatch(p, E, bf16, false, (hipStream_t)stream);
    if (rc || !rev || natoms == 0) return rc;
    hipLaunchKernelGGL(gather_kernel, dim3(grid_cap(natoms, 4)), dim3(256), 0, (hipStream_t)stream, rev, counts,
                       p.fslot, fcenter, force, ntypes, mnl, natoms);
    return check_launch("gather_kernel");
}
