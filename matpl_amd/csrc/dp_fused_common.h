// Device-side building blocks shared by the fused DP kernels (dp_fused.hip) and the
// training kernels (dp_train.hip): parameter block, weight layouts, the embedding net on
// the matrix cores, geometry and neighbor compaction, wave reductions. The layout
// conventions are described at the top of dp_fused.hip.
#pragma once
#include "common.h"

namespace matpl {
namespace {

constexpr int WAVE = 64;
constexpr int TILE = 32;
constexpr int FUSED_WAVES = 4;     // wavefronts per block (default; the net-mode kernels also build 2 and 8)
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
    const uint32_t *frame_types;// [F] bit t set when type t is in the frame's type list (nullptr: all T)
    const int32_t *list;        // [N][T][mnl] 1-based within frame, 0 = empty
    const int32_t *shift;       // [N][T][mnl] packed image
    const int32_t *counts;      // [N][T] entries used per bucket
    const int32_t *rev;         // [N][T][mnl] partner slot of each used entry (bwd gather mode)
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
    uint32_t *gbuf;             // store mode: [N][T][cap][2 halves][NRP] (G, dG/ds) as fp16 pairs
    int64_t natoms;
    int T, mnl, m;
    int cap;                    // LDS tile capacity per type = min(m, mnl)
    double rc2;                 // model cutoff^2 (FP64 test)
    float rmin, rmax, scale;    // scale = 1 / (m * T), used when frame_types is null
    float iw;                   // 1 / (rmax - rmin)
    int xcd;                    // 1: XCD-aware block order (grid a multiple of NXCD)
    // compressed embedding (table mode): G(s) by piecewise polynomials, K coefficients per
    // feature and interval, highest power first, laid out [T*T][rows][32 features][K], so that
    // the 32 lanes of a half-wave read one interval's row in one coalesced load
    const float *ctab;
    int crows, cnf;             // table rows; rows of the fine part
    float cx0, cdx, cidx, cxfe; // first knot, fine spacing, 1/dx, end of the fine part
};

constexpr int NXCD = 8;

// Blocks are dispatched to the XCDs round-robin (block b on XCD b % 8). With
// xcd set, block b takes logical position (b % 8) * (G / 8) + b / 8, so that
// each XCD, and its L2, works through one contiguous range of atoms.
__device__ __forceinline__ int64_t logical_block(int xcd) {
    const unsigned b = blockIdx.x, G = gridDim.x;
    return xcd ? (int64_t)(b % NXCD) * (G / NXCD) + b / NXCD : (int64_t)b;
}

// In BF16 mode the forward weights and biases of the embedding net come
// pre-multiplied by TANH_K = 2 log2(e) (matpl_amd/dp_fused.py), so a layer's
// pre-activation y = TANH_K z and tanh z = 1 - 2 / (2^y + 1): one hardware
// exp2 and one reciprocal, saturating correctly at +-inf.
constexpr float TANH_K = 2.8853900817779268f;

template <bool BF16>
__device__ __forceinline__ float act(float y) {
    if constexpr (BF16) return 1.0f - 2.0f * __builtin_amdgcn_rcpf(__builtin_amdgcn_exp2f(y) + 1.0f);
    else return tanhf(y);
}

// s(r) and ds/dr, dp_dp.py:244-249 (FP32); iw = 1 / (rmax - rmin)
__device__ __forceinline__ void switch_fn(float r, float rmin, float rmax, float iw, float &s, float &ds) {
    const float ir = __builtin_amdgcn_rcpf(r);
    if (r < rmin) {
        s = ir;
        ds = -s * s;
    } else if (r < rmax) {
        const float u = (r - rmin) * iw;
        const float vv = u * u * u * (-6.0f * u * u + 15.0f * u - 10.0f) + 1.0f;
        const float dvv = -30.0f * u * u * (u - 1.0f) * (u - 1.0f) * iw;
        s = vv * ir;
        ds = (dvv - s) * ir;
    } else {
        s = 0.0f;
        ds = 0.0f;
    }
}

// Weight blocks. Per type pair: NVEC vectors [2 halves][16] floats and NMAT
// matrices of 64 lanes x 16 elements (bf16 or f32), pre-permuted for the
// accumulator layout. When the weights of all pairs fit, each block stages
// them in LDS once; the matrices are then stored as [chunk][lane][16 bytes]
// so that every ds_read_b128 of a wavefront reads 64 consecutive 16-byte
// chunks (no bank conflicts), and the vectors are read as broadcasts (all
// lanes of a half read the same address).
template <bool BF16>
struct MatLayout {
    static constexpr int ES = BF16 ? 2 : 4;                 // element size
    static constexpr int CPL = 16 * ES / 16;                // 16-byte chunks per lane
    static constexpr int BLK = 64 * 16 * ES;                // bytes per matrix block
};

template <bool BF16, bool LW>
__device__ __forceinline__ void mfma_layer(const unsigned char *blk, int lane, const float x[16], floatx16 &acc) {
    if constexpr (BF16) {
        bf16x8 w0, w1;
        if constexpr (LW) {
            w0 = *reinterpret_cast<const bf16x8 *>(blk + lane * 16);
            w1 = *reinterpret_cast<const bf16x8 *>(blk + 1024 + lane * 16);
        } else {
            const bf16x8 *w = reinterpret_cast<const bf16x8 *>(blk) + lane * 2;
            w0 = w[0];
            w1 = w[1];
        }
        bf16x8 b0, b1;
#pragma unroll
        for (int q = 0; q < 8; q++) { b0[q] = (__bf16)x[q]; b1[q] = (__bf16)x[8 + q]; }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w0, b0, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w1, b1, acc, 0, 0, 0);
    } else {
        float w[16];
        if constexpr (LW) {
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const float4 v = *reinterpret_cast<const float4 *>(blk + (c * 64 + lane) * 16);
                w[4 * c] = v.x; w[4 * c + 1] = v.y; w[4 * c + 2] = v.z; w[4 * c + 3] = v.w;
            }
        } else {
            const float *wp = reinterpret_cast<const float *>(blk) + lane * 16;
#pragma unroll
            for (int r = 0; r < 16; r++) w[r] = wp[r];
        }
#pragma unroll
        for (int r = 0; r < 16; r++) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(w[r], x[r], acc, 0, 0, 0);
    }
}

template <bool BF16>
__device__ __forceinline__ const unsigned char *mat_ptr(const unsigned char *base, int pair, int which) {
    return base + ((size_t)pair * NMAT + which) * MatLayout<BF16>::BLK;
}

__device__ __forceinline__ const float *vec_ptr(const float *base, int pair, int which, int h) {
    return base + (((size_t)pair * NVEC + which) * 2 + h) * 16;
}

// Bytes of LDS the staged weights of T*T pairs take.
template <bool BF16>
__host__ __device__ inline size_t weight_lds_bytes(int T) {
    return (size_t)T * T * (NVEC * 32 * sizeof(float) + NMAT * MatLayout<BF16>::BLK);
}

// Copies every pair's weights from global memory into LDS (chunked matrix layout).
template <bool BF16>
__device__ void stage_weights(const FusedParams &p, float *svec, unsigned char *smat) {
    using ML = MatLayout<BF16>;
    const int npair = p.T * p.T;
    const int nv = npair * NVEC * 32;
    for (int x = threadIdx.x; x < nv; x += blockDim.x) svec[x] = p.vec[x];
    const int nchunk = npair * NMAT * 64 * ML::CPL;
    const uint4 *g = reinterpret_cast<const uint4 *>(p.mat);
    uint4 *l = reinterpret_cast<uint4 *>(smat);
    for (int u = threadIdx.x; u < nchunk; u += blockDim.x) {
        const int blk = u / (64 * ML::CPL), rem = u - blk * (64 * ML::CPL);
        const int ln = rem / ML::CPL, c = rem - ln * ML::CPL;
        l[blk * 64 * ML::CPL + c * 64 + ln] = g[u];
    }
    __syncthreads();
}

// Embedding forward for one tile. s = embedding input of this lane's neighbor.
// Out: h1 = tanh(layer1), t2, t3 = tanh of layers 2/3, G = output.
template <int E, bool BF16, bool LW>
__device__ __forceinline__ void embed_fwd(const float *vecb, const unsigned char *matb, int pair, int lane, float s,
                                          float h1[16], float t2[16], float t3[16], float G[16]) {
    constexpr int NR = NReg<E>::value;
    const int h = lane >> 5;
    const float *w1 = vec_ptr(vecb, pair, V_W1, h), *b1 = vec_ptr(vecb, pair, V_B1, h);
    const float *b2 = vec_ptr(vecb, pair, V_B2, h), *b3 = vec_ptr(vecb, pair, V_B3, h);
    const float *r2 = vec_ptr(vecb, pair, V_R2, h), *r3 = vec_ptr(vecb, pair, V_R3, h);
#pragma unroll
    for (int r = 0; r < 16; r++) h1[r] = (r < NR) ? act<BF16>(s * w1[r] + b1[r]) : 0.0f;
    floatx16 acc;
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = b2[r];
    mfma_layer<BF16, LW>(mat_ptr<BF16>(matb, pair, M_W2F), lane, h1, acc);
    float h2[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        t2[r] = (r < NR) ? act<BF16>(acc[r]) : 0.0f;
        h2[r] = (r < NR) ? h1[r] + r2[r] * t2[r] : 0.0f;     // padding rows stay literal zeros
    }
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = b3[r];
    mfma_layer<BF16, LW>(mat_ptr<BF16>(matb, pair, M_W3F), lane, h2, acc);
#pragma unroll
    for (int r = 0; r < 16; r++) {
        t3[r] = (r < NR) ? act<BF16>(acc[r]) : 0.0f;
        G[r] = (r < NR) ? h2[r] + r3[r] * t3[r] : 0.0f;
    }
}

// Embedding forward with its tangent in s: G and G' = dG/ds for this lane's features (forward mode
// through the net; the same layer rules and operation order as embed_fwd for G). Pre-activation
// tangents are true values: in BF16 mode the stored forward weights carry TANH_K.
template <int E, bool BF16, bool LW>
__device__ __forceinline__ void embed_fwd_tan(const float *vecb, const unsigned char *matb, int pair, int lane,
                                              float s, float G[16], float Gd[16]) {
    constexpr int NR = NReg<E>::value;
    constexpr float IK = BF16 ? 1.0f / TANH_K : 1.0f;
    const int h = lane >> 5;
    const float *w1 = vec_ptr(vecb, pair, V_W1, h), *b1 = vec_ptr(vecb, pair, V_B1, h);
    const float *b2 = vec_ptr(vecb, pair, V_B2, h), *b3 = vec_ptr(vecb, pair, V_B3, h);
    const float *r2 = vec_ptr(vecb, pair, V_R2, h), *r3 = vec_ptr(vecb, pair, V_R3, h);
    float h1[16], hd1[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        h1[r] = (r < NR) ? act<BF16>(s * w1[r] + b1[r]) : 0.0f;
        hd1[r] = (r < NR) ? (1.0f - h1[r] * h1[r]) * (w1[r] * IK) : 0.0f;
    }
    floatx16 acc, tac;
#pragma unroll
    for (int r = 0; r < 16; r++) { acc[r] = b2[r]; tac[r] = 0.0f; }
    mfma_layer<BF16, LW>(mat_ptr<BF16>(matb, pair, M_W2F), lane, h1, acc);
    mfma_layer<BF16, LW>(mat_ptr<BF16>(matb, pair, M_W2F), lane, hd1, tac);
    float h2[16], hd2[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const float t2 = (r < NR) ? act<BF16>(acc[r]) : 0.0f;
        h2[r] = (r < NR) ? h1[r] + r2[r] * t2 : 0.0f;
        hd2[r] = (r < NR) ? hd1[r] + r2[r] * ((1.0f - t2 * t2) * (tac[r] * IK)) : 0.0f;
    }
#pragma unroll
    for (int r = 0; r < 16; r++) { acc[r] = b3[r]; tac[r] = 0.0f; }
    mfma_layer<BF16, LW>(mat_ptr<BF16>(matb, pair, M_W3F), lane, h2, acc);
    mfma_layer<BF16, LW>(mat_ptr<BF16>(matb, pair, M_W3F), lane, hd2, tac);
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const float t3 = (r < NR) ? act<BF16>(acc[r]) : 0.0f;
        G[r] = (r < NR) ? h2[r] + r3[r] * t3 : 0.0f;
        Gd[r] = (r < NR) ? hd2[r] + r3[r] * ((1.0f - t3 * t3) * (tac[r] * IK)) : 0.0f;
    }
}

// Store mode: per compacted neighbor slot (atom i, bucket t, index k) and lane half h, the lane's
// registers r < NRP (NR rounded up to 4; registers r >= NR hold zeros) of G and G' = dG/ds: in BF16
// mode one dword per register, (G[r], G'[r]) as two fp16; in FP32 mode two dwords, G[r] and G'[r]
// as floats. A lane moves its slot's words in dwordx4 accesses.
template <int E, bool BF16>
struct GStore {
    static constexpr int NRP = (NReg<E>::value + 3) & ~3;
    static constexpr int WORDS = BF16 ? NRP : 2 * NRP;     // dwords per slot and half
};

__device__ __forceinline__ uint32_t pack_half2(float a, float b) {
    const _Float16 x = (_Float16)a, y = (_Float16)b;
    return (uint32_t)__builtin_bit_cast(uint16_t, x) | ((uint32_t)__builtin_bit_cast(uint16_t, y) << 16);
}

template <int E, bool BF16>
__device__ __forceinline__ size_t gslot(const FusedParams &p, int64_t i, int t, int k, int h) {
    return ((((size_t)i * p.T + t) * p.cap + k) * 2 + h) * GStore<E, BF16>::WORDS;
}

template <int E, bool BF16>
__device__ __forceinline__ void store_g(const FusedParams &p, int64_t i, int t, int k, int h, const float G[16],
                                        const float Gd[16]) {
    constexpr int NRP = GStore<E, BF16>::NRP;
    uint4 *o = reinterpret_cast<uint4 *>(p.gbuf + gslot<E, BF16>(p, i, t, k, h));
    if constexpr (BF16) {
#pragma unroll
        for (int q = 0; q < NRP / 4; q++)
            o[q] = make_uint4(pack_half2(G[4 * q], Gd[4 * q]), pack_half2(G[4 * q + 1], Gd[4 * q + 1]),
                              pack_half2(G[4 * q + 2], Gd[4 * q + 2]), pack_half2(G[4 * q + 3], Gd[4 * q + 3]));
    } else {
#pragma unroll
        for (int q = 0; q < NRP / 4; q++) {
            o[2 * q] = make_uint4(__float_as_uint(G[4 * q]), __float_as_uint(G[4 * q + 1]),
                                  __float_as_uint(G[4 * q + 2]), __float_as_uint(G[4 * q + 3]));
            o[2 * q + 1] = make_uint4(__float_as_uint(Gd[4 * q]), __float_as_uint(Gd[4 * q + 1]),
                                      __float_as_uint(Gd[4 * q + 2]), __float_as_uint(Gd[4 * q + 3]));
        }
    }
}

template <int E, bool BF16>
__device__ __forceinline__ void load_g(const FusedParams &p, int64_t i, int t, int k, int h, float G[16],
                                       float Gd[16]) {
    constexpr int NRP = GStore<E, BF16>::NRP;
    const uint4 *o = reinterpret_cast<const uint4 *>(p.gbuf + gslot<E, BF16>(p, i, t, k, h));
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if (q < NRP / 4) {
            if constexpr (BF16) {
                const uint4 v = o[q];
                const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    G[4 * q + u] = (float)__builtin_bit_cast(_Float16, (uint16_t)(w[u] & 0xffffu));
                    Gd[4 * q + u] = (float)__builtin_bit_cast(_Float16, (uint16_t)(w[u] >> 16));
                }
            } else {
                const uint4 a = o[2 * q], b = o[2 * q + 1];
                G[4 * q] = __uint_as_float(a.x); G[4 * q + 1] = __uint_as_float(a.y);
                G[4 * q + 2] = __uint_as_float(a.z); G[4 * q + 3] = __uint_as_float(a.w);
                Gd[4 * q] = __uint_as_float(b.x); Gd[4 * q + 1] = __uint_as_float(b.y);
                Gd[4 * q + 2] = __uint_as_float(b.z); Gd[4 * q + 3] = __uint_as_float(b.w);
            }
        } else {
#pragma unroll
            for (int u = 0; u < 4; u++) { G[4 * q + u] = 0.0f; Gd[4 * q + u] = 0.0f; }
        }
    }
}

// Compressed embedding: G = P_k(s - x_k) and dG/ds = P_k'(s - x_k) for the interval k holding s
// (intervals of dx up to the end of the fine part, then 10 dx; s outside the table extrapolates the
// first or last interval), as the reference's calc_compress_polynomial (dp_dp.py:572-601) with the
// table of ckpt_compress.py. table_row finds the interval: the float offset of its row and s - x_k.
template <int CK>
__device__ __forceinline__ void table_row(const FusedParams &p, int pair, float s, int &rowoff, float &x) {
    int row;
    float knot;
    if (s < p.cxfe) {
        row = (int)floorf((s - p.cx0) * p.cidx);
        row = row < 0 ? 0 : row;
        knot = p.cx0 + (float)row * p.cdx;
    } else {
        row = p.cnf + (int)floorf((s - p.cxfe) * (0.1f * p.cidx));
        row = row > p.crows - 1 ? p.crows - 1 : row;
        knot = p.cxfe + (float)(row - p.cnf) * (10.0f * p.cdx);
    }
    x = s - knot;
    rowoff = (pair * p.crows + row) * (32 * CK);
}

// Value and derivative of one feature's polynomial (coefficients c[0..CK), highest power first) at
// x, by one Horner pass.
template <int CK>
__device__ __forceinline__ void horner(const float *c, float x, float &v, float &d) {
    float cc[CK];
    if constexpr (CK == 4) {
        const float4 q = *reinterpret_cast<const float4 *>(c);
        cc[0] = q.x; cc[1] = q.y; cc[2] = q.z; cc[3] = q.w;
    } else {
#pragma unroll
        for (int u = 0; u < CK / 2; u++) {
            const float2 q = reinterpret_cast<const float2 *>(c)[u];
            cc[2 * u] = q.x;
            cc[2 * u + 1] = q.y;
        }
    }
    v = cc[0];
    d = 0.0f;
#pragma unroll
    for (int kk = 1; kk < CK; kk++) {
        d = d * x + v;
        v = v * x + cc[kk];
    }
}

// The features this lane holds in the accumulator layout (register r: feature f(r, h)).
template <int E, int CK>
__device__ __forceinline__ void table_eval(const FusedParams &p, int pair, float s, int h, float G[16],
                                           float Gd[16]) {
    constexpr int NR = NReg<E>::value;
    int rowoff;
    float x;
    table_row<CK>(p, pair, s, rowoff, x);
#pragma unroll
    for (int r = 0; r < 16; r++) {
        G[r] = 0.0f;
        Gd[r] = 0.0f;
        if (r < NR) horner<CK>(p.ctab + rowoff + feat(r, h) * CK, x, G[r], Gd[r]);
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

// Compacts the in-cutoff neighbors of atom i into LDS: nb[t*cap + c] =
// (dx, dy, dz, bucket slot k as float bits); nv[t] = count. Used bucket
// entries beyond the cutoff get a zero per-slot force when fslot is given.
// Returns false on overflow.
//
// The used entries of all buckets are taken as one sequence (bucket t's
// entries at offsets off[t]..off[t+1]), two per lane per pass, and every load
// of a pass (list, shift, then the neighbors' positions) is issued before any
// is used: a skin list of ~105 entries in two buckets takes one pass of two
// dependent load rounds instead of three bucket-by-bucket passes of two each.
// Within a bucket the entries keep their order, so nb is the same as a
// bucket-by-bucket compaction.
__device__ __forceinline__ bool compact(const FusedParams &p, const AtomGeo &g, int64_t i, int lane, float4 *nb,
                                        int *nv, float4 *fslot) {
    int off[MAXT + 1], n[MAXT];
    off[0] = 0;
#pragma unroll
    for (int t = 0; t < MAXT; t++) {
        off[t + 1] = off[t] + (t < p.T ? p.counts[i * p.T + t] : 0);
        n[t] = 0;
    }
    const int total = off[MAXT];
    const uint64_t below = (1ull << lane) - 1ull;
    for (int e0 = 0; e0 < total; e0 += 2 * WAVE) {
        int tt[2];
        int64_t slot[2];
        double dx[2], dy[2], dz[2];
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int e = e0 + u * WAVE + lane;
            int t = 0;
#pragma unroll
            for (int q = 1; q < MAXT; q++) t += (e >= off[q]) ? 1 : 0;
            tt[u] = e < total ? t : -1;
            slot[u] = (i * p.T + t) * (int64_t)p.mnl + (e - off[t]);
            dx[u] = dy[u] = dz[u] = 0.0;
            if (e < total) neighbor_disp(p, g, slot[u], dx[u], dy[u], dz[u]);
        }
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const double r2 = dx[u] * dx[u] + dy[u] * dy[u] + dz[u] * dz[u];
            const bool ok = tt[u] >= 0 && r2 < p.rc2 && r2 > 1e-10;
            const int k = (int)(slot[u] - (i * p.T + (tt[u] < 0 ? 0 : tt[u])) * (int64_t)p.mnl);
            if (tt[u] >= 0 && !ok && fslot) {
                const int r = p.rev[slot[u]];    // -1 if the list is not symmetric (reported by its status)
                if (r >= 0) fslot[r] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
            const float4 e = make_float4((float)dx[u], (float)dy[u], (float)dz[u], __int_as_float(k));
#pragma unroll
            for (int t = 0; t < MAXT; t++) {
                if (t < p.T) {
                    const bool mine = ok && tt[u] == t;
                    const uint64_t bal = __ballot(mine);
                    const int idx = n[t] + __popcll(bal & below);
                    if (mine && idx < p.cap) nb[t * p.cap + idx] = e;
                    n[t] += __popcll(bal);
                }
            }
        }
    }
    bool ok_all = true;
#pragma unroll
    for (int t = 0; t < MAXT; t++) {
        if (t < p.T) {
            // more than m in-cutoff neighbors is an error the caller reports; only cap of them are used
            if (lane == 0) nv[t] = n[t] < p.cap ? n[t] : p.cap;
            if (n[t] > p.m) ok_all = false;
        }
    }
    return ok_all;
}

// The type list of atom i's frame: the mask of the neighbor types whose buckets take part, and the
// descriptor scale 1 / (m * type_nums) of dp_dp.py:413 (type_nums = types in the list).
__device__ __forceinline__ uint32_t frame_type_mask(const FusedParams &p, int64_t i, float &scale) {
    if (!p.frame_types) {
        scale = p.scale;
        return 0xffffffffu;
    }
    const uint32_t mk = p.frame_types[p.atom_frame ? p.atom_frame[i] : 0] & ((1u << p.T) - 1u);
    scale = 1.0f / (float)(p.m * __popc(mk));
    return mk;
}

// Normalized Ri of a compacted neighbor (FP32).
__device__ __forceinline__ void ri_of(const float4 e, const float *ti, float rmin, float rmax, float iw, float Ri[4],
                                      float &r, float &s, float &ds) {
    r = sqrtf(e.x * e.x + e.y * e.y + e.z * e.z);
    switch_fn(r, rmin, rmax, iw, s, ds);
    const float ir = __builtin_amdgcn_rcpf(r);
    Ri[0] = (s - ti[TI_DAVG]) * ti[TI_ISTD];
    Ri[1] = (s * e.x * ir - ti[TI_DAVG + 1]) * ti[TI_ISTD + 1];
    Ri[2] = (s * e.y * ir - ti[TI_DAVG + 2]) * ti[TI_ISTD + 2];
    Ri[3] = (s * e.z * ir - ti[TI_DAVG + 3]) * ti[TI_ISTD + 3];
}

// Sum over the 32 lanes of each half-wave, without LDS: DPP sums within each
// row of 16 lanes (quad_perm, row_half_mirror, row_mirror), row_bcast:15 adds
// row 0 into row 1 and row 2 into row 3, and the two half totals (lanes 31
// and 63) are read back as scalars.
__device__ __forceinline__ float half_sum(float v) {
    v += __builtin_amdgcn_update_dpp(0.0f, v, 0xB1, 0xF, 0xF, false);    // quad_perm [1,0,3,2]
    v += __builtin_amdgcn_update_dpp(0.0f, v, 0x4E, 0xF, 0xF, false);    // quad_perm [2,3,0,1]
    v += __builtin_amdgcn_update_dpp(0.0f, v, 0x141, 0xF, 0xF, false);   // row_half_mirror
    v += __builtin_amdgcn_update_dpp(0.0f, v, 0x140, 0xF, 0xF, false);   // row_mirror
    v += __builtin_amdgcn_update_dpp(0.0f, v, 0x142, 0xA, 0xF, false);   // row_bcast:15 -> rows 1, 3
    const float s0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 31));
    const float s1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
    return (threadIdx.x & 32) ? s1 : s0;
}

// Value of lane l ^ 32 (the same lane of the other half-wave), via v_permlane32_swap.
__device__ __forceinline__ float other_half(float v) {
    const unsigned u = __float_as_uint(v);
    const auto r = __builtin_amdgcn_permlane32_swap(u, u, false, false);
    return __uint_as_float((threadIdx.x & 32) ? r[0] : r[1]);
}

// x and y summed over the two half-waves, valid on lanes 0-31 (lanes 32-63: unspecified). One
// v_permlane32_swap exchanges x's upper half with y's lower half, so that one add gives the x sum
// on lanes 0-31 and the y sum on lanes 32-63; a second swap brings the y sum down. Four ops for two
// values where other_half takes five for one.
__device__ __forceinline__ void fold_halves2(float &x, float &y) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(y), false, false);
    const float s = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    const auto q = __builtin_amdgcn_permlane32_swap(__float_as_uint(s), __float_as_uint(s), false, false);
    x = s;
    y = __uint_as_float(q[1]);
}

// Sum over all 64 lanes, without LDS.
__device__ __forceinline__ float wave_sum_dpp(float v) {
    v = half_sum(v);
    return v + other_half(v);
}

// The three tanh-layer gates r (1 - t^2) of a tile, kept from the forward
// recompute until the backward uses them: FP32, or packed bf16 pairs (half
// the registers; measured slower on gfx950 at no register saving for E = 25,
// so the kernels use FP32).
template <bool PACK>
struct Gate16;
template <>
struct Gate16<true> {
    uint32_t v[8];
    __device__ __forceinline__ void set(const float g[16]) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const __bf16 a = (__bf16)g[2 * j], b = (__bf16)g[2 * j + 1];
            v[j] = (uint32_t)__builtin_bit_cast(uint16_t, a) | ((uint32_t)__builtin_bit_cast(uint16_t, b) << 16);
        }
    }
    __device__ __forceinline__ float get(int r) const {
        const uint32_t w = v[r >> 1];
        return (r & 1) ? __uint_as_float(w & 0xffff0000u) : __uint_as_float(w << 16);
    }
};
template <>
struct Gate16<false> {
    float v[16];
    __device__ __forceinline__ void set(const float g[16]) {
#pragma unroll
        for (int r = 0; r < 16; r++) v[r] = g[r];
    }
    __device__ __forceinline__ float get(int r) const { return v[r]; }
};

// LDS layout of a fused block: [staged weights (LW only)][neighbor tiles of
// each wave: T*mnl float4][per-wave bucket counts: MAXT ints][bwd only: per
// wave dA of the current atom [32 features][4] floats, virial 9 doubles]
// [BF16 fwd only: per wave the tile's Ri^T [4][32] and G^T [32][32 or 40] in bf16].
// (table mode: G^T rows of 40 instead of 32, which spreads the coalesced evaluation's writes over
// the LDS banks, and 32 (row offset, x) pairs)
constexpr int XT_ELEMS = 4 * 32 + 32 * 40 + 2 * 64;
struct BlockLds {
    const float *vec;
    const unsigned char *mat;
    float4 *nb;
    int *nvs;
    float4 *dAs;
    double *vacc;
    __bf16 *xt;
};

template <bool BF16, bool LW, bool BWD, int W = FUSED_WAVES>
__device__ __forceinline__ BlockLds block_lds(const FusedParams &p, float4 *smem, int wid) {
    unsigned char *base = reinterpret_cast<unsigned char *>(smem);
    BlockLds b;
    size_t off = 0;
    if constexpr (LW) {
        float *sv = reinterpret_cast<float *>(base);
        unsigned char *sm = base + (size_t)p.T * p.T * NVEC * 32 * sizeof(float);
        stage_weights<BF16>(p, sv, sm);
        b.vec = sv;
        b.mat = sm;
        off = weight_lds_bytes<BF16>(p.T);
    } else {
        b.vec = p.vec;
        b.mat = reinterpret_cast<const unsigned char *>(p.mat);
    }
    float4 *nball = reinterpret_cast<float4 *>(base + off);
    b.nb = nball + (size_t)wid * p.T * p.cap;
    int *nvall = reinterpret_cast<int *>(nball + (size_t)W * p.T * p.cap);
    b.nvs = nvall + wid * MAXT;
    float4 *dAall = reinterpret_cast<float4 *>(nvall + W * MAXT);
    b.dAs = dAall + wid * 32;
    b.vacc = reinterpret_cast<double *>(dAall + W * 32) + wid * 9;
    // forward: the transposed tile sits where the backward keeps dA (16-byte aligned)
    b.xt = reinterpret_cast<__bf16 *>(dAall) + (size_t)wid * XT_ELEMS;
    (void)BWD;
    return b;
}

template <bool BF16, bool LW, bool BWD>
inline size_t fused_lds_bytes(int T, int cap, int W = FUSED_WAVES) {
    size_t b = (LW ? weight_lds_bytes<BF16>(T) : 0) + sizeof(float4) * W * (size_t)T * cap + sizeof(int) * W * MAXT;
    if (BWD) b += sizeof(float4) * W * 32 + sizeof(double) * W * 9;
    if (!BWD && BF16) b += sizeof(__bf16) * W * XT_ELEMS;
    return b;
}

}  // namespace
}  // namespace matpl
