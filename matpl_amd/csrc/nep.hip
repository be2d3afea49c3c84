// NEP descriptor path on gfx950: neighbor search, radial and angular
// (3/4/5-body) features, their first and second derivatives, and the NEP
// force / virial scatter. FP64 and FP32 instantiations.
//
// Semantics (the reference's src/op contract, with the maths restated):
//   neighbor  : calculateNepNeighbor.cu:166-281 / cpu_calculate_nepneighbor.cpp:97-203
//               every atom of a frame against every atom of the frame in every
//               replica (ia,ib,ic) < num_cell, self excluded in replica 0,
//               minimum image in the expanded box; lists ordered (n2, ia, ib, ic),
//               global indices, -1 padded; Ri = [r, x, y, z] (or [x, y, z])
//   f_k(r)    : (T_k(x)+1)/2 fc(r), x = 2(r/rc-1)^2-1, fc = (1+cos(pi r/rc))/2 (r < rc)
//   radial    : q_n = sum_j sum_k c2[ti,tj,n,k] f_k(r_ij)         (calculateNepFeat.cu:5-96)
//   angular   : s_nlm = sum_j g_n(r_ij) b_lm(r_ij/r), g_n = sum_k c3[ti,tj,n,k] f_k
//               q = C3B-weighted s^2 per L, 4-body, 5-body         (nep3_small_box.cuh:68-146)
// Derivatives are taken with r and (x, y, z) as independent inputs, the
// reference's convention for the [N, S, 4] gradient of the descriptor ops.
// Neighbor loops stop at the first -1 entry as the reference's do.
//
// Layout: one wavefront per atom, lanes over neighbor slots; per-atom sums
// over slots (s_nlm, their tangents) go through LDS atomics; per-type
// coefficient gradients are accumulated per block in LDS and added to global
// memory once per block (persistent grid).
#include "common.h"

#pragma clang fp contract(off)

namespace matpl {
namespace {

constexpr int WAVE = 64;
constexpr int NEP_WAVES = 4;
constexpr int NB_MAX = 20;      // basis functions
constexpr int NM_MAX = 8;       // radial functions per descriptor part
constexpr int NLM = 24;
constexpr double PI_D = 3.14159265358979323846;

template <typename T>
__device__ __forceinline__ void fk_fkp(T r, T rc, int nb, T *f, T *fp) {
    const T rcinv = T(1) / rc;
    T fc = T(0), fcp = T(0);
    if (r < rc) {
        const T a = T(PI_D) * r * rcinv;
        fc = T(0.5) * cos(a) + T(0.5);
        fcp = -T(0.5 * PI_D) * sin(a) * rcinv;
    }
    const T d = r * rcinv - T(1);
    const T x = T(2) * d * d - T(1);
    const T xp = T(4) * d * rcinv;
    T t0 = T(1), t1 = x, tp0 = T(0), tp1 = T(1);
    for (int k = 0; k < nb; k++) {
        T t, tp;
        if (k == 0) { t = t0; tp = tp0; }
        else if (k == 1) { t = t1; tp = tp1; }
        else {
            t = T(2) * x * t1 - t0;
            tp = T(2) * t1 + T(2) * x * tp1 - tp0;
            t0 = t1; t1 = t; tp0 = tp1; tp1 = tp;
        }
        f[k] = (t + T(1)) * T(0.5) * fc;
        fp[k] = T(0.5) * (tp * xp * fc + (t + T(1)) * fcp);
    }
}

// 24 real harmonics of u = (x,y,z)/r (accumulate_s order) and their partial
// derivatives with respect to (r, x, y, z) taken independently.
template <typename T>
__device__ __forceinline__ void blm_and_grad(T r, T X, T Y, T Z, T b[NLM], T db[NLM][4]) {
    const T ir = T(1) / r;
    const T x = X * ir, y = Y * ir, z = Z * ir;
    const T x2 = x * x, y2 = y * y, z2 = z * z, xy2 = x2 - y2;
    // value and d/du_(x,y,z)
    T du[NLM][3];
#define SET(k, v, a, bb, c) { b[k] = (v); du[k][0] = (a); du[k][1] = (bb); du[k][2] = (c); }
    SET(0, z, T(0), T(0), T(1));
    SET(1, x, T(1), T(0), T(0));
    SET(2, y, T(0), T(1), T(0));
    SET(3, T(3) * z2 - T(1), T(0), T(0), T(6) * z);
    SET(4, x * z, z, T(0), x);
    SET(5, y * z, T(0), z, y);
    SET(6, xy2, T(2) * x, -T(2) * y, T(0));
    SET(7, T(2) * x * y, T(2) * y, T(2) * x, T(0));
    SET(8, (T(5) * z2 - T(3)) * z, T(0), T(0), T(15) * z2 - T(3));
    SET(9, (T(5) * z2 - T(1)) * x, T(5) * z2 - T(1), T(0), T(10) * x * z);
    SET(10, (T(5) * z2 - T(1)) * y, T(0), T(5) * z2 - T(1), T(10) * y * z);
    SET(11, xy2 * z, T(2) * x * z, -T(2) * y * z, xy2);
    SET(12, T(2) * x * y * z, T(2) * y * z, T(2) * x * z, T(2) * x * y);
    SET(13, (x2 - T(3) * y2) * x, T(3) * x2 - T(3) * y2, -T(6) * x * y, T(0));
    SET(14, (T(3) * x2 - y2) * y, T(6) * x * y, T(3) * x2 - T(3) * y2, T(0));
    SET(15, (T(35) * z2 - T(30)) * z2 + T(3), T(0), T(0), T(140) * z2 * z - T(60) * z);
    SET(16, (T(7) * z2 - T(3)) * x * z, (T(7) * z2 - T(3)) * z, T(0), T(21) * x * z2 - T(3) * x);
    SET(17, (T(7) * z2 - T(3)) * y * z, T(0), (T(7) * z2 - T(3)) * z, T(21) * y * z2 - T(3) * y);
    SET(18, (T(7) * z2 - T(1)) * xy2, T(2) * x * (T(7) * z2 - T(1)), -T(2) * y * (T(7) * z2 - T(1)),
        T(14) * z * xy2);
    SET(19, (T(7) * z2 - T(1)) * x * y * T(2), T(2) * y * (T(7) * z2 - T(1)), T(2) * x * (T(7) * z2 - T(1)),
        T(28) * x * y * z);
    SET(20, (x2 - T(3) * y2) * x * z, (T(3) * x2 - T(3) * y2) * z, -T(6) * x * y * z, (x2 - T(3) * y2) * x);
    SET(21, (T(3) * x2 - y2) * y * z, T(6) * x * y * z, (T(3) * x2 - T(3) * y2) * z, (T(3) * x2 - y2) * y);
    SET(22, xy2 * xy2 - T(4) * x2 * y2, T(4) * x2 * x - T(12) * x * y2, T(4) * y2 * y - T(12) * x2 * y, T(0));
    SET(23, T(4) * x * y * xy2, T(12) * x2 * y - T(4) * y2 * y, T(4) * x2 * x - T(12) * x * y2, T(0));
#undef SET
#pragma unroll
    for (int k = 0; k < NLM; k++) {
        db[k][1] = du[k][0] * ir;
        db[k][2] = du[k][1] * ir;
        db[k][3] = du[k][2] * ir;
        db[k][0] = -(x * du[k][0] + y * du[k][1] + z * du[k][2]) * ir;
    }
}

// ------------------------------------------------------------ q(s) and its gradient
// Weighted C3B (factor 2 on m != 0, nep_cpu.cpp:93-107), C4B, C5B (nep_cpu.h:8-16).
__constant__ double kC3B[NLM] = {
    0.238732414637843, 0.238732414637844, 0.238732414637844, 0.099471839432435, 1.193662073189216,
    1.193662073189216, 0.298415518297304, 0.298415518297304, 0.139260575205408, 0.208890862808112,
    0.208890862808112, 2.088908628081126, 2.088908628081126, 0.348151438013522, 0.348151438013522,
    0.011190581936149, 0.447623277445956, 0.447623277445956, 0.223811638722978, 0.223811638722978,
    3.133362942121690, 3.133362942121690, 0.391670367765212, 0.391670367765212};
__constant__ double kC4B[5] = {-0.007499480826664, -0.134990654879954, 0.067495327439977, 0.404971964639861,
                               -0.809943929279723};
__constant__ double kC5B[3] = {0.026596810706114, 0.053193621412227, 0.026596810706114};
__constant__ int kLstart[5] = {0, 3, 8, 15, 24};

// Forward-mode dual number (value, tangent) for Hessian-vector products.
template <typename T>
struct Dual {
    T v, e;
    __device__ Dual() {}
    __device__ Dual(T a) : v(a), e(T(0)) {}
    __device__ Dual(T a, T b) : v(a), e(b) {}
};
template <typename T> __device__ __forceinline__ Dual<T> operator+(Dual<T> a, Dual<T> b) { return {a.v + b.v, a.e + b.e}; }
template <typename T> __device__ __forceinline__ Dual<T> operator-(Dual<T> a, Dual<T> b) { return {a.v - b.v, a.e - b.e}; }
template <typename T> __device__ __forceinline__ Dual<T> operator*(Dual<T> a, Dual<T> b) { return {a.v * b.v, a.v * b.e + a.e * b.v}; }
template <typename T> __device__ __forceinline__ Dual<T> operator*(T a, Dual<T> b) { return {a * b.v, a * b.e}; }

// q for one radial index: q3[l] (l < l3), then 4-body, then 5-body (nq outputs).
template <typename T, typename S>
__device__ __forceinline__ int q_of_s(const S s[NLM], int l3, int l4, int l5, S q[6]) {
    int nq = 0;
    for (int l = 0; l < l3; l++) {
        S acc(T(0));
        for (int m = kLstart[l]; m < kLstart[l + 1]; m++) acc = acc + T(kC3B[m]) * (s[m] * s[m]);
        q[nq++] = acc;
    }
    if (l4 > 0) {
        q[nq++] = T(kC4B[0]) * (s[3] * s[3] * s[3]) + T(kC4B[1]) * (s[3] * (s[4] * s[4] + s[5] * s[5])) +
                  T(kC4B[2]) * (s[3] * (s[6] * s[6] + s[7] * s[7])) + T(kC4B[3]) * (s[6] * (s[5] * s[5] - s[4] * s[4])) +
                  T(kC4B[4]) * (s[4] * s[5] * s[7]);
    }
    if (l5 > 0) {
        const S s0 = s[0] * s[0], s12 = s[1] * s[1] + s[2] * s[2];
        q[nq++] = T(kC5B[0]) * (s0 * s0) + T(kC5B[1]) * (s0 * s12) + T(kC5B[2]) * (s12 * s12);
    }
    return nq;
}

// D[m] = sum_F gq[F] dq_F/ds_m (gq in q_of_s order).
template <typename T, typename S>
__device__ __forceinline__ void dq_ds(const S s[NLM], const T *gq, int l3, int l4, int l5, S D[NLM]) {
    for (int m = 0; m < NLM; m++) D[m] = S(T(0));
    int f = 0;
    for (int l = 0; l < l3; l++, f++)
        for (int m = kLstart[l]; m < kLstart[l + 1]; m++) D[m] = D[m] + (T(2) * T(kC3B[m]) * gq[f]) * s[m];
    if (l4 > 0) {
        const T g = gq[f++];
        const T c0 = T(kC4B[0]), c1 = T(kC4B[1]), c2 = T(kC4B[2]), c3 = T(kC4B[3]), c4 = T(kC4B[4]);
        D[3] = D[3] + g * (T(3) * c0 * (s[3] * s[3]) + c1 * (s[4] * s[4] + s[5] * s[5]) + c2 * (s[6] * s[6] + s[7] * s[7]));
        D[4] = D[4] + g * (T(2) * c1 * (s[3] * s[4]) - T(2) * c3 * (s[6] * s[4]) + c4 * (s[5] * s[7]));
        D[5] = D[5] + g * (T(2) * c1 * (s[3] * s[5]) + T(2) * c3 * (s[6] * s[5]) + c4 * (s[4] * s[7]));
        D[6] = D[6] + g * (T(2) * c2 * (s[3] * s[6]) + c3 * (s[5] * s[5] - s[4] * s[4]));
        D[7] = D[7] + g * (T(2) * c2 * (s[3] * s[7]) + c4 * (s[4] * s[5]));
    }
    if (l5 > 0) {
        const T g = gq[f++];
        const T c0 = T(kC5B[0]), c1 = T(kC5B[1]), c2 = T(kC5B[2]);
        const S s12 = s[1] * s[1] + s[2] * s[2];
        D[0] = D[0] + g * (T(4) * c0 * (s[0] * s[0] * s[0]) + T(2) * c1 * (s[0] * s12));
        D[1] = D[1] + g * (T(2) * c1 * (s[0] * s[0] * s[1]) + T(4) * c2 * (s12 * s[1]));
        D[2] = D[2] + g * (T(2) * c1 * (s[0] * s[0] * s[2]) + T(4) * c2 * (s12 * s[2]));
    }
}

// position of feature (group f, radial index n) in the angular feature vector
__device__ __forceinline__ int aidx(int f, int n, int na) { return f * na + n; }

// ------------------------------------------------------------------ neighbor search
struct NepGeo {
    const double *box;       // [F][18] expanded box + inverse (column-major)
    const double *box_orig;  // [F][9]
    const int64_t *num_cell; // [F][3]
};

__device__ __forceinline__ void apply_mic(const double *h, double &x, double &y, double &z) {
    double sx = h[9] * x + h[10] * y + h[11] * z;
    double sy = h[12] * x + h[13] * y + h[14] * z;
    double sz = h[15] * x + h[16] * y + h[17] * z;
    sx -= nearbyint(sx);
    sy -= nearbyint(sy);
    sz -= nearbyint(sz);
    x = h[0] * sx + h[1] * sy + h[2] * sz;
    y = h[3] * sx + h[4] * sy + h[5] * sz;
    z = h[6] * sx + h[7] * sy + h[8] * sz;
}

__device__ __forceinline__ int wave_incl_scan(int v) {
    const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        int u = __shfl_up(v, off, WAVE);
        if (lane >= off) v += u;
    }
    return v;
}

// mode 0: lists; 1: counts; 2: per-type counts. One wavefront per atom.
__global__ void __launch_bounds__(256)
nep_neighbor_kernel(NepGeo g, const int64_t *__restrict__ frame_start, const int32_t *__restrict__ atom_frame,
                    const int64_t *__restrict__ type_map, const double *__restrict__ pos, double rc_r, double rc_a,
                    int64_t maxr, int64_t maxa, int64_t n, int mode, int ntypes, int rij4, int64_t *__restrict__ NN_r,
                    int64_t *__restrict__ NL_r, int64_t *__restrict__ NN_a, int64_t *__restrict__ NL_a,
                    double *__restrict__ Ri_r, double *__restrict__ Ri_a, int32_t *__restrict__ status) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t n1 = blockIdx.x * (int64_t)(blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (n1 >= n) return;
    const int f = atom_frame[n1];
    const int64_t N1 = frame_start[f], N2 = frame_start[f + 1];
    const double *h = g.box + 18 * (int64_t)f, *bo = g.box_orig + 9 * (int64_t)f;
    const int64_t *nc = g.num_cell + 3 * (int64_t)f;
    const int64_t nrep = nc[0] * nc[1] * nc[2];
    const double x1 = pos[3 * n1], y1 = pos[3 * n1 + 1], z1 = pos[3 * n1 + 2];
    const double rr2 = rc_r * rc_r, ra2 = rc_a * rc_a;
    const int w = rij4 ? 4 : 3;
    int64_t baser = 0, basea = 0;
    int64_t tr[16], ta[16];
    for (int t = 0; t < 16; t++) tr[t] = ta[t] = 0;
    bool overflow = false;
    for (int64_t c0 = N1; c0 < N2; c0 += WAVE) {
        const int64_t n2 = c0 + lane;
        const bool valid = n2 < N2;
        int cr = 0, ca = 0;
        // type_map is only given (and only read) for the per-type counts of mode 2
        const int t2 = (valid && mode == 2) ? (int)type_map[n2] : 0;
        // pass 1: counts
        if (valid) {
            for (int64_t rep = 0; rep < nrep; rep++) {
                const int64_t ia = rep / (nc[1] * nc[2]), ib = (rep / nc[2]) % nc[1], ic = rep % nc[2];
                if (ia == 0 && ib == 0 && ic == 0 && n1 == n2) continue;
                const double dx = bo[0] * (double)ia + bo[1] * (double)ib + bo[2] * (double)ic;
                const double dy = bo[3] * (double)ia + bo[4] * (double)ib + bo[5] * (double)ic;
                const double dz = bo[6] * (double)ia + bo[7] * (double)ib + bo[8] * (double)ic;
                double x12 = pos[3 * n2] + dx - x1, y12 = pos[3 * n2 + 1] + dy - y1, z12 = pos[3 * n2 + 2] + dz - z1;
                apply_mic(h, x12, y12, z12);
                const double d2 = x12 * x12 + y12 * y12 + z12 * z12;
                if (d2 < rr2) cr++;
                if (d2 < ra2) ca++;
            }
        }
        if (mode == 2) {
            for (int t = 0; t < ntypes; t++) {
                const int vr = (valid && t2 == t) ? cr : 0, va = (valid && t2 == t) ? ca : 0;
                const int sr = wave_incl_scan(vr), sa = wave_incl_scan(va);
                tr[t] += __shfl(sr, WAVE - 1, WAVE);
                ta[t] += __shfl(sa, WAVE - 1, WAVE);
            }
            continue;
        }
        const int ir = wave_incl_scan(cr), ia_ = wave_incl_scan(ca);
        const int64_t offr = baser + ir - cr, offa = basea + ia_ - ca;
        baser += __shfl(ir, WAVE - 1, WAVE);
        basea += __shfl(ia_, WAVE - 1, WAVE);
        if (mode == 1) continue;
        if (baser > maxr || basea > maxa) { overflow = true; continue; }
        // pass 2: write in (n2, replica) order
        if (valid && (cr > 0 || ca > 0)) {
            int64_t kr = offr, ka = offa;
            for (int64_t rep = 0; rep < nrep; rep++) {
                const int64_t ia = rep / (nc[1] * nc[2]), ib = (rep / nc[2]) % nc[1], ic = rep % nc[2];
                if (ia == 0 && ib == 0 && ic == 0 && n1 == n2) continue;
                const double dx = bo[0] * (double)ia + bo[1] * (double)ib + bo[2] * (double)ic;
                const double dy = bo[3] * (double)ia + bo[4] * (double)ib + bo[5] * (double)ic;
                const double dz = bo[6] * (double)ia + bo[7] * (double)ib + bo[8] * (double)ic;
                double x12 = pos[3 * n2] + dx - x1, y12 = pos[3 * n2 + 1] + dy - y1, z12 = pos[3 * n2 + 2] + dz - z1;
                apply_mic(h, x12, y12, z12);
                const double d2 = x12 * x12 + y12 * y12 + z12 * z12;
                if (d2 < rr2) {
                    NL_r[n1 * maxr + kr] = n2;
                    double *o = Ri_r + (n1 * maxr + kr) * w;
                    if (rij4) { o[0] = sqrt(d2); o[1] = x12; o[2] = y12; o[3] = z12; }
                    else { o[0] = x12; o[1] = y12; o[2] = z12; }
                    kr++;
                }
                if (d2 < ra2) {
                    NL_a[n1 * maxa + ka] = n2;
                    double *o = Ri_a + (n1 * maxa + ka) * w;
                    if (rij4) { o[0] = sqrt(d2); o[1] = x12; o[2] = y12; o[3] = z12; }
                    else { o[0] = x12; o[1] = y12; o[2] = z12; }
                    ka++;
                }
            }
        }
    }
    if (lane == 0) {
        if (mode == 2) {
            for (int t = 0; t < ntypes; t++) { NN_r[n1 * ntypes + t] = tr[t]; NN_a[n1 * ntypes + t] = ta[t]; }
        } else {
            NN_r[n1] = baser;
            NN_a[n1] = basea;
            if (overflow) atomicCAS(status, 0, (int)(n1 + 1));
        }
    }
    if (mode != 0) return;
    // pad the tails (-1 / 0)
    for (int64_t k = (baser < maxr ? baser : maxr) + lane; k < maxr; k += WAVE) {
        NL_r[n1 * maxr + k] = -1;
        for (int q = 0; q < w; q++) Ri_r[(n1 * maxr + k) * w + q] = 0.0;
    }
    for (int64_t k = (basea < maxa ? basea : maxa) + lane; k < maxa; k += WAVE) {
        NL_a[n1 * maxa + k] = -1;
        for (int q = 0; q < w; q++) Ri_a[(n1 * maxa + k) * w + q] = 0.0;
    }
}

// ------------------------------------------------------------------ feature kernels
struct FeatArgs {
    const void *c;          // [T][T][nm][nb]
    const void *Ri;         // [N][S][4]
    const int64_t *NL;      // [N][S]
    const int64_t *atom_map;
    void *feats;            // [N][ld] (written at column offset col0)
    void *s_save;           // [N][nm][24] angular sums (angular only)
    const void *gfeat;      // [N][ld] upstream gradient of feats
    const void *gg;         // [N][S][4] upstream gradient of the position gradient (2nd order)
    void *gRi;              // [N][S][4] out
    void *gc;               // [T][T][nm][nb] out (accumulated)
    void *ggfeat;           // [N][ld] out (2nd order)
    int64_t N;
    int S, T, nm, nb, ld, col0, l3, l4, l5;
    double rc;
};

// first used slot count (loop stops at the first -1)
__device__ __forceinline__ int slot_count(const int64_t *NL, int64_t i, int S) {
    int lo = 0, hi = S;   // NL entries are a valid prefix followed by -1
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (NL[i * S + mid] >= 0) lo = mid + 1; else hi = mid;
    }
    return lo;
}

template <typename T>
__device__ __forceinline__ void block_flush_add(T *lds, T *glob, int count) {
    __syncthreads();
    for (int x = threadIdx.x; x < count; x += blockDim.x) {
        const T v = lds[x];
        if (v != T(0)) atomic_add(glob + x, v);
    }
}

// mode: 0 = forward features, 1 = backward (gRi, gc), 2 = second order (ggfeat, gc)
template <typename T, int MODE>
__global__ void __launch_bounds__(NEP_WAVES * WAVE)
nep_radial_kernel(FeatArgs a) {
    extern __shared__ unsigned char smem_raw[];
    T *gcl = reinterpret_cast<T *>(smem_raw);                // [T*T*nm*nb] block accumulator (MODE > 0)
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    const int ncoef = a.T * a.T * a.nm * a.nb;
    if (MODE > 0) {
        for (int x = threadIdx.x; x < ncoef; x += blockDim.x) gcl[x] = T(0);
        __syncthreads();
    }
    const T *c = reinterpret_cast<const T *>(a.c);
    const T *Ri = reinterpret_cast<const T *>(a.Ri);
    for (int64_t i = (int64_t)blockIdx.x * NEP_WAVES + wid; i < a.N; i += (int64_t)gridDim.x * NEP_WAVES) {
        const int ti = (int)a.atom_map[i];
        const int cnt = slot_count(a.NL, i, a.S);
        T acc[NM_MAX];
        for (int n = 0; n < a.nm; n++) acc[n] = T(0);
        T g[NM_MAX];
        if (MODE > 0)
            for (int n = 0; n < a.nm; n++) g[n] = reinterpret_cast<const T *>(a.gfeat)[i * a.ld + a.col0 + n];
        for (int k = lane; k < a.S; k += WAVE) {
            T *gRi = MODE == 1 ? reinterpret_cast<T *>(a.gRi) + (i * a.S + k) * 4 : nullptr;
            if (k >= cnt) {
                if (MODE == 1) { gRi[0] = gRi[1] = gRi[2] = gRi[3] = T(0); }
                continue;
            }
            const int tj = (int)a.atom_map[a.NL[i * a.S + k]];
            const T r = Ri[(i * a.S + k) * 4];
            T f[NB_MAX], fp[NB_MAX];
            fk_fkp<T>(r, (T)a.rc, a.nb, f, fp);
            const T *cij = c + ((int64_t)(ti * a.T + tj) * a.nm) * a.nb;
            T *gc = gcl + ((ti * a.T + tj) * a.nm) * a.nb;
            if (MODE == 0) {
                for (int n = 0; n < a.nm; n++) {
                    T s = T(0);
                    for (int q = 0; q < a.nb; q++) s += cij[n * a.nb + q] * f[q];
                    acc[n] += s;
                }
            } else if (MODE == 1) {
                T d = T(0);
                for (int n = 0; n < a.nm; n++) {
                    T s = T(0);
                    for (int q = 0; q < a.nb; q++) {
                        s += cij[n * a.nb + q] * fp[q];
                        atomicAdd(gc + n * a.nb + q, g[n] * f[q]);
                    }
                    d += g[n] * s;
                }
                gRi[0] = d; gRi[1] = gRi[2] = gRi[3] = T(0);
            } else {
                const T gg = reinterpret_cast<const T *>(a.gg)[(i * a.S + k) * 4];
                for (int n = 0; n < a.nm; n++) {
                    T s = T(0);
                    for (int q = 0; q < a.nb; q++) {
                        s += cij[n * a.nb + q] * fp[q];
                        atomicAdd(gc + n * a.nb + q, gg * g[n] * fp[q]);
                    }
                    acc[n] += gg * s;
                }
            }
        }
        if (MODE == 0 || MODE == 2) {
            T *out = reinterpret_cast<T *>(MODE == 0 ? a.feats : a.ggfeat);
            for (int n = 0; n < a.nm; n++) {
                const T v = wave_sum(acc[n]);
                if (lane == 0) out[i * a.ld + a.col0 + n] = v;
            }
        }
    }
    if (MODE > 0) block_flush_add(gcl, reinterpret_cast<T *>(a.gc), ncoef);
}

// angular: per-atom LDS arrays of nm*24 values (s, D, P, E)
template <typename T, int MODE>
__global__ void __launch_bounds__(NEP_WAVES * WAVE)
nep_angular_kernel(FeatArgs a) {
    extern __shared__ unsigned char smem_raw[];
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    const int ncoef = a.T * a.T * a.nm * a.nb;
    const int per = a.nm * NLM;
    T *gcl = reinterpret_cast<T *>(smem_raw);                   // [ncoef]
    T *wv = gcl + ncoef + (size_t)wid * 3 * per;                // per wave: arrays A, B, C of nm*24
    T *A = wv, *B = wv + per, *C = wv + 2 * per;
    if (MODE > 0) {
        for (int x = threadIdx.x; x < ncoef; x += blockDim.x) gcl[x] = T(0);
        __syncthreads();
    }
    const T *c = reinterpret_cast<const T *>(a.c);
    const T *Ri = reinterpret_cast<const T *>(a.Ri);
    const T *gfe = reinterpret_cast<const T *>(a.gfeat);
    const int nf = a.l3 + (a.l4 > 0) + (a.l5 > 0);
    for (int64_t i = (int64_t)blockIdx.x * NEP_WAVES + wid; i < a.N; i += (int64_t)gridDim.x * NEP_WAVES) {
        const int ti = (int)a.atom_map[i];
        const int cnt = slot_count(a.NL, i, a.S);
        const T *s_at = reinterpret_cast<const T *>(a.s_save) + i * per;
        // ---- stage per-atom arrays
        if (MODE == 0) {
            for (int x = lane; x < per; x += WAVE) A[x] = T(0);           // A = s
        } else {
            for (int x = lane; x < per; x += WAVE) { A[x] = s_at[x]; C[x] = T(0); }
            __builtin_amdgcn_wave_barrier();
            // B = D = dE/ds (per radial index n, lanes over n)
            for (int n = lane; n < a.nm; n += WAVE) {
                T gq[6];
                for (int f = 0; f < nf; f++) gq[f] = gfe[i * a.ld + a.col0 + aidx(f, n, a.nm)];
                T D[NLM];
                dq_ds<T, T>(A + n * NLM, gq, a.l3, a.l4, a.l5, D);
                for (int m = 0; m < NLM; m++) B[n * NLM + m] = D[m];
            }
        }
        __builtin_amdgcn_wave_barrier();
        // ---- slot pass 1
        for (int k = lane; k < a.S; k += WAVE) {
            T *gRi = MODE == 1 ? reinterpret_cast<T *>(a.gRi) + (i * a.S + k) * 4 : nullptr;
            if (k >= cnt) {
                if (MODE == 1) { gRi[0] = gRi[1] = gRi[2] = gRi[3] = T(0); }
                continue;
            }
            const int tj = (int)a.atom_map[a.NL[i * a.S + k]];
            const T *e = Ri + (i * a.S + k) * 4;
            const T r = e[0];
            T f[NB_MAX], fp[NB_MAX];
            fk_fkp<T>(r, (T)a.rc, a.nb, f, fp);
            T b[NLM], db[NLM][4];
            blm_and_grad<T>(r, e[1], e[2], e[3], b, db);
            const T *cij = c + ((int64_t)(ti * a.T + tj) * a.nm) * a.nb;
            T *gc = gcl + ((ti * a.T + tj) * a.nm) * a.nb;
            T gr[4] = {T(0), T(0), T(0), T(0)};
            const T *gg = MODE == 2 ? reinterpret_cast<const T *>(a.gg) + (i * a.S + k) * 4 : nullptr;
            for (int n = 0; n < a.nm; n++) {
                T gn = T(0), gnp = T(0);
                for (int q = 0; q < a.nb; q++) { gn += cij[n * a.nb + q] * f[q]; gnp += cij[n * a.nb + q] * fp[q]; }
                if (MODE == 0) {
                    for (int m = 0; m < NLM; m++) atomicAdd(A + n * NLM + m, gn * b[m]);
                } else if (MODE == 1) {
                    T Db = T(0);
                    for (int m = 0; m < NLM; m++) {
                        const T D = B[n * NLM + m];
                        Db += D * b[m];
                        gr[0] += D * (gnp * b[m] + gn * db[m][0]);
                        gr[1] += D * gn * db[m][1];
                        gr[2] += D * gn * db[m][2];
                        gr[3] += D * gn * db[m][3];
                    }
                    for (int q = 0; q < a.nb; q++) atomicAdd(gc + n * a.nb + q, Db * f[q]);
                } else {
                    // C = P = sum_slots gg . d(g_n b_m)/dRi ; c-gradient of the explicit term
                    for (int m = 0; m < NLM; m++) {
                        const T t = gg[0] * (gnp * b[m] + gn * db[m][0]) +
                                    gn * (gg[1] * db[m][1] + gg[2] * db[m][2] + gg[3] * db[m][3]);
                        atomicAdd(C + n * NLM + m, t);
                    }
                    for (int q = 0; q < a.nb; q++) {
                        T acc = T(0);
                        for (int m = 0; m < NLM; m++) {
                            const T dfb = gg[0] * (fp[q] * b[m] + f[q] * db[m][0]) +
                                          f[q] * (gg[1] * db[m][1] + gg[2] * db[m][2] + gg[3] * db[m][3]);
                            acc += B[n * NLM + m] * dfb;
                        }
                        atomicAdd(gc + n * a.nb + q, acc);
                    }
                }
            }
            if (MODE == 1) { gRi[0] = gr[0]; gRi[1] = gr[1]; gRi[2] = gr[2]; gRi[3] = gr[3]; }
        }
        __builtin_amdgcn_wave_barrier();
        if (MODE == 0) {
            T *fo = reinterpret_cast<T *>(a.feats);
            T *so = reinterpret_cast<T *>(a.s_save) + i * per;
            for (int x = lane; x < per; x += WAVE) so[x] = A[x];
            for (int n = lane; n < a.nm; n += WAVE) {
                T q[6];
                const int nq = q_of_s<T, T>(A + n * NLM, a.l3, a.l4, a.l5, q);
                for (int f = 0; f < nq; f++) fo[i * a.ld + a.col0 + aidx(f, n, a.nm)] = q[f];
            }
        } else if (MODE == 2) {
            // ggfeat[F] = J P (directional derivative of q along P); E = H(gq) P  -> stored in A
            T *go = reinterpret_cast<T *>(a.ggfeat);
            for (int n = lane; n < a.nm; n += WAVE) {
                Dual<T> sd[NLM];
                for (int m = 0; m < NLM; m++) sd[m] = Dual<T>(A[n * NLM + m], C[n * NLM + m]);
                Dual<T> q[6];
                const int nq = q_of_s<T, Dual<T>>(sd, a.l3, a.l4, a.l5, q);
                for (int f = 0; f < nq; f++) go[i * a.ld + a.col0 + aidx(f, n, a.nm)] = q[f].e;
                T gq[6];
                for (int f = 0; f < nf; f++) gq[f] = gfe[i * a.ld + a.col0 + aidx(f, n, a.nm)];
                Dual<T> D[NLM];
                dq_ds<T, Dual<T>>(sd, gq, a.l3, a.l4, a.l5, D);
                for (int m = 0; m < NLM; m++) C[n * NLM + m] = D[m].e;     // E = (H P)_m
            }
            __builtin_amdgcn_wave_barrier();
            // c-gradient through s: sum_slots sum_m E_nm f_k b_m
            for (int k = lane; k < cnt; k += WAVE) {
                const int tj = (int)a.atom_map[a.NL[i * a.S + k]];
                const T *e = Ri + (i * a.S + k) * 4;
                T f[NB_MAX], fp[NB_MAX];
                fk_fkp<T>(e[0], (T)a.rc, a.nb, f, fp);
                T b[NLM], db[NLM][4];
                blm_and_grad<T>(e[0], e[1], e[2], e[3], b, db);
                T *gc = gcl + ((ti * a.T + tj) * a.nm) * a.nb;
                for (int n = 0; n < a.nm; n++) {
                    T Eb = T(0);
                    for (int m = 0; m < NLM; m++) Eb += C[n * NLM + m] * b[m];
                    for (int q = 0; q < a.nb; q++) atomicAdd(gc + n * a.nb + q, Eb * f[q]);
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (MODE > 0) block_flush_add(gcl, reinterpret_cast<T *>(a.gc), ncoef);
}

// ------------------------------------------------------------------ force / virial ops
// calculateNepForce: F[i] += ... in place: F[j] += dE . Ri_d for every slot j >= 0
// (calculateNepForce.cu:25-70). dE [N][S][4], Ri_d [N][S][4][3], NL int64 [N][S].
template <typename T>
__global__ void nep_force_kernel(const int64_t *__restrict__ NL, const T *__restrict__ dE, const T *__restrict__ Ri_d,
                                 T *__restrict__ F, int64_t nslots) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < nslots;
         s += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = NL[s];
        if (j < 0) continue;
        const T *e = dE + 4 * s, *d = Ri_d + 12 * s;
#pragma unroll
        for (int a = 0; a < 3; a++)
            atomic_add(F + 3 * j + a, e[0] * d[a] + e[1] * d[3 + a] + e[2] * d[6 + a] + e[3] * d[9 + a]);
    }
}

// calculateNepVirial: atom_virial[j][3a+c] += (dE.Ri_d)_a Rij_c; virial per frame
// (calculateNepVirial.cu:19-96).
template <typename T>
__global__ void nep_virial_kernel(const int64_t *__restrict__ NL, const T *__restrict__ dE, const T *__restrict__ Rij,
                                  const T *__restrict__ Ri_d, T *__restrict__ atom_virial, int64_t nslots) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < nslots;
         s += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = NL[s];
        if (j < 0) continue;
        const T *e = dE + 4 * s, *d = Ri_d + 12 * s;
        T f[3];
#pragma unroll
        for (int a = 0; a < 3; a++) f[a] = e[0] * d[a] + e[1] * d[3 + a] + e[2] * d[6 + a] + e[3] * d[9 + a];
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int cc = 0; cc < 3; cc++) atomic_add(atom_virial + 9 * j + 3 * a + cc, f[a] * Rij[3 * s + cc]);
    }
}

// per-frame sum of atom_virial rows: one block per (frame, component)
template <typename T>
__global__ void __launch_bounds__(256) nep_virial_reduce_kernel(const T *__restrict__ atom_virial,
                                                                const int64_t *__restrict__ frame_start,
                                                                T *__restrict__ virial) {
    const int f = blockIdx.y, v = blockIdx.x;
    __shared__ T red[256];
    T acc = T(0);
    for (int64_t r = frame_start[f] + threadIdx.x; r < frame_start[f + 1]; r += 256) acc += atom_virial[9 * r + v];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) virial[9 * f + v] = red[0];
}

// backward of calculateNepForce wrt dE (calculateNepForce.cu:74-122): the centre
// term -gF_i . Ri_d[s][w] plus the neighbor term gF_j . Ri_d[s][w] (j >= 0)
template <typename T>
__global__ void nep_force_grad_kernel(const int64_t *__restrict__ NL, const T *__restrict__ Ri_d,
                                      const T *__restrict__ gF, T *__restrict__ g, int S, int64_t nslots) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < nslots;
         s += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = s / S;
        const int64_t j = NL[s];
        const T *d = Ri_d + 12 * s;
        T q[3] = {-gF[3 * i], -gF[3 * i + 1], -gF[3 * i + 2]};
        if (j >= 0) { q[0] += gF[3 * j]; q[1] += gF[3 * j + 1]; q[2] += gF[3 * j + 2]; }
#pragma unroll
        for (int w = 0; w < 4; w++) g[4 * s + w] = q[0] * d[3 * w] + q[1] * d[3 * w + 1] + q[2] * d[3 * w + 2];
    }
}

// backward of calculateNepVirial wrt dE: g[s][w] = sum_{a,c} gV[f(i)][3a+c] Rij[s][c] Ri_d[s][w][a]
template <typename T>
__global__ void nep_virial_grad_kernel(const int64_t *__restrict__ NL, const T *__restrict__ Rij,
                                       const T *__restrict__ Ri_d, const T *__restrict__ gV,
                                       const int32_t *__restrict__ atom_frame, int S, T *__restrict__ g,
                                       int64_t nslots) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < nslots;
         s += (int64_t)gridDim.x * blockDim.x) {
        T o[4] = {T(0), T(0), T(0), T(0)};
        if (NL[s] >= 0) {
            const T *gv = gV + 9 * (int64_t)atom_frame[s / S];
            T q[3];
#pragma unroll
            for (int a = 0; a < 3; a++) q[a] = gv[3 * a] * Rij[3 * s] + gv[3 * a + 1] * Rij[3 * s + 1] + gv[3 * a + 2] * Rij[3 * s + 2];
            const T *d = Ri_d + 12 * s;
#pragma unroll
            for (int w = 0; w < 4; w++) o[w] = q[0] * d[3 * w] + q[1] * d[3 * w + 1] + q[2] * d[3 * w + 2];
        }
#pragma unroll
        for (int w = 0; w < 4; w++) g[4 * s + w] = o[w];
    }
}

template <typename T>
int feat_launch(FeatArgs &a, int angular, int mode, hipStream_t st) {
    MATPL_CHECK_ARG(a.nb >= 1 && a.nb <= NB_MAX && a.nm >= 1 && a.nm <= NM_MAX, "nep_feat: basis sizes out of range");
    MATPL_CHECK_ARG(a.T >= 1 && a.T <= 16, "nep_feat: ntypes out of range");
    MATPL_CHECK_ARG(!angular || (a.l3 >= 0 && a.l3 <= 4 && (a.l3 == 4 || (a.l4 == 0 && a.l5 == 0))),
                    "nep_feat: l_max_3 must be 4 when 4/5-body terms are on");
    MATPL_CHECK_ARG(!angular || a.l5 == 0 || a.l4 > 0, "nep_feat: the 5-body term needs the 4-body term");
    if (a.N == 0) return OK;
    const size_t ncoef = (size_t)a.T * a.T * a.nm * a.nb;
    size_t lds = (mode > 0 ? ncoef : 0) * sizeof(T);
    if (angular) lds = ncoef * sizeof(T) + (size_t)NEP_WAVES * 3 * a.nm * NLM * sizeof(T);
    MATPL_CHECK_ARG(lds <= 64 * 1024, "nep_feat: LDS %zu too large", lds);
    int64_t blocks = (a.N + NEP_WAVES - 1) / NEP_WAVES;
    if (mode > 0 && blocks > 512) blocks = 512;      // few blocks -> few global coefficient atomics
    if (blocks > 8192) blocks = 8192;
    dim3 grid((unsigned)blocks), block(NEP_WAVES * WAVE);
    if (!angular) {
        if (mode == 0) hipLaunchKernelGGL((nep_radial_kernel<T, 0>), grid, block, lds, st, a);
        else if (mode == 1) hipLaunchKernelGGL((nep_radial_kernel<T, 1>), grid, block, lds, st, a);
        else hipLaunchKernelGGL((nep_radial_kernel<T, 2>), grid, block, lds, st, a);
        return check_launch("nep_radial_kernel");
    }
    if (mode == 0) hipLaunchKernelGGL((nep_angular_kernel<T, 0>), grid, block, lds, st, a);
    else if (mode == 1) hipLaunchKernelGGL((nep_angular_kernel<T, 1>), grid, block, lds, st, a);
    else hipLaunchKernelGGL((nep_angular_kernel<T, 2>), grid, block, lds, st, a);
    return check_launch("nep_angular_kernel");
}

template <typename T>
int force_launch(const int64_t *NL, const T *dE, const T *Ri_d, T *F, int64_t nslots, hipStream_t st) {
    if (nslots == 0) return OK;
    hipLaunchKernelGGL(nep_force_kernel<T>, dim3(grid_cap(nslots)), dim3(256), 0, st, NL, dE, Ri_d, F, nslots);
    return check_launch("nep_force_kernel");
}

template <typename T>
int virial_launch(const int64_t *NL, const T *dE, const T *Rij, const T *Ri_d, const int64_t *frame_start,
                  int64_t nframes, int64_t N, int S, T *virial, T *atom_virial, hipStream_t st) {
    int rc = fill_async<T>(atom_virial, T(0), N * 9, st);
    if (rc) return rc;
    if (N * S > 0) {
        hipLaunchKernelGGL(nep_virial_kernel<T>, dim3(grid_cap(N * S)), dim3(256), 0, st, NL, dE, Rij, Ri_d,
                           atom_virial, N * S);
        if ((rc = check_launch("nep_virial_kernel"))) return rc;
    }
    if (nframes == 0) return OK;
    hipLaunchKernelGGL(nep_virial_reduce_kernel<T>, dim3(9, (unsigned)nframes), dim3(256), 0, st, atom_virial,
                       frame_start, virial);
    return check_launch("nep_virial_reduce_kernel");
}

}  // namespace
}  // namespace matpl

using namespace matpl;

MATPL_EXPORT int matpl_nep_neighbor(const double *box, const double *box_orig, const int64_t *num_cell,
                                    const int64_t *frame_start, const int32_t *atom_frame, const int64_t *type_map,
                                    const double *pos, double rc_r, double rc_a, int64_t maxr, int64_t maxa,
                                    int64_t n, int mode, int ntypes, int rij4, int64_t *NN_r, int64_t *NL_r,
                                    int64_t *NN_a, int64_t *NL_a, double *Ri_r, double *Ri_a, int32_t *status,
                                    void *stream) {
    MATPL_CHECK_ARG(mode >= 0 && mode <= 2 && n >= 0, "nep_neighbor: bad mode/size");
    MATPL_CHECK_ARG(mode != 2 || (ntypes >= 1 && ntypes <= 16), "nep_neighbor: ntypes in [1,16] for typed counts");
    if (n == 0) return OK;
    MATPL_CHECK_ARG(box && box_orig && num_cell && frame_start && atom_frame && pos && NN_r && NN_a,
                    "nep_neighbor: null pointer");
    MATPL_CHECK_ARG(mode != 0 || (NL_r && NL_a && Ri_r && Ri_a && status), "nep_neighbor: null list output");
    MATPL_CHECK_ARG(mode != 2 || type_map, "nep_neighbor: typed counts need type_map");
    NepGeo g{box, box_orig, num_cell};
    hipLaunchKernelGGL(nep_neighbor_kernel, dim3(ceil_div(n, 4)), dim3(256), 0, (hipStream_t)stream, g, frame_start,
                       atom_frame, type_map, pos, rc_r, rc_a, maxr, maxa, n, mode, ntypes, rij4, NN_r, NL_r, NN_a,
                       NL_a, Ri_r, Ri_a, status);
    return check_launch("nep_neighbor_kernel");
}

MATPL_EXPORT int matpl_nep_feat(int f64, int angular, int mode, const void *c, const void *Ri, const int64_t *NL,
                                const int64_t *atom_map, int64_t N, int S, int T, int nm, int nb, int ld, int col0,
                                int l3, int l4, int l5, double rc, void *feats, void *s_save, const void *gfeat,
                                const void *gg, void *gRi, void *gc, void *ggfeat, void *stream) {
    FeatArgs a{};
    a.c = c; a.Ri = Ri; a.NL = NL; a.atom_map = atom_map; a.feats = feats; a.s_save = s_save; a.gfeat = gfeat;
    a.gg = gg; a.gRi = gRi; a.gc = gc; a.ggfeat = ggfeat;
    a.N = N; a.S = S; a.T = T; a.nm = nm; a.nb = nb; a.ld = ld; a.col0 = col0; a.l3 = l3; a.l4 = l4; a.l5 = l5;
    a.rc = rc;
    MATPL_CHECK_ARG(mode >= 0 && mode <= 2, "nep_feat: bad mode");
    MATPL_CHECK_ARG(N == 0 || (c && Ri && NL && atom_map), "nep_feat: null input");
    MATPL_CHECK_ARG(mode != 0 || N == 0 || (feats && (!angular || s_save)), "nep_feat: null output");
    MATPL_CHECK_ARG(mode != 1 || N == 0 || (gfeat && gRi && gc && (!angular || s_save)), "nep_feat: null grad io");
    MATPL_CHECK_ARG(mode != 2 || N == 0 || (gfeat && gg && gc && ggfeat && (!angular || s_save)),
                    "nep_feat: null 2nd-order io");
    hipStream_t st = (hipStream_t)stream;
    return f64 ? feat_launch<double>(a, angular, mode, st) : feat_launch<float>(a, angular, mode, st);
}

MATPL_EXPORT int matpl_nep_force(int f64, const int64_t *NL, const void *dE, const void *Ri_d, void *F, int64_t N, int S,
                                 void *stream) {
    MATPL_CHECK_ARG(N == 0 || (NL && dE && Ri_d && F), "nep_force: null pointer");
    hipStream_t st = (hipStream_t)stream;
    if (f64) return force_launch<double>(NL, (const double *)dE, (const double *)Ri_d, (double *)F, N * S, st);
    return force_launch<float>(NL, (const float *)dE, (const float *)Ri_d, (float *)F, N * S, st);
}

MATPL_EXPORT int matpl_nep_virial(int f64, const int64_t *NL, const void *dE, const void *Rij, const void *Ri_d,
                                  const int64_t *frame_start, int64_t nframes, int64_t N, int S, void *virial,
                                  void *atom_virial, void *stream) {
    MATPL_CHECK_ARG(N == 0 || (NL && dE && Rij && Ri_d && frame_start && virial && atom_virial),
                    "nep_virial: null pointer");
    hipStream_t st = (hipStream_t)stream;
    if (f64)
        return virial_launch<double>(NL, (const double *)dE, (const double *)Rij, (const double *)Ri_d, frame_start,
                                     nframes, N, S, (double *)virial, (double *)atom_virial, st);
    return virial_launch<float>(NL, (const float *)dE, (const float *)Rij, (const float *)Ri_d, frame_start, nframes,
                                N, S, (float *)virial, (float *)atom_virial, st);
}

MATPL_EXPORT int matpl_nep_force_grad(int f64, const int64_t *NL, const void *Ri_d, const void *gF, void *g, int64_t N,
                                      int S, void *stream) {
    if (N * S == 0) return OK;
    MATPL_CHECK_ARG(NL && Ri_d && gF && g, "nep_force_grad: null pointer");
    hipStream_t st = (hipStream_t)stream;
    if (f64)
        hipLaunchKernelGGL(nep_force_grad_kernel<double>, dim3(grid_cap(N * S)), dim3(256), 0, st, NL,
                           (const double *)Ri_d, (const double *)gF, (double *)g, S, N * S);
    else
        hipLaunchKernelGGL(nep_force_grad_kernel<float>, dim3(grid_cap(N * S)), dim3(256), 0, st, NL,
                           (const float *)Ri_d, (const float *)gF, (float *)g, S, N * S);
    return check_launch("nep_force_grad_kernel");
}

MATPL_EXPORT int matpl_nep_virial_grad(int f64, const int64_t *NL, const void *Rij, const void *Ri_d, const void *gV,
                                       const int32_t *atom_frame, void *g, int64_t N, int S, void *stream) {
    if (N * S == 0) return OK;
    MATPL_CHECK_ARG(NL && Rij && Ri_d && gV && atom_frame && g, "nep_virial_grad: null pointer");
    hipStream_t st = (hipStream_t)stream;
    if (f64)
        hipLaunchKernelGGL(nep_virial_grad_kernel<double>, dim3(grid_cap(N * S)), dim3(256), 0, st, NL,
                           (const double *)Rij, (const double *)Ri_d, (const double *)gV, atom_frame, S, (double *)g,
                           N * S);
    else
        hipLaunchKernelGGL(nep_virial_grad_kernel<float>, dim3(grid_cap(N * S)), dim3(256), 0, st, NL,
                           (const float *)Rij, (const float *)Ri_d, (const float *)gV, atom_frame, S, (float *)g,
                           N * S);
    return check_launch("nep_virial_grad_kernel");
}
