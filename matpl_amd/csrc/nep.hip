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
// Layout: one wavefront per atom, lanes over neighbor slots for the per-pair
// work; sums over slots (s_nlm, their tangents, coefficient gradients) run
// transposed with lanes over outputs, so lanes never add to one address; the
// per-type coefficient gradients are accumulated per block in LDS and added
// to global memory once per block (persistent grid).
#include "common.h"

#pragma clang fp contract(off)

namespace matpl {
namespace {

constexpr int WAVE = 64;
constexpr int NEP_WAVES = 4;
constexpr int NB_MAX = 20;      // basis functions
constexpr int NM_MAX = 8;       // radial functions per descriptor part
constexpr int NLM = 24;
constexpr int WS = NLM + 1;     // LDS row stride of a slot's harmonics (odd: lanes writing their rows hit distinct banks)
constexpr double PI_D = 3.14159265358979323846;

// f_k(r) and f_k'(r) for k = 0, 1, 2, ... in order, one per next() call
// (Chebyshev recursion, nothing stored per k).
template <typename T>
struct Basis {
    T x, xp, fc, fcp, t0, t1, tp0, tp1;
    int k;
    __device__ __forceinline__ Basis(T r, T rc) {
        const T rcinv = T(1) / rc;
        fc = T(0);
        fcp = T(0);
        if (r < rc) {
            const T a = T(PI_D) * r * rcinv;
            fc = T(0.5) * cos(a) + T(0.5);
            fcp = -T(0.5 * PI_D) * sin(a) * rcinv;
        }
        const T d = r * rcinv - T(1);
        x = T(2) * d * d - T(1);
        xp = T(4) * d * rcinv;
        t0 = T(1); t1 = x; tp0 = T(0); tp1 = T(1);
        k = 0;
    }
    __device__ __forceinline__ void next(T &f, T &fp) {
        T t, tp;
        if (k == 0) { t = t0; tp = tp0; }
        else if (k == 1) { t = t1; tp = tp1; }
        else {
            t = T(2) * x * t1 - t0;
            tp = T(2) * t1 + T(2) * x * tp1 - tp0;
            t0 = t1; t1 = t; tp0 = tp1; tp1 = tp;
        }
        k++;
        f = (t + T(1)) * T(0.5) * fc;
        fp = T(0.5) * (tp * xp * fc + (t + T(1)) * fcp);
    }
};

// The 24 real harmonics b_k(u) of a unit vector u (accumulate_s order): calls
// fn(k, b_k, db_k/dux, db_k/duy, db_k/duz) for k = 0..23. With r and (x, y, z)
// as independent inputs and u = (x, y, z)/r:
//   db_k/dx_w = (db_k/du_w) / r,   db_k/dr = -(u . db_k/du) / r.
template <typename T, typename Fn>
__device__ __forceinline__ void blm_each(T x, T y, T z, Fn &&fn) {
    const T x2 = x * x, y2 = y * y, z2 = z * z, xy2 = x2 - y2;
    fn(0, z, T(0), T(0), T(1));
    fn(1, x, T(1), T(0), T(0));
    fn(2, y, T(0), T(1), T(0));
    fn(3, T(3) * z2 - T(1), T(0), T(0), T(6) * z);
    fn(4, x * z, z, T(0), x);
    fn(5, y * z, T(0), z, y);
    fn(6, xy2, T(2) * x, -T(2) * y, T(0));
    fn(7, T(2) * x * y, T(2) * y, T(2) * x, T(0));
    fn(8, (T(5) * z2 - T(3)) * z, T(0), T(0), T(15) * z2 - T(3));
    fn(9, (T(5) * z2 - T(1)) * x, T(5) * z2 - T(1), T(0), T(10) * x * z);
    fn(10, (T(5) * z2 - T(1)) * y, T(0), T(5) * z2 - T(1), T(10) * y * z);
    fn(11, xy2 * z, T(2) * x * z, -T(2) * y * z, xy2);
    fn(12, T(2) * x * y * z, T(2) * y * z, T(2) * x * z, T(2) * x * y);
    fn(13, (x2 - T(3) * y2) * x, T(3) * x2 - T(3) * y2, -T(6) * x * y, T(0));
    fn(14, (T(3) * x2 - y2) * y, T(6) * x * y, T(3) * x2 - T(3) * y2, T(0));
    fn(15, (T(35) * z2 - T(30)) * z2 + T(3), T(0), T(0), T(140) * z2 * z - T(60) * z);
    fn(16, (T(7) * z2 - T(3)) * x * z, (T(7) * z2 - T(3)) * z, T(0), T(21) * x * z2 - T(3) * x);
    fn(17, (T(7) * z2 - T(3)) * y * z, T(0), (T(7) * z2 - T(3)) * z, T(21) * y * z2 - T(3) * y);
    fn(18, (T(7) * z2 - T(1)) * xy2, T(2) * x * (T(7) * z2 - T(1)), -T(2) * y * (T(7) * z2 - T(1)),
       T(14) * z * xy2);
    fn(19, (T(7) * z2 - T(1)) * x * y * T(2), T(2) * y * (T(7) * z2 - T(1)), T(2) * x * (T(7) * z2 - T(1)),
       T(28) * x * y * z);
    fn(20, (x2 - T(3) * y2) * x * z, (T(3) * x2 - T(3) * y2) * z, -T(6) * x * y * z, (x2 - T(3) * y2) * x);
    fn(21, (T(3) * x2 - y2) * y * z, T(6) * x * y * z, (T(3) * x2 - T(3) * y2) * z, (T(3) * x2 - y2) * y);
    fn(22, xy2 * xy2 - T(4) * x2 * y2, T(4) * x2 * x - T(12) * x * y2, T(4) * y2 * y - T(12) * x2 * y, T(0));
    fn(23, T(4) * x * y * xy2, T(12) * x2 * y - T(4) * y2 * y, T(4) * x2 * x - T(12) * x * y2, T(0));
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
// The 4/5-body terms need l3 = 4, so their slots are q[4] and q[5]. Every index
// below is a compile-time constant once unrolled, keeping s, q and D in registers.
constexpr int LSTART[5] = {0, 3, 8, 15, 24};
template <typename T, typename S>
__device__ __forceinline__ int q_of_s(const S s[NLM], int l3, int l4, int l5, S q[6]) {
#pragma unroll
    for (int l = 0; l < 4; l++) {
        S acc(T(0));
#pragma unroll
        for (int m = LSTART[l]; m < LSTART[l + 1]; m++) acc = acc + T(kC3B[m]) * (s[m] * s[m]);
        q[l] = acc;
    }
    if (l4 > 0) {
        q[4] = T(kC4B[0]) * (s[3] * s[3] * s[3]) + T(kC4B[1]) * (s[3] * (s[4] * s[4] + s[5] * s[5])) +
                  T(kC4B[2]) * (s[3] * (s[6] * s[6] + s[7] * s[7])) + T(kC4B[3]) * (s[6] * (s[5] * s[5] - s[4] * s[4])) +
                  T(kC4B[4]) * (s[4] * s[5] * s[7]);
    }
    if (l5 > 0) {
        const S s0 = s[0] * s[0], s12 = s[1] * s[1] + s[2] * s[2];
        q[5] = T(kC5B[0]) * (s0 * s0) + T(kC5B[1]) * (s0 * s12) + T(kC5B[2]) * (s12 * s12);
    }
    return l3 + (l4 > 0) + (l5 > 0);
}

// D[m] = sum_F gq[F] dq_F/ds_m (gq in q_of_s order).
template <typename T, typename S>
__device__ __forceinline__ void dq_ds(const S s[NLM], const T *gq, int l3, int l4, int l5, S D[NLM]) {
#pragma unroll
    for (int m = 0; m < NLM; m++) D[m] = S(T(0));
#pragma unroll
    for (int l = 0; l < 4; l++)
        if (l < l3)
#pragma unroll
            for (int m = LSTART[l]; m < LSTART[l + 1]; m++) D[m] = D[m] + (T(2) * T(kC3B[m]) * gq[l]) * s[m];
    if (l4 > 0) {
        const T g = gq[4];
        const T c0 = T(kC4B[0]), c1 = T(kC4B[1]), c2 = T(kC4B[2]), c3 = T(kC4B[3]), c4 = T(kC4B[4]);
        D[3] = D[3] + g * (T(3) * c0 * (s[3] * s[3]) + c1 * (s[4] * s[4] + s[5] * s[5]) + c2 * (s[6] * s[6] + s[7] * s[7]));
        D[4] = D[4] + g * (T(2) * c1 * (s[3] * s[4]) - T(2) * c3 * (s[6] * s[4]) + c4 * (s[5] * s[7]));
        D[5] = D[5] + g * (T(2) * c1 * (s[3] * s[5]) + T(2) * c3 * (s[6] * s[5]) + c4 * (s[4] * s[7]));
        D[6] = D[6] + g * (T(2) * c2 * (s[3] * s[6]) + c3 * (s[5] * s[5] - s[4] * s[4]));
        D[7] = D[7] + g * (T(2) * c2 * (s[3] * s[7]) + c4 * (s[4] * s[5]));
    }
    if (l5 > 0) {
        const T g = gq[5];
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

// Displacement of replica (ia, ib, ic) of the atom at (px, py, pz) from n1, minimum image in the
// expanded box h, and its square; the expression order is the reference's (calculateNepNeighbor.cu),
// so that the lists and [r, x, y, z] are bit-identical.
__device__ __forceinline__ double replica_disp(const double *h, const double *bo, double px, double py, double pz,
                                               int ia, int ib, int ic, double x1, double y1, double z1, double &x12,
                                               double &y12, double &z12) {
    const double dx = bo[0] * (double)ia + bo[1] * (double)ib + bo[2] * (double)ic;
    const double dy = bo[3] * (double)ia + bo[4] * (double)ib + bo[5] * (double)ic;
    const double dz = bo[6] * (double)ia + bo[7] * (double)ib + bo[8] * (double)ic;
    x12 = px + dx - x1;
    y12 = py + dy - y1;
    z12 = pz + dz - z1;
    apply_mic(h, x12, y12, z12);
    return x12 * x12 + y12 * y12 + z12 * z12;
}

// (R = float: the FP64 values rounded once, as a cast of the double rows would)
template <typename R>
__device__ __forceinline__ void put_entry(int64_t *NL, R *Ri, int64_t idx, int w, int rij4, int64_t n2, double d2,
                                          double x, double y, double z) {
    NL[idx] = n2;
    R *o = Ri + idx * w;
    if (rij4) { o[0] = (R)sqrt(d2); o[1] = (R)x; o[2] = (R)y; o[3] = (R)z; }
    else { o[0] = (R)x; o[1] = (R)y; o[2] = (R)z; }
}

// MODE 0: lists; 1: counts; 2: per-type counts. One wavefront per atom n1, lanes over the frame's
// atoms n2, each with its replicas (ia, ib, ic) in the reference's order; a list holds the hits in
// (n2, replica) order. With one replica (the box is at least 2 rc thick) a lane has at most one hit
// per list and one pass both counts and writes; several replicas take a counting pass, a scan and a
// writing pass. The next chunk's positions are loaded before the current chunk is tested, and the
// per-type counts of MODE 2 are held by lane t for type t, so that the lists and counts modes keep
// their register count low (several wavefronts per SIMD hide the load latency).
template <int MODE, typename R = double>
__global__ void __launch_bounds__(256)
nep_neighbor_kernel(NepGeo g, const int64_t *__restrict__ frame_start, const int32_t *__restrict__ atom_frame,
                    const int64_t *__restrict__ type_map, const double *__restrict__ pos, double rc_r, double rc_a,
                    int64_t maxr, int64_t maxa, int64_t n, int ntypes, int rij4, int64_t *__restrict__ NN_r,
                    int64_t *__restrict__ NL_r, int64_t *__restrict__ NN_a, int64_t *__restrict__ NL_a,
                    R *__restrict__ Ri_r, R *__restrict__ Ri_a, int32_t *__restrict__ status) {
    const int lane = threadIdx.x & (WAVE - 1);
    // the wave index is uniform: the atom, its frame and the frame's boxes stay in scalar registers
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const int64_t n1 = blockIdx.x * (int64_t)(blockDim.x / WAVE) + wid;
    if (n1 >= n) return;
    const int f = atom_frame[n1];
    const int64_t N1 = frame_start[f], N2 = frame_start[f + 1];
    double h[18], bo[9];
#pragma unroll
    for (int q = 0; q < 18; q++) h[q] = g.box[18 * (int64_t)f + q];
#pragma unroll
    for (int q = 0; q < 9; q++) bo[q] = g.box_orig[9 * (int64_t)f + q];
    const int64_t *nc = g.num_cell + 3 * (int64_t)f;
    const int na = (int)nc[0], nb = (int)nc[1], ncc = (int)nc[2];
    const bool one = na == 1 && nb == 1 && ncc == 1;
    const double x1 = pos[3 * n1], y1 = pos[3 * n1 + 1], z1 = pos[3 * n1 + 2];
    const double rr2 = rc_r * rc_r, ra2 = rc_a * rc_a;
    const int w = rij4 ? 4 : 3;
    const uint64_t below = (1ull << lane) - 1ull;
    int64_t baser = 0, basea = 0;
    int64_t mytr = 0, myta = 0;     // MODE 2: the counts of type `lane`
    bool overflow = false;
    double px = 0.0, py = 0.0, pz = 0.0;
    if (N1 + lane < N2) {
        const int64_t n2 = N1 + lane;
        px = pos[3 * n2]; py = pos[3 * n2 + 1]; pz = pos[3 * n2 + 2];
    }
    for (int64_t c0 = N1; c0 < N2; c0 += WAVE) {
        const int64_t n2 = c0 + lane;
        const bool valid = n2 < N2;
        const double cx = px, cy = py, cz = pz;
        if (n2 + WAVE < N2) {
            const int64_t m2 = n2 + WAVE;
            px = pos[3 * m2]; py = pos[3 * m2 + 1]; pz = pos[3 * m2 + 2];
        }
        int cr = 0, ca = 0;
        // type_map is only given (and only read) for the per-type counts of MODE 2
        const int t2 = (MODE == 2 && valid) ? (int)type_map[n2] : 0;
        if (one) {
            // one replica: count and, in MODE 0, write in the same pass
            double x12 = 0.0, y12 = 0.0, z12 = 0.0, d2 = 0.0;
            if (valid && n1 != n2) {
                d2 = replica_disp(h, bo, cx, cy, cz, 0, 0, 0, x1, y1, z1, x12, y12, z12);
                cr = d2 < rr2 ? 1 : 0;
                ca = d2 < ra2 ? 1 : 0;
            }
            if constexpr (MODE == 2) {
                for (int t = 0; t < ntypes; t++) {
                    const int ur = __popcll(__ballot(cr && t2 == t)), ua = __popcll(__ballot(ca && t2 == t));
                    if (lane == t) { mytr += ur; myta += ua; }
                }
            } else {
                const uint64_t br = __ballot(cr), ba = __ballot(ca);
                const int64_t kr = baser + __popcll(br & below), ka = basea + __popcll(ba & below);
                baser += __popcll(br);
                basea += __popcll(ba);
                if constexpr (MODE == 0) {
                    if (baser > maxr || basea > maxa) overflow = true;
                    if (cr && kr < maxr) put_entry(NL_r, Ri_r, n1 * maxr + kr, w, rij4, n2, d2, x12, y12, z12);
                    if (ca && ka < maxa) put_entry(NL_a, Ri_a, n1 * maxa + ka, w, rij4, n2, d2, x12, y12, z12);
                }
            }
            continue;
        }
        // pass 1: counts
        if (valid) {
            for (int ia = 0; ia < na; ia++)
                for (int ib = 0; ib < nb; ib++)
                    for (int ic = 0; ic < ncc; ic++) {
                        if (ia == 0 && ib == 0 && ic == 0 && n1 == n2) continue;
                        double x12, y12, z12;
                        const double d2 = replica_disp(h, bo, cx, cy, cz, ia, ib, ic, x1, y1, z1, x12, y12, z12);
                        if (d2 < rr2) cr++;
                        if (d2 < ra2) ca++;
                    }
        }
        if constexpr (MODE == 2) {
            for (int t = 0; t < ntypes; t++) {
                const int vr = (valid && t2 == t) ? cr : 0, va = (valid && t2 == t) ? ca : 0;
                const int sr = wave_incl_scan(vr), sa = wave_incl_scan(va);
                const int ur = __shfl(sr, WAVE - 1, WAVE), ua = __shfl(sa, WAVE - 1, WAVE);
                if (lane == t) { mytr += ur; myta += ua; }
            }
            continue;
        }
        const int ir = wave_incl_scan(cr), ia_ = wave_incl_scan(ca);
        const int sumr = __shfl(ir, WAVE - 1, WAVE), suma = __shfl(ia_, WAVE - 1, WAVE);
        const int64_t offr = baser + ir - cr, offa = basea + ia_ - ca;
        baser += sumr;
        basea += suma;
        if constexpr (MODE == 1) continue;
        if (baser > maxr || basea > maxa) { overflow = true; continue; }
        // pass 2: write in (n2, replica) order
        if (valid && (cr > 0 || ca > 0)) {
            int64_t kr = offr, ka = offa;
            for (int ia = 0; ia < na; ia++)
                for (int ib = 0; ib < nb; ib++)
                    for (int ic = 0; ic < ncc; ic++) {
                        if (ia == 0 && ib == 0 && ic == 0 && n1 == n2) continue;
                        double x12, y12, z12;
                        const double d2 = replica_disp(h, bo, cx, cy, cz, ia, ib, ic, x1, y1, z1, x12, y12, z12);
                        if (d2 < rr2) put_entry(NL_r, Ri_r, n1 * maxr + kr++, w, rij4, n2, d2, x12, y12, z12);
                        if (d2 < ra2) put_entry(NL_a, Ri_a, n1 * maxa + ka++, w, rij4, n2, d2, x12, y12, z12);
                    }
        }
    }
    if constexpr (MODE == 2) {
        if (lane < ntypes) { NN_r[n1 * ntypes + lane] = mytr; NN_a[n1 * ntypes + lane] = myta; }
        return;
    }
    if (lane == 0) {
        NN_r[n1] = baser;
        NN_a[n1] = basea;
        if (overflow) atomicCAS(status, 0, (int)(n1 + 1));
    }
    if constexpr (MODE != 0) return;
    // pad the tails (-1 / 0)
    for (int64_t k = (baser < maxr ? baser : maxr) + lane; k < maxr; k += WAVE) {
        NL_r[n1 * maxr + k] = -1;
        for (int q = 0; q < w; q++) Ri_r[(n1 * maxr + k) * w + q] = R(0);
    }
    for (int64_t k = (basea < maxa ? basea : maxa) + lane; k < maxa; k += WAVE) {
        NL_a[n1 * maxa + k] = -1;
        for (int q = 0; q < w; q++) Ri_a[(n1 * maxa + k) * w + q] = R(0);
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
    int pw;                 // coefficient-gradient tables per wavefront (plain adds) instead of per block (LDS atomics)
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

// The block's coefficient-gradient tables (ntab of count entries: one per wavefront, or one shared)
// summed and added to global memory, one atomic per entry.
template <typename T>
__device__ __forceinline__ void block_flush_add(T *lds, T *glob, int count, int ntab) {
    __syncthreads();
    for (int x = threadIdx.x; x < count; x += blockDim.x) {
        T v = lds[x];
        for (int w = 1; w < ntab; w++) v += lds[w * count + x];
        if (v != T(0)) atomic_add(glob + x, v);
    }
}

// Feature kernels. Radial: one wavefront per centre atom; angular: a wavefront
// walks a contiguous range of atoms, several atoms' slots to a chunk (below).
// Per-neighbor work runs with lanes over the list entries (slots) and writes
// per-slot factors to a per-wave LDS chunk. Every sum over slots then runs transposed, with lanes
// over its outputs reading the chunk as broadcasts, so no two lanes of a
// wavefront ever add to the same address:
//   radial  q_n                  lanes over slots, wave sums (registers)
//   radial  dL/dc[ti][tj][n][q]  = g_n sum_{k: tj} v_k[q]: lanes over (tj, q)
//   angular s_nm, P_nm           = sum_k u_k[n] w_k[m]: lanes over (n, m)
//   angular dL/dc                = sum_k u_k[n] v_k[q]: lanes over (n, q), one
//                                  LDS add per lane per slot into the block's table
// The block's coefficient-gradient table goes to global memory with one atomic
// per entry at the end.
// mode: 0 = forward features, 1 = backward (gRi, gc), 2 = second order (ggfeat, gc)
constexpr int RCH = WAVE;        // radial slots per chunk
constexpr int ANG_WAVES = 2;     // wavefronts per angular block (LDS per wave is large)

template <typename T, int MODE>
__global__ void __launch_bounds__(NEP_WAVES * WAVE)
nep_radial_kernel(FeatArgs a) {
    extern __shared__ unsigned char smem_raw[];
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    const int ncoef = a.T * a.T * a.nm * a.nb;
    const int ntab = (MODE > 0 && a.pw) ? NEP_WAVES : 1;
    T *gcl = reinterpret_cast<T *>(smem_raw);                    // [ntab][T*T*nm*nb] tables (MODE > 0)
    T *gtab = gcl + (ntab > 1 ? wid : 0) * ncoef;                // this wave's table
    // per-wave chunk: v[RCH][nb] and the neighbor type of each slot (MODE > 0)
    T *vch = gcl + (size_t)ntab * ncoef + (size_t)wid * RCH * a.nb;
    int *tch = reinterpret_cast<int *>(gcl + (size_t)ntab * ncoef + (size_t)NEP_WAVES * RCH * a.nb) + wid * RCH;
    if (MODE > 0) {
        for (int x = threadIdx.x; x < ntab * ncoef; x += blockDim.x) gcl[x] = T(0);
        __syncthreads();
    }
    const T *c = reinterpret_cast<const T *>(a.c);
    const T *Ri = reinterpret_cast<const T *>(a.Ri);
    for (int64_t i = (int64_t)blockIdx.x * NEP_WAVES + wid; i < a.N; i += (int64_t)gridDim.x * NEP_WAVES) {
        const int ti = (int)a.atom_map[i];
        const int cnt = slot_count(a.NL, i, a.S);
        T acc[NM_MAX], g[NM_MAX];
#pragma unroll
        for (int n = 0; n < NM_MAX; n++) {
            acc[n] = T(0);
            g[n] = (MODE > 0 && n < a.nm) ? reinterpret_cast<const T *>(a.gfeat)[i * a.ld + a.col0 + n] : T(0);
        }
        if (MODE == 1)
            for (int k = cnt + lane; k < a.S; k += WAVE) {
                T *gRi = reinterpret_cast<T *>(a.gRi) + (i * a.S + k) * 4;
                gRi[0] = gRi[1] = gRi[2] = gRi[3] = T(0);
            }
        for (int k0 = 0; k0 < cnt; k0 += RCH) {
            const int k = k0 + lane;
            if (k < cnt) {
                const int tj = (int)a.atom_map[a.NL[i * a.S + k]];
                const T r = Ri[(i * a.S + k) * 4];
                const T gg = MODE == 2 ? reinterpret_cast<const T *>(a.gg)[(i * a.S + k) * 4] : T(0);
                const T *cij = c + ((int64_t)(ti * a.T + tj) * a.nm) * a.nb;
                Basis<T> bs(r, (T)a.rc);
                T sn[NM_MAX];
#pragma unroll
                for (int n = 0; n < NM_MAX; n++) sn[n] = T(0);
                for (int q = 0; q < a.nb; q++) {
                    T f, fp;
                    bs.next(f, fp);
                    const T fq = MODE == 0 ? f : fp;
#pragma unroll
                    for (int n = 0; n < NM_MAX; n++)
                        if (n < a.nm) sn[n] += cij[n * a.nb + q] * fq;
                    if (MODE == 1) vch[lane * a.nb + q] = f;            // dc = g_n f_q
                    if (MODE == 2) vch[lane * a.nb + q] = gg * fp;      // dc = g_n gg fp_q
                }
                if (MODE == 1) {
                    T d = T(0);
#pragma unroll
                    for (int n = 0; n < NM_MAX; n++)
                        if (n < a.nm) d += g[n] * sn[n];
                    T *gRi = reinterpret_cast<T *>(a.gRi) + (i * a.S + k) * 4;
                    gRi[0] = d; gRi[1] = gRi[2] = gRi[3] = T(0);
                } else {
#pragma unroll
                    for (int n = 0; n < NM_MAX; n++) acc[n] += MODE == 0 ? sn[n] : gg * sn[n];
                }
                if (MODE > 0) tch[lane] = tj;
            }
            if (MODE > 0) {
                // dc[ti][tj][n][q] += g_n sum_{slots of type tj} v[q]: lanes over (tj, q)
                __builtin_amdgcn_wave_barrier();
                const int nk = cnt - k0 < RCH ? cnt - k0 : RCH;
                for (int o = lane; o < a.T * a.nb; o += WAVE) {
                    const int tj = o / a.nb, q = o - tj * a.nb;
                    T sv = T(0);
                    for (int kk = 0; kk < nk; kk++)
                        if (tch[kk] == tj) sv += vch[kk * a.nb + q];
                    if (sv != T(0)) {
                        T *gcp = gtab + ((ti * a.T + tj) * a.nm) * a.nb + q;
                        // lanes own distinct (tj, q): a wave's own table needs no atomics
#pragma unroll
                        for (int n = 0; n < NM_MAX; n++)
                            if (n < a.nm) {
                                if (ntab > 1) gcp[n * a.nb] += g[n] * sv;
                                else atomicAdd(gcp + n * a.nb, g[n] * sv);
                            }
                    }
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
        if (MODE == 0 || MODE == 2) {
            T *out = reinterpret_cast<T *>(MODE == 0 ? a.feats : a.ggfeat);
#pragma unroll
            for (int n = 0; n < NM_MAX; n++) {
                if (n < a.nm) {
                    const T v = wave_sum(acc[n]);
                    if (lane == 0) out[i * a.ld + a.col0 + n] = v;
                }
            }
        }
    }
    if (MODE > 0) block_flush_add(gcl, reinterpret_cast<T *>(a.gc), ncoef, ntab);
}

// Angular features (mode 0): s_nm = sum over an atom's slots of g_n(r) b_m(u), with the slots of up
// to AB_ATOMS atoms in one 64-lane chunk (an atom has ~16 angular neighbors at rc = 4 A). Lanes take
// slots for g_n and b_m (written to per-wave LDS rows), then lanes own (n, m) outputs and sum each
// atom's rows; s goes to s_save (the derivative passes read it) and q_of_s(s) to the features.
constexpr int AB_ATOMS = 4;      // atoms whose angular slots share one chunk

__host__ __device__ inline int angf_wave_elems(int nm) { return AB_ATOMS * nm * NLM + WAVE * (nm + WS); }

template <typename T>
__global__ void __launch_bounds__(ANG_WAVES * WAVE)
nep_angular_fwd_kernel(FeatArgs a) {
    constexpr int NRED = (NM_MAX * NLM + WAVE - 1) / WAVE;
    extern __shared__ unsigned char smem_raw[];
    const int lane = threadIdx.x & (WAVE - 1);
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const int per = a.nm * NLM;
    const int nm = a.nm, nb = a.nb;
    T *Sq = reinterpret_cast<T *>(smem_raw) + (size_t)wid * angf_wave_elems(nm);   // [AB_ATOMS][nm][24]
    T *u1 = Sq + AB_ATOMS * per;                                                    // [WAVE][nm] g_n
    T *w1 = u1 + WAVE * nm;                                                         // [WAVE][WS] b_m
    const T *c = reinterpret_cast<const T *>(a.c);
    const T *Ri = reinterpret_cast<const T *>(a.Ri);
    T *fo = reinterpret_cast<T *>(a.feats);
    T *so = reinterpret_cast<T *>(a.s_save);
    const int64_t nw = (int64_t)gridDim.x * ANG_WAVES, gw = (int64_t)blockIdx.x * ANG_WAVES + wid;
    const int64_t i1 = a.N * (gw + 1) / nw;
    for (int64_t i = a.N * gw / nw; i < i1;) {
        const int64_t il = i + (lane < AB_ATOMS ? lane : 0);
        const int c_l = (lane < AB_ATOMS && il < i1) ? slot_count(a.NL, il, a.S) : -1;
        int off[AB_ATOMS + 1];
        off[0] = 0;
        int na = 0;
        bool open = true;
#pragma unroll
        for (int q = 0; q < AB_ATOMS; q++) {
            const int cq = __builtin_amdgcn_readlane(c_l, q);
            open = open && (q == 0 || (cq >= 0 && off[q] + cq <= WAVE));
            if (open) na = q + 1;
            off[q + 1] = off[q] + (open ? cq : 0);
        }
        const int tot = off[AB_ATOMS];
        T red[AB_ATOMS][NRED];
#pragma unroll
        for (int q = 0; q < AB_ATOMS; q++)
#pragma unroll
            for (int j = 0; j < NRED; j++) red[q][j] = T(0);
        for (int e0 = 0; e0 < tot; e0 += WAVE) {
            const int nk = tot - e0 < WAVE ? tot - e0 : WAVE;
            if (lane < nk) {
                const int e = e0 + lane;
                int q = 0;
#pragma unroll
                for (int qq = 1; qq < AB_ATOMS; qq++) q += (e >= off[qq]) ? 1 : 0;
                const int64_t ia = i + q;
                const int k = e - off[q];
                const int ti = (int)a.atom_map[ia];
                const int tj = (int)a.atom_map[a.NL[ia * a.S + k]];
                const T *er = Ri + (ia * a.S + k) * 4;
                const T r = er[0], ir = T(1) / r;
                const T *cij = c + ((int64_t)(ti * a.T + tj) * nm) * nb;
                T gn[NM_MAX];
                {
                    Basis<T> bs(r, (T)a.rc);
#pragma unroll
                    for (int n = 0; n < NM_MAX; n++) gn[n] = T(0);
                    for (int qb = 0; qb < nb; qb++) {
                        T f, fp;
                        bs.next(f, fp);
#pragma unroll
                        for (int n = 0; n < NM_MAX; n++)
                            if (n < nm) gn[n] += cij[n * nb + qb] * f;
                    }
                }
#pragma unroll
                for (int n = 0; n < NM_MAX; n++)
                    if (n < nm) u1[lane * nm + n] = gn[n];
                blm_each<T>(er[1] * ir, er[2] * ir, er[3] * ir, [&](int m, T bv, T, T, T) { w1[lane * WS + m] = bv; });
            }
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int q = 0; q < AB_ATOMS; q++) {
                if (q >= na) break;
                const int k0 = off[q] - e0 > 0 ? off[q] - e0 : 0;
                const int k1 = off[q + 1] - e0 < nk ? off[q + 1] - e0 : nk;
#pragma unroll
                for (int j = 0; j < NRED; j++) {
                    const int o = lane + j * WAVE;
                    if (o < per) {
                        const int n = o / NLM, m = o - n * NLM;
                        T sacc = T(0);
                        for (int kk = k0; kk < k1; kk++) sacc += u1[kk * nm + n] * w1[kk * WS + m];
                        red[q][j] += sacc;
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
#pragma unroll
        for (int q = 0; q < AB_ATOMS; q++) {
            if (q >= na) break;
#pragma unroll
            for (int j = 0; j < NRED; j++) {
                const int o = lane + j * WAVE;
                if (o < per) {
                    Sq[q * per + o] = red[q][j];
                    so[(i + q) * per + o] = red[q][j];
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        if (lane < na * nm) {
            const int q = lane / nm, n = lane - q * nm;
            T qv[6];
            const int nq = q_of_s<T, T>(Sq + q * per + n * NLM, a.l3, a.l4, a.l5, qv);
#pragma unroll
            for (int f = 0; f < 6; f++)
                if (f < nq) fo[(i + q) * a.ld + a.col0 + aidx(f, n, nm)] = qv[f];
        }
        __builtin_amdgcn_wave_barrier();
        i += na;
    }
}

// ------------------------------------------------------------------ force / virial ops
// Neighbor terms of calculateNepForce (NV = 3: out[j] += f, calculateNepForce.cu:25-70) and of
// calculateNepVirial (NV = 9: out[j][3a+c] += f_a Rij[s][c], calculateNepVirial.cu:19-96), with
// f = sum_w dE[s][w] Ri_d[s][w][:] for every slot s whose NL[s] = j >= 0. dE [N][S][4],
// Ri_d [N][S][4][3], Rij [N][S][3], NL int64 [N][S]; out is accumulated into.
// A block takes SC_CH consecutive atoms: a first pass over their slots finds the range [lo, hi) of
// the atoms they name, the terms are added into an LDS copy of out over that range, and the copy's
// nonzero entries are added to out once per block. Per-slot global atomics, which on this GPU are
// carried out beyond the per-XCD L2 caches, remain only for blocks whose range exceeds the copy.
constexpr int SC_CH = 64;       // atoms per block

template <typename T, int NV>
__global__ void __launch_bounds__(256)
nep_scatter_kernel(const int64_t *__restrict__ NL, const T *__restrict__ dE, const T *__restrict__ Ri_d,
                   const T *__restrict__ Rij, T *__restrict__ out, int64_t N, int S, int64_t rcap) {
    extern __shared__ double sc_smem[];
    T *acc = reinterpret_cast<T *>(sc_smem);
    __shared__ int64_t wlo[4], whi[4];
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    const int64_t i0 = blockIdx.x * (int64_t)SC_CH, i1 = i0 + SC_CH < N ? i0 + SC_CH : N;
    const int64_t s0 = i0 * S, s1 = i1 * S;
    int64_t mn = INT64_MAX, mx = -1;
    for (int64_t s = s0 + threadIdx.x; s < s1; s += blockDim.x) {
        const int64_t j = NL[s];
        if (j >= 0) { mn = j < mn ? j : mn; mx = j > mx ? j : mx; }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const int64_t a = __shfl_xor(mn, off, WAVE), b = __shfl_xor(mx, off, WAVE);
        mn = a < mn ? a : mn;
        mx = b > mx ? b : mx;
    }
    if (lane == 0) { wlo[wid] = mn; whi[wid] = mx; }
    __syncthreads();
    mn = wlo[0]; mx = whi[0];
    for (int w = 1; w < 4; w++) { mn = wlo[w] < mn ? wlo[w] : mn; mx = whi[w] > mx ? whi[w] : mx; }
    const int64_t lo = mn, hi = mx + 1;
    const bool lds = mx >= 0 && hi - lo <= rcap;
    if (lds)
        for (int64_t r = threadIdx.x; r < NV * (hi - lo); r += blockDim.x) acc[r] = T(0);
    __syncthreads();
    for (int64_t s = s0 + threadIdx.x; s < s1; s += blockDim.x) {
        const int64_t j = NL[s];
        if (j < 0) continue;
        const T *e = dE + 4 * s, *d = Ri_d + 12 * s;
        T v[NV];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const T f = e[0] * d[a] + e[1] * d[3 + a] + e[2] * d[6 + a] + e[3] * d[9 + a];
            if constexpr (NV == 3) {
                v[a] = f;
            } else {
#pragma unroll
                for (int c = 0; c < 3; c++) v[3 * a + c] = f * Rij[3 * s + c];
            }
        }
#pragma unroll
        for (int c = 0; c < NV; c++) {
            if (lds) atomicAdd(acc + NV * (j - lo) + c, v[c]);
            else atomic_add(out + NV * j + c, v[c]);
        }
    }
    __syncthreads();
    if (lds)
        for (int64_t r = threadIdx.x; r < NV * (hi - lo); r += blockDim.x) {
            const T x = acc[r];
            if (x != T(0)) atomic_add(out + NV * lo + r, x);
        }
}

// Centre term of the NEP force, F[i] = -sum_{k, w} dE[i][k][w] Ri_d[i][k][w][:],
// over all S slots (the product src/model/nep_net.py:608 forms with torch.matmul).
// One wavefront per atom; lanes over slots.
template <typename T>
__global__ void __launch_bounds__(256) nep_centre_kernel(const T *__restrict__ dE, const T *__restrict__ Ri_d,
                                                         int64_t N, int S, T *__restrict__ F) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t i = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / WAVE;
    if (i >= N) return;
    T f[3] = {T(0), T(0), T(0)};
    for (int k = lane; k < S; k += WAVE) {
        const int64_t s = i * S + k;
        const T *e = dE + 4 * s, *d = Ri_d + 12 * s;
#pragma unroll
        for (int a = 0; a < 3; a++) f[a] += e[0] * d[a] + e[1] * d[3 + a] + e[2] * d[6 + a] + e[3] * d[9 + a];
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const T v = wave_sum(f[a]);
        if (lane == 0) F[3 * i + a] = -v;
    }
}

// Per-frame virial only: sum over the frame's list entries (j >= 0) of
// (dE . Ri_d)_a Rij_c, the row sum of the per-atom virial without forming it.
// Two passes without atomics (deterministic): one wavefront per atom, lanes
// over slots, writes the atom's centre row rows[i][9]; nep_virial_reduce_kernel
// then sums each frame's rows.
template <typename T>
__global__ void __launch_bounds__(256) nep_centre_virial_kernel(const int64_t *__restrict__ NL,
                                                                const T *__restrict__ dE, const T *__restrict__ Rij,
                                                                const T *__restrict__ Ri_d, int64_t N, int S,
                                                                T *__restrict__ rows) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t i = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / WAVE;
    if (i >= N) return;
    T v[9];
#pragma unroll
    for (int c = 0; c < 9; c++) v[c] = T(0);
    for (int k = lane; k < S; k += WAVE) {
        const int64_t s = i * S + k;
        if (NL[s] < 0) continue;
        const T *e = dE + 4 * s, *d = Ri_d + 12 * s, *r = Rij + 3 * s;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const T fa = e[0] * d[a] + e[1] * d[3 + a] + e[2] * d[6 + a] + e[3] * d[9 + a];
#pragma unroll
            for (int c = 0; c < 3; c++) v[3 * a + c] += fa * r[c];
        }
    }
#pragma unroll
    for (int c = 0; c < 9; c++) {
        const T x = wave_sum(v[c]);
        if (lane == 0) rows[9 * i + c] = x;
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

// ------------------------------------------------------------------ fused force + frame virial
// F and the per-frame virial of the NEP model from dE and Ri = [r, x, y, z] directly: the position
// derivative Ri_d (row 0 = (x, y, z)/r, rows 1-3 = identity, zero where |r| <= 1e-5; nep_net.py:440-468)
// is formed in registers, so f(s) = dE[s][0] (x, y, z)/r + dE[s][1:4] per used slot s of atom i.
// calculateNepForce's terms (CalcOps.cpp:75-158): F[i] -= f(s), F[j] += f(s); calculateNepVirial's
// frame sum (calculateNepVirial.cu:19-96): V[frame] += f(s) (x) Rij[s].
// A block takes `ch` consecutive atoms i (one wavefront per atom at a time). With LDS copies
// (rcap > 0) it adds the neighbor terms of their slots into an LDS copy of F over the frames those
// atoms belong to (every j of a slot lies in its atom's frame), then writes that copy to its own row
// of `part`; fv_reduce_kernel adds the rows of the blocks covering each atom. A block whose frames
// exceed the LDS copy adds its terms to F with global atomics instead, as does any entry naming an
// atom outside its block's frames, and so does every block without LDS copies (rcap = 0).
// Unlike per-slot global atomics, which on this GPU are carried out beyond the per-XCD L2 caches, the
// LDS adds stay on the CU; the copies pay when many small frames fill the GPU with blocks (fv_plan).
constexpr int FV_CH = 64;        // atoms per block with LDS copies
constexpr int FV_CH_DIRECT = 16; // atoms per block without
constexpr int FV_WAVES = 16;     // one LDS copy shared by many wavefronts, each with few atoms

template <typename T>
struct V4 { T x, y, z, w; };

template <typename T>
__device__ __forceinline__ V4<T> ld4(const T *p) {
    if constexpr (sizeof(T) == 4) {
        const float4 v = *reinterpret_cast<const float4 *>(p);
        return {v.x, v.y, v.z, v.w};
    } else {
        const double2 a = reinterpret_cast<const double2 *>(p)[0], b = reinterpret_cast<const double2 *>(p)[1];
        return {a.x, a.y, b.x, b.y};
    }
}

// the frames of atoms [i0, i1): [lo, hi)
__device__ __forceinline__ void fv_range(const int64_t *frame_start, const int32_t *atom_frame, int64_t N, int64_t b,
                                         int ch, int64_t &lo, int64_t &hi) {
    const int64_t i0 = b * ch, i1 = i0 + ch < N ? i0 + ch : N;
    lo = frame_start[atom_frame[i0]];
    hi = frame_start[atom_frame[i1 - 1] + 1];
}

template <typename T>
__global__ void __launch_bounds__(FV_WAVES * WAVE)
nep_fv_kernel(const int64_t *__restrict__ NL, const T *__restrict__ dE, const T *__restrict__ Ri,
              const int64_t *__restrict__ frame_start, const int32_t *__restrict__ atom_frame, int64_t N, int S,
              int ch, int64_t rcap, T *__restrict__ F, T *__restrict__ part, T *__restrict__ rows) {
    extern __shared__ double fv_smem[];
    T *acc = reinterpret_cast<T *>(fv_smem);
    const int lane = threadIdx.x & (WAVE - 1), wid = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const int64_t b = blockIdx.x, i0 = b * ch, i1 = i0 + ch < N ? i0 + ch : N;
    int64_t lo, hi;
    fv_range(frame_start, atom_frame, N, b, ch, lo, hi);
    const bool lds = hi - lo <= rcap;
    if (lds)
        for (int64_t r = threadIdx.x; r < 3 * (hi - lo); r += blockDim.x) acc[r] = T(0);
    __syncthreads();
    for (int64_t i = i0 + wid; i < i1; i += FV_WAVES) {
        T c[3] = {T(0), T(0), T(0)}, v[9];
#pragma unroll
        for (int q = 0; q < 9; q++) v[q] = T(0);
        for (int k0 = 0; k0 < S; k0 += 2 * WAVE) {
            // the loads of two passes are issued before either is used
            int64_t jj[2];
            V4<T> qq[2], ee[2];
#pragma unroll
            for (int u = 0; u < 2; u++) {
                const int k = k0 + u * WAVE + lane;
                jj[u] = -1;
                if (k < S) {
                    const int64_t s = i * S + k;
                    jj[u] = NL[s];
                    qq[u] = ld4(Ri + 4 * s);
                    ee[u] = ld4(dE + 4 * s);
                }
            }
#pragma unroll
            for (int u = 0; u < 2; u++) {
                const int64_t j = jj[u];
                const V4<T> q = qq[u], e = ee[u];
                if (j < 0 || !(fabs(q.x) > T(1e-5))) continue;
                const T f[3] = {e.x * (q.y / q.x) + e.y, e.x * (q.z / q.x) + e.z, e.x * (q.w / q.x) + e.w};
                const T x[3] = {q.y, q.z, q.w};
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    c[a] += f[a];
#pragma unroll
                    for (int cc = 0; cc < 3; cc++) v[3 * a + cc] += f[a] * x[cc];
                    if (lds && j >= lo && j < hi) atomicAdd(acc + 3 * (j - lo) + a, f[a]);
                    else atomic_add(F + 3 * j + a, f[a]);
                }
            }
        }
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const T x = wave_sum(c[a]);
            if (lane == 0) {
                if (lds) atomicAdd(acc + 3 * (i - lo) + a, -x);
                else atomic_add(F + 3 * i + a, -x);
            }
        }
#pragma unroll
        for (int q = 0; q < 9; q++) {
            const T x = wave_sum(v[q]);
            if (lane == 0) rows[9 * i + q] = x;
        }
    }
    __syncthreads();
    if (lds)
        for (int64_t r = threadIdx.x; r < 3 * (hi - lo); r += blockDim.x) part[b * 3 * rcap + r] = acc[r];
}

// F[j] += the LDS copies of the blocks whose atoms share j's frame (blocks fv_range-covering it)
template <typename T>
__global__ void __launch_bounds__(256)
nep_fv_reduce_kernel(const int64_t *__restrict__ frame_start, const int32_t *__restrict__ atom_frame, int64_t N,
                     int ch, int64_t rcap, const T *__restrict__ part, T *__restrict__ F) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < N; j += (int64_t)gridDim.x * blockDim.x) {
        const int f = atom_frame[j];
        const int64_t fs = frame_start[f], fe = frame_start[f + 1];
        if (fe - fs > rcap) continue;             // no block covering this frame used LDS
        T s[3] = {T(0), T(0), T(0)};
        for (int64_t b = fs / ch; b <= (fe - 1) / ch; b++) {
            int64_t lo, hi;
            fv_range(frame_start, atom_frame, N, b, ch, lo, hi);
            if (hi - lo > rcap) continue;
            const T *p = part + b * 3 * rcap + 3 * (j - lo);
            s[0] += p[0]; s[1] += p[1]; s[2] += p[2];
        }
        F[3 * j] += s[0]; F[3 * j + 1] += s[1]; F[3 * j + 2] += s[2];
    }
}

// Backward of nep_fv_kernel wrt dE: with q = gF[j] - gF[i] + gV[frame(i)] Rij[s] (either upstream
// gradient may be absent), g[s] = [q . (x, y, z)/r, q] on used slots with |r| > 1e-5, else 0
// (calculateNepForce.cu:74-122 and the virial's transpose, in one pass). One wavefront per atom.
template <typename T>
__global__ void __launch_bounds__(256)
nep_fv_grad_kernel(const int64_t *__restrict__ NL, const T *__restrict__ Ri, const T *__restrict__ gF,
                   const T *__restrict__ gV, const int32_t *__restrict__ atom_frame, int64_t N, int S,
                   T *__restrict__ g) {
    const int lane = threadIdx.x & (WAVE - 1), wid = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    for (int64_t i = blockIdx.x * (int64_t)(blockDim.x / WAVE) + wid; i < N;
         i += (int64_t)gridDim.x * (blockDim.x / WAVE)) {
        T fi[3] = {T(0), T(0), T(0)}, gv[9];
        if (gF) { fi[0] = gF[3 * i]; fi[1] = gF[3 * i + 1]; fi[2] = gF[3 * i + 2]; }
#pragma unroll
        for (int q = 0; q < 9; q++) gv[q] = gV ? gV[9 * (int64_t)atom_frame[i] + q] : T(0);
        for (int k = lane; k < S; k += WAVE) {
            const int64_t s = i * S + k;
            const int64_t j = NL[s];
            const V4<T> r = ld4(Ri + 4 * s);
            T o[4] = {T(0), T(0), T(0), T(0)};
            if (j >= 0 && fabs(r.x) > T(1e-5)) {
                const T x[3] = {r.y, r.z, r.w};
                T q[3];
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    q[a] = gv[3 * a] * x[0] + gv[3 * a + 1] * x[1] + gv[3 * a + 2] * x[2];
                    if (gF) q[a] += gF[3 * j + a] - fi[a];
                }
                o[0] = q[0] * (x[0] / r.x) + q[1] * (x[1] / r.x) + q[2] * (x[2] / r.x);
                o[1] = q[0]; o[2] = q[1]; o[3] = q[2];
            }
            if constexpr (sizeof(T) == 4) {
                *reinterpret_cast<float4 *>(g + 4 * s) = make_float4(o[0], o[1], o[2], o[3]);
            } else {
                reinterpret_cast<double2 *>(g + 4 * s)[0] = make_double2(o[0], o[1]);
                reinterpret_cast<double2 *>(g + 4 * s)[1] = make_double2(o[2], o[3]);
            }
        }
    }
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

// The angular descriptor's first derivative (mode 1: gRi and dL/dc), batched like the features: per
// batch the dE/ds rows (B) of its atoms in LDS, then per slot (lanes) gRi = sum_nm D_nm d(g_n b_m)/dRi
// and the rows of dc[n][q] = f_q (D b)_n, added with lanes over (n, q) into the table row of each
// slot's (centre type, neighbor type).
__host__ __device__ inline int angb_wave_elems(int nm, int nb) { return AB_ATOMS * nm * NLM + WAVE * (nm + nb); }

template <typename T>
__global__ void __launch_bounds__(ANG_WAVES * WAVE)
nep_angular_bwd_kernel(FeatArgs a) {
    extern __shared__ unsigned char smem_raw[];
    const int lane = threadIdx.x & (WAVE - 1);
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const int ncoef = a.T * a.T * a.nm * a.nb;
    const int per = a.nm * NLM;
    const int nm = a.nm, nb = a.nb, nmb = nm * nb;
    const int ntab = a.pw ? ANG_WAVES : 1;
    T *gcl = reinterpret_cast<T *>(smem_raw);                                  // [ntab][ncoef]
    T *gtab = gcl + (ntab > 1 ? wid : 0) * ncoef;                             // this wave's table
    T *Bq = gcl + (size_t)ntab * ncoef + (size_t)wid * angb_wave_elems(nm, nb);   // [AB_ATOMS][nm][24]
    T *u1 = Bq + AB_ATOMS * per;                                               // [WAVE][nm]
    T *v1 = u1 + WAVE * nm;                                                    // [WAVE][nb]
    int *trow = reinterpret_cast<int *>(gcl + (size_t)ntab * ncoef + (size_t)ANG_WAVES * angb_wave_elems(nm, nb)) +
                wid * WAVE;                                                    // [WAVE] table row per slot
    for (int x = threadIdx.x; x < ntab * ncoef; x += blockDim.x) gcl[x] = T(0);
    __syncthreads();
    const T *c = reinterpret_cast<const T *>(a.c);
    const T *Ri = reinterpret_cast<const T *>(a.Ri);
    const T *gfe = reinterpret_cast<const T *>(a.gfeat);
    T *gRiA = reinterpret_cast<T *>(a.gRi);
    const int nf = a.l3 + (a.l4 > 0) + (a.l5 > 0);
    // this wavefront's contiguous range of atoms
    const int64_t nw = (int64_t)gridDim.x * ANG_WAVES, gw = (int64_t)blockIdx.x * ANG_WAVES + wid;
    const int64_t i1 = a.N * (gw + 1) / nw;
    for (int64_t i = a.N * gw / nw; i < i1;) {
        const int64_t il = i + (lane < AB_ATOMS ? lane : 0);
        const int c_l = (lane < AB_ATOMS && il < i1) ? slot_count(a.NL, il, a.S) : -1;
        int off[AB_ATOMS + 1];
        off[0] = 0;
        int na = 0;
        bool open = true;
#pragma unroll
        for (int q = 0; q < AB_ATOMS; q++) {
            const int cq = __builtin_amdgcn_readlane(c_l, q);
            open = open && (q == 0 || (cq >= 0 && off[q] + cq <= WAVE));
            if (open) na = q + 1;
            off[q + 1] = off[q] + (open ? cq : 0);
        }
        const int tot = off[AB_ATOMS];
        // unused slots of each atom get gRi = 0
        for (int q = 0; q < na; q++) {
            const int64_t ia = i + q;
            for (int k = off[q + 1] - off[q] + lane; k < a.S; k += WAVE) {
                T *g = gRiA + (ia * a.S + k) * 4;
                g[0] = g[1] = g[2] = g[3] = T(0);
            }
        }
        // B = D = dE/ds of each atom of the batch (lanes over (atom, n))
        if (lane < na * nm) {
            const int q = lane / nm, n = lane - q * nm;
            const int64_t ia = i + q;
            T gq[6];
#pragma unroll
            for (int f = 0; f < 6; f++) gq[f] = f < nf ? gfe[ia * a.ld + a.col0 + aidx(f, n, nm)] : T(0);
            T D[NLM];
            dq_ds<T, T>(reinterpret_cast<const T *>(a.s_save) + ia * per + n * NLM, gq, a.l3, a.l4, a.l5, D);
            for (int m = 0; m < NLM; m++) Bq[q * per + n * NLM + m] = D[m];
        }
        __builtin_amdgcn_wave_barrier();
        for (int e0 = 0; e0 < tot; e0 += WAVE) {
            const int nk = tot - e0 < WAVE ? tot - e0 : WAVE;
            const int e = e0 + lane;
            if (lane < nk) {
                int q = 0;
#pragma unroll
                for (int qq = 1; qq < AB_ATOMS; qq++) q += (e >= off[qq]) ? 1 : 0;
                const int64_t ia = i + q;
                const int k = e - off[q];
                const int ti = (int)a.atom_map[ia];
                const int tj = (int)a.atom_map[a.NL[ia * a.S + k]];
                const T *er = Ri + (ia * a.S + k) * 4;
                const T r = er[0], ir = T(1) / r;
                const T ux = er[1] * ir, uy = er[2] * ir, uz = er[3] * ir;
                const T *cij = c + ((int64_t)(ti * a.T + tj) * nm) * nb;
                trow[lane] = ti * a.T + tj;
                // g_n, g'_n over the streamed basis, f_q to this slot's v1 row
                T gn[NM_MAX], gnp[NM_MAX];
                {
                    Basis<T> bs(r, (T)a.rc);
#pragma unroll
                    for (int n = 0; n < NM_MAX; n++) { gn[n] = T(0); gnp[n] = T(0); }
                    for (int qb = 0; qb < nb; qb++) {
                        T f, fp;
                        bs.next(f, fp);
#pragma unroll
                        for (int n = 0; n < NM_MAX; n++)
                            if (n < nm) { gn[n] += cij[n * nb + qb] * f; gnp[n] += cij[n * nb + qb] * fp; }
                        v1[lane * nb + qb] = f;
                    }
                }
                // gRi = sum_nm D_nm d(g_n b_m)/dRi; dc[n][q] = f_q (D b)_n
                const T *B = Bq + q * per;
                T b[NLM];
                blm_each<T>(ux, uy, uz, [&](int m, T bv, T, T, T) { b[m] = bv; });
                T Dg[NLM], Db[NM_MAX];
#pragma unroll
                for (int n = 0; n < NM_MAX; n++) Db[n] = T(0);
                T gr0 = T(0);
#pragma unroll
                for (int m = 0; m < NLM; m++) {
                    T dg = T(0), dgp = T(0);
#pragma unroll
                    for (int n = 0; n < NM_MAX; n++)
                        if (n < nm) {
                            const T Bv = B[n * NLM + m];
                            dg += Bv * gn[n];
                            dgp += Bv * gnp[n];
                            Db[n] += Bv * b[m];
                        }
                    Dg[m] = dg;
                    gr0 += dgp * b[m];
                }
#pragma unroll
                for (int n = 0; n < NM_MAX; n++)
                    if (n < nm) u1[lane * nm + n] = Db[n];
                T V0 = T(0), V1 = T(0), V2 = T(0);
                blm_each<T>(ux, uy, uz, [&](int m, T, T d0, T d1, T d2) {
                    V0 += Dg[m] * d0; V1 += Dg[m] * d1; V2 += Dg[m] * d2;
                });
                T *g = gRiA + (ia * a.S + k) * 4;
                g[0] = gr0 - (ux * V0 + uy * V1 + uz * V2) * ir;
                g[1] = V0 * ir;
                g[2] = V1 * ir;
                g[3] = V2 * ir;
            }
            __builtin_amdgcn_wave_barrier();
            // dc[row of slot kk][n][q] += u1[kk][n] v1[kk][q]: lanes over (n, q), no two lanes on one entry
            for (int o = lane; o < nmb; o += WAVE) {
                const int n = o / nb, qb = o - n * nb;
                for (int kk = 0; kk < nk; kk++) {
                    const T d = u1[kk * nm + n] * v1[kk * nb + qb];
                    if (ntab > 1) gtab[trow[kk] * nmb + o] += d;
                    else atomicAdd(gtab + trow[kk] * nmb + o, d);
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
        i += na;
    }
    block_flush_add(gcl, reinterpret_cast<T *>(a.gc), ncoef, ntab);
}

// The angular descriptor's second-order pass (mode 2: ggfeat and dL/dc) with the slots of up to AB
// atoms in one chunk of CHB lanes, as nep_angular_bwd_kernel does for mode 1: per batch the dE/ds
// rows (B) of its atoms, a first pass over the slots (the directional derivative P of each atom's
// s, summed per atom with lanes over (n, m), and the explicit coefficient gradient), then per atom
// ggfeat = J P and E = H P, and a second pass (the coefficient gradient through s). Float takes 48
// slots of up to 4 atoms; double keeps 16 slots of one atom (its per-slot rows, 96 values of 8
// bytes, leave no room for more in LDS).
template <typename T>
struct Ang2Shape {
    static constexpr int CHB = sizeof(T) == 4 ? 48 : 16;
    static constexpr int AB = sizeof(T) == 4 ? 4 : 1;
};

template <typename T>
__host__ __device__ inline int angb2_wave_elems(int nm, int nb) {
    return Ang2Shape<T>::AB * 2 * nm * NLM + Ang2Shape<T>::CHB * (4 * nm + 2 * WS + 2 * nb);
}

template <typename T>
__global__ void __launch_bounds__(ANG_WAVES * WAVE)
nep_angular_bwd2_kernel(FeatArgs a) {
    constexpr int CHB = Ang2Shape<T>::CHB, AB = Ang2Shape<T>::AB;
    constexpr int NRED = (NM_MAX * NLM + WAVE - 1) / WAVE;
    extern __shared__ unsigned char smem_raw[];
    const int lane = threadIdx.x & (WAVE - 1);
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const int ncoef = a.T * a.T * a.nm * a.nb;
    const int per = a.nm * NLM;
    const int nm = a.nm, nb = a.nb, nmb = nm * nb;
    const int ntab = a.pw ? ANG_WAVES : 1;
    T *gcl = reinterpret_cast<T *>(smem_raw);                                  // [ntab][ncoef]
    T *gtab = gcl + (ntab > 1 ? wid : 0) * ncoef;                             // this wave's table
    T *Bq = gcl + (size_t)ntab * ncoef + (size_t)wid * angb2_wave_elems<T>(nm, nb);   // [AB][nm][24] D
    T *Cq = Bq + AB * per;                                                    // [AB][nm][24] P, then E
    T *u1 = Cq + AB * per, *u2 = u1 + CHB * nm;                                // g'_n, g_n (pass 2: (E b)_n)
    T *w1 = u2 + CHB * nm, *w2 = w1 + CHB * WS;                                // beta_m, delta_m
    T *z1 = w2 + CHB * WS, *z2 = z1 + CHB * nm;                                // (B beta)_n, (B delta)_n
    T *v1 = z2 + CHB * nm, *v2 = v1 + CHB * nb;                                // f'_q, f_q (pass 2: f_q)
    int *trow = reinterpret_cast<int *>(gcl + (size_t)ntab * ncoef + (size_t)ANG_WAVES * angb2_wave_elems<T>(nm, nb)) +
                wid * CHB;                                                     // [CHB] table row per slot
    for (int x = threadIdx.x; x < ntab * ncoef; x += blockDim.x) gcl[x] = T(0);
    __syncthreads();
    const T *c = reinterpret_cast<const T *>(a.c);
    const T *Ri = reinterpret_cast<const T *>(a.Ri);
    const T *gfe = reinterpret_cast<const T *>(a.gfeat);
    const T *sv = reinterpret_cast<const T *>(a.s_save);
    T *go = reinterpret_cast<T *>(a.ggfeat);
    const int nf = a.l3 + (a.l4 > 0) + (a.l5 > 0);
    // dc[row of slot kk][n][q] += p1[kk][n] q1[kk][q] (+ p2 q2): lanes over (n, q)
    auto add_dc_rows = [&](int nk, const T *p1, const T *q1, const T *p2, const T *q2) {
        for (int o = lane; o < nmb; o += WAVE) {
            const int n = o / nb, qb = o - n * nb;
            for (int kk = 0; kk < nk; kk++) {
                T d = p1[kk * nm + n] * q1[kk * nb + qb];
                if (p2) d += p2[kk * nm + n] * q2[kk * nb + qb];
                if (ntab > 1) gtab[trow[kk] * nmb + o] += d;
                else atomicAdd(gtab + trow[kk] * nmb + o, d);
            }
        }
    };
    const int64_t nw = (int64_t)gridDim.x * ANG_WAVES, gw = (int64_t)blockIdx.x * ANG_WAVES + wid;
    const int64_t i1 = a.N * (gw + 1) / nw;
    for (int64_t i = a.N * gw / nw; i < i1;) {
        const int64_t il = i + (lane < AB ? lane : 0);
        const int c_l = (lane < AB && il < i1) ? slot_count(a.NL, il, a.S) : -1;
        int off[AB + 1];
        off[0] = 0;
        int na = 0;
        bool open = true;
#pragma unroll
        for (int q = 0; q < AB; q++) {
            const int cq = __builtin_amdgcn_readlane(c_l, q);
            open = open && (q == 0 || (cq >= 0 && off[q] + cq <= CHB));
            if (open) na = q + 1;
            off[q + 1] = off[q] + (open ? cq : 0);
        }
        const int tot = off[AB];
        auto atom_of = [&](int e) {
            int q = 0;
#pragma unroll
            for (int qq = 1; qq < AB; qq++) q += (e >= off[qq]) ? 1 : 0;
            return q;
        };
        // B = D = dE/ds of each atom (lanes over (atom, n))
        if (lane < na * nm) {
            const int q = lane / nm, n = lane - q * nm;
            const int64_t ia = i + q;
            T gq[6];
#pragma unroll
            for (int f = 0; f < 6; f++) gq[f] = f < nf ? gfe[ia * a.ld + a.col0 + aidx(f, n, nm)] : T(0);
            T D[NLM];
            dq_ds<T, T>(sv + ia * per + n * NLM, gq, a.l3, a.l4, a.l5, D);
            for (int m = 0; m < NLM; m++) Bq[q * per + n * NLM + m] = D[m];
        }
        __builtin_amdgcn_wave_barrier();
        T red[AB][NRED];
#pragma unroll
        for (int q = 0; q < AB; q++)
#pragma unroll
            for (int j = 0; j < NRED; j++) red[q][j] = T(0);
        // pass 1: P_nm = sum_slots gg . d(g_n b_m)/dRi = g'_n beta_m + g_n delta_m with beta_m = gg0 b_m,
        // delta_m = (db_m/du . (gg_xyz - gg0 u)) / r; explicit c-gradient
        // sum_m B_nm dP_nm/dc[n][q] = f'_q (B beta)_n + f_q (B delta)_n
        for (int e0 = 0; e0 < tot; e0 += CHB) {
            const int nk = tot - e0 < CHB ? tot - e0 : CHB;
            if (lane < nk) {
                const int e = e0 + lane, q = atom_of(e);
                const int64_t ia = i + q;
                const int k = e - off[q];
                const int ti = (int)a.atom_map[ia];
                const int tj = (int)a.atom_map[a.NL[ia * a.S + k]];
                const T *er = Ri + (ia * a.S + k) * 4;
                const T r = er[0], ir = T(1) / r;
                const T ux = er[1] * ir, uy = er[2] * ir, uz = er[3] * ir;
                const T *cij = c + ((int64_t)(ti * a.T + tj) * nm) * nb;
                trow[lane] = ti * a.T + tj;
                const T *gg = reinterpret_cast<const T *>(a.gg) + (ia * a.S + k) * 4;
                const T g0 = gg[0];
                const T gx = gg[1] - g0 * ux, gy = gg[2] - g0 * uy, gz = gg[3] - g0 * uz;
                T gn[NM_MAX], gnp[NM_MAX];
                {
                    Basis<T> bs(r, (T)a.rc);
#pragma unroll
                    for (int n = 0; n < NM_MAX; n++) { gn[n] = T(0); gnp[n] = T(0); }
                    for (int qb = 0; qb < nb; qb++) {
                        T f, fp;
                        bs.next(f, fp);
#pragma unroll
                        for (int n = 0; n < NM_MAX; n++)
                            if (n < nm) { gn[n] += cij[n * nb + qb] * f; gnp[n] += cij[n * nb + qb] * fp; }
                        v2[lane * nb + qb] = f;
                        v1[lane * nb + qb] = fp;
                    }
                }
                T be[NLM], de[NLM];
                blm_each<T>(ux, uy, uz, [&](int m, T bv, T d0, T d1, T d2) {
                    be[m] = g0 * bv;
                    de[m] = (d0 * gx + d1 * gy + d2 * gz) * ir;
                });
#pragma unroll
                for (int m = 0; m < NLM; m++) {
                    w1[lane * WS + m] = be[m];
                    w2[lane * WS + m] = de[m];
                }
                const T *B = Bq + q * per;
#pragma unroll
                for (int n = 0; n < NM_MAX; n++)
                    if (n < nm) {
                        u1[lane * nm + n] = gnp[n];
                        u2[lane * nm + n] = gn[n];
                        T sb = T(0), sd = T(0);
#pragma unroll
                        for (int m = 0; m < NLM; m++) {
                            sb += B[n * NLM + m] * be[m];
                            sd += B[n * NLM + m] * de[m];
                        }
                        z1[lane * nm + n] = sb;
                        z2[lane * nm + n] = sd;
                    }
            }
            __builtin_amdgcn_wave_barrier();
            // P of each atom over its slots of this chunk: lanes own (n, m) outputs
#pragma unroll
            for (int q = 0; q < AB; q++) {
                if (q >= na) break;
                const int k0 = off[q] - e0 > 0 ? off[q] - e0 : 0;
                const int k1 = off[q + 1] - e0 < nk ? off[q + 1] - e0 : nk;
#pragma unroll
                for (int j = 0; j < NRED; j++) {
                    const int o = lane + j * WAVE;
                    if (o < per) {
                        const int n = o / NLM, m = o - n * NLM;
                        T sacc = T(0);
                        for (int kk = k0; kk < k1; kk++)
                            sacc += u1[kk * nm + n] * w1[kk * WS + m] + u2[kk * nm + n] * w2[kk * WS + m];
                        red[q][j] += sacc;
                    }
                }
            }
            add_dc_rows(nk, z1, v1, z2, v2);
            __builtin_amdgcn_wave_barrier();
        }
        // per atom: ggfeat = J P (directional derivative of q along P); E = H(gq) P
#pragma unroll
        for (int q = 0; q < AB; q++) {
            if (q >= na) break;
#pragma unroll
            for (int j = 0; j < NRED; j++)
                if (lane + j * WAVE < per) Cq[q * per + lane + j * WAVE] = red[q][j];
        }
        __builtin_amdgcn_wave_barrier();
        if (lane < na * nm) {
            const int q = lane / nm, n = lane - q * nm;
            const int64_t ia = i + q;
            Dual<T> sd[NLM];
            for (int m = 0; m < NLM; m++) sd[m] = Dual<T>(sv[ia * per + n * NLM + m], Cq[q * per + n * NLM + m]);
            Dual<T> qv[6];
            const int nq = q_of_s<T, Dual<T>>(sd, a.l3, a.l4, a.l5, qv);
#pragma unroll
            for (int f = 0; f < 6; f++)
                if (f < nq) go[ia * a.ld + a.col0 + aidx(f, n, nm)] = qv[f].e;
            T gq[6];
#pragma unroll
            for (int f = 0; f < 6; f++) gq[f] = f < nf ? gfe[ia * a.ld + a.col0 + aidx(f, n, nm)] : T(0);
            Dual<T> D[NLM];
            dq_ds<T, Dual<T>>(sd, gq, a.l3, a.l4, a.l5, D);
            for (int m = 0; m < NLM; m++) Cq[q * per + n * NLM + m] = D[m].e;     // E = (H P)_m
        }
        __builtin_amdgcn_wave_barrier();
        // pass 2: c-gradient through s, dc[ti][tj][n][q] += (E b)_n f_q over the slots
        for (int e0 = 0; e0 < tot; e0 += CHB) {
            const int nk = tot - e0 < CHB ? tot - e0 : CHB;
            if (lane < nk) {
                const int e = e0 + lane, q = atom_of(e);
                const int64_t ia = i + q;
                const int k = e - off[q];
                const int ti = (int)a.atom_map[ia];
                trow[lane] = ti * a.T + (int)a.atom_map[a.NL[ia * a.S + k]];
                const T *er = Ri + (ia * a.S + k) * 4;
                const T ir = T(1) / er[0];
                Basis<T> bs(er[0], (T)a.rc);
                for (int qb = 0; qb < nb; qb++) {
                    T f, fp;
                    bs.next(f, fp);
                    v1[lane * nb + qb] = f;
                }
                T b[NLM];
                blm_each<T>(er[1] * ir, er[2] * ir, er[3] * ir, [&](int m, T bv, T, T, T) { b[m] = bv; });
                const T *E = Cq + q * per;
#pragma unroll
                for (int n = 0; n < NM_MAX; n++)
                    if (n < nm) {
                        T Eb = T(0);
#pragma unroll
                        for (int m = 0; m < NLM; m++) Eb += E[n * NLM + m] * b[m];
                        u1[lane * nm + n] = Eb;
                    }
            }
            __builtin_amdgcn_wave_barrier();
            add_dc_rows(nk, u1, v1, nullptr, nullptr);
            __builtin_amdgcn_wave_barrier();
        }
        i += na;
    }
    block_flush_add(gcl, reinterpret_cast<T *>(a.gc), ncoef, ntab);
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
    // LDS of the block for ntab coefficient-gradient tables
    auto lds_of = [&](size_t ntab) -> size_t {
        if (!angular)
            return mode > 0 ? (ntab * ncoef + (size_t)NEP_WAVES * RCH * a.nb) * sizeof(T) +
                                  (size_t)NEP_WAVES * RCH * sizeof(int)
                            : 0;
        if (mode == 1)
            return (ntab * ncoef + (size_t)ANG_WAVES * angb_wave_elems(a.nm, a.nb)) * sizeof(T) +
                   (size_t)ANG_WAVES * WAVE * sizeof(int);
        if (mode == 2)
            return (ntab * ncoef + (size_t)ANG_WAVES * angb2_wave_elems<T>(a.nm, a.nb)) * sizeof(T) +
                   (size_t)ANG_WAVES * Ang2Shape<T>::CHB * sizeof(int);
        return (size_t)ANG_WAVES * angf_wave_elems(a.nm) * sizeof(T);
    };
    // float: a table per wavefront (plain adds) when they fit, else one per block (LDS atomics).
    // Measured on the NEP training batch: float LDS atomic adds keep the LDS busy 3-4x longer than
    // double ones for the same instruction count (the per-wave tables take the float step from
    // 9.46 to 9.05 ms), while for double the larger tables cost occupancy (10.3 -> 11.8 ms).
    const size_t nw = angular ? ANG_WAVES : NEP_WAVES;
    a.pw = (mode > 0 && sizeof(T) == 4 && lds_of(nw) <= 64 * 1024) ? 1 : 0;
    const size_t lds = lds_of(a.pw ? nw : 1);
    MATPL_CHECK_ARG(lds <= 64 * 1024, "nep_feat: LDS %zu too large", lds);
    const int waves = angular ? ANG_WAVES : NEP_WAVES;
    int64_t blocks = (a.N + waves - 1) / waves;
    // gradient modes: a bounded grid (4,096 wavefronts, 16 per CU) keeps the global coefficient
    // atomics few (one per table entry and block) while hiding the kernels' memory waits (on the
    // NEP training batch 2,048 wavefronts left them waiting ~65 % of their cycles: float step 8.9
    // -> 8.3 ms at 4,096; 8,192 gives the same time with twice the atomics)
    if (mode > 0 && blocks > 4096 / waves) blocks = 4096 / waves;
    if (blocks > 8192) blocks = 8192;
    dim3 grid((unsigned)blocks), block(waves * WAVE);
    if (!angular) {
        if (mode == 0) hipLaunchKernelGGL((nep_radial_kernel<T, 0>), grid, block, lds, st, a);
        else if (mode == 1) hipLaunchKernelGGL((nep_radial_kernel<T, 1>), grid, block, lds, st, a);
        else hipLaunchKernelGGL((nep_radial_kernel<T, 2>), grid, block, lds, st, a);
        return check_launch("nep_radial_kernel");
    }
    if (mode == 1) {
        hipLaunchKernelGGL((nep_angular_bwd_kernel<T>), grid, block, lds, st, a);
        return check_launch("nep_angular_bwd_kernel");
    }
    if (mode == 2) {
        hipLaunchKernelGGL((nep_angular_bwd2_kernel<T>), grid, block, lds, st, a);
        return check_launch("nep_angular_bwd2_kernel");
    }
    hipLaunchKernelGGL((nep_angular_fwd_kernel<T>), grid, block, lds, st, a);
    return check_launch("nep_angular_fwd_kernel");
}

template <typename T, int NV>
int scatter_launch(const int64_t *NL, const T *dE, const T *Ri_d, const T *Rij, T *out, int64_t N, int S,
                   hipStream_t st) {
    if (N * S == 0) return OK;
    // LDS copy of at most 63 KB, so that with the kernel's own 64 bytes the block stays within 64 KB
    // (and of no more atoms than there are)
    int64_t rcap = 64512 / (NV * (int64_t)sizeof(T));
    if (rcap > N) rcap = N;
    hipLaunchKernelGGL((nep_scatter_kernel<T, NV>), dim3(ceil_div(N, SC_CH)), dim3(256), NV * rcap * sizeof(T), st, NL,
                       dE, Ri_d, Rij, out, N, S, rcap);
    return check_launch("nep_scatter_kernel");
}

template <typename T>
int virial_launch(const int64_t *NL, const T *dE, const T *Rij, const T *Ri_d, const int64_t *frame_start,
                  int64_t nframes, int64_t N, int S, T *virial, T *atom_virial, hipStream_t st) {
    int rc = fill_async<T>(atom_virial, T(0), N * 9, st);
    if (rc) return rc;
    if ((rc = scatter_launch<T, 9>(NL, dE, Ri_d, Rij, atom_virial, N, S, st))) return rc;
    if (nframes == 0) return OK;
    hipLaunchKernelGGL(nep_virial_reduce_kernel<T>, dim3(9, (unsigned)nframes), dim3(256), 0, st, atom_virial,
                       frame_start, virial);
    return check_launch("nep_virial_reduce_kernel");
}

template <typename T>
int centre_launch(const T *dE, const T *Ri_d, int64_t N, int S, T *F, hipStream_t st) {
    if (N == 0) return OK;
    if (S == 0) return fill_async<T>(F, T(0), N * 3, st);
    hipLaunchKernelGGL(nep_centre_kernel<T>, dim3((unsigned)ceil_div(N, 4)), dim3(256), 0, st, dE, Ri_d, N, S, F);
    return check_launch("nep_centre_kernel");
}

template <typename T>
int frame_virial_launch(const int64_t *NL, const T *dE, const T *Rij, const T *Ri_d, const int64_t *frame_start,
                        int64_t nframes, int64_t N, int S, T *rows, T *virial, hipStream_t st) {
    if (N == 0 || S == 0) return fill_async<T>(virial, T(0), nframes * 9, st);
    hipLaunchKernelGGL(nep_centre_virial_kernel<T>, dim3((unsigned)ceil_div(N, 4)), dim3(256), 0, st, NL, dE, Rij,
                       Ri_d, N, S, rows);
    int rc = check_launch("nep_centre_virial_kernel");
    if (rc) return rc;
    hipLaunchKernelGGL(nep_virial_reduce_kernel<T>, dim3(9, (unsigned)nframes), dim3(256), 0, st, rows, frame_start,
                       virial);
    return check_launch("nep_virial_reduce_kernel");
}

}  // namespace
}  // namespace matpl

using namespace matpl;

MATPL_EXPORT int matpl_nep_centre_force(int f64, const void *dE, const void *Ri_d, int64_t N, int S, void *F,
                                        void *stream) {
    MATPL_CHECK_ARG(N >= 0 && S >= 0, "nep_centre_force: bad sizes");
    if (N == 0) return OK;
    MATPL_CHECK_ARG(F && (S == 0 || (dE && Ri_d)), "nep_centre_force: null pointer");
    hipStream_t st = (hipStream_t)stream;
    return f64 ? centre_launch<double>((const double *)dE, (const double *)Ri_d, N, S, (double *)F, st)
               : centre_launch<float>((const float *)dE, (const float *)Ri_d, N, S, (float *)F, st);
}

MATPL_EXPORT int matpl_nep_frame_virial(int f64, const int64_t *NL, const void *dE, const void *Rij, const void *Ri_d,
                                        const int64_t *frame_start, int64_t nframes, int64_t N, int S, void *rows,
                                        void *virial, void *stream) {
    MATPL_CHECK_ARG(N >= 0 && S >= 0 && nframes >= 0, "nep_frame_virial: bad sizes");
    if (nframes == 0) return OK;
    MATPL_CHECK_ARG(virial && (N == 0 || S == 0 || (NL && dE && Rij && Ri_d && frame_start && rows)),
                    "nep_frame_virial: null pointer");
    hipStream_t st = (hipStream_t)stream;
    return f64 ? frame_virial_launch<double>(NL, (const double *)dE, (const double *)Rij, (const double *)Ri_d,
                                             frame_start, nframes, N, S, (double *)rows, (double *)virial, st)
               : frame_virial_launch<float>(NL, (const float *)dE, (const float *)Rij, (const float *)Ri_d,
                                            frame_start, nframes, N, S, (float *)rows, (float *)virial, st);
}

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
    const dim3 grid(ceil_div(n, 4)), block(256);
    hipStream_t st = (hipStream_t)stream;
    if (mode == 0)
        hipLaunchKernelGGL(nep_neighbor_kernel<0>, grid, block, 0, st, g, frame_start, atom_frame, type_map, pos, rc_r,
                           rc_a, maxr, maxa, n, ntypes, rij4, NN_r, NL_r, NN_a, NL_a, Ri_r, Ri_a, status);
    else if (mode == 1)
        hipLaunchKernelGGL(nep_neighbor_kernel<1>, grid, block, 0, st, g, frame_start, atom_frame, type_map, pos, rc_r,
                           rc_a, maxr, maxa, n, ntypes, rij4, NN_r, NL_r, NN_a, NL_a, Ri_r, Ri_a, status);
    else
        hipLaunchKernelGGL(nep_neighbor_kernel<2>, grid, block, 0, st, g, frame_start, atom_frame, type_map, pos, rc_r,
                           rc_a, maxr, maxa, n, ntypes, rij4, NN_r, NL_r, NN_a, NL_a, Ri_r, Ri_a, status);
    return check_launch("nep_neighbor_kernel");
}

// matpl_nep_neighbor's lists (mode 0) with the [r, x, y, z] rows in float: each value the FP64 one
// rounded once, as casting matpl_nep_neighbor's double rows would give (for float models).
MATPL_EXPORT int matpl_nep_neighbor_ri32(const double *box, const double *box_orig, const int64_t *num_cell,
                                         const int64_t *frame_start, const int32_t *atom_frame, const double *pos,
                                         double rc_r, double rc_a, int64_t maxr, int64_t maxa, int64_t n, int rij4,
                                         int64_t *NN_r, int64_t *NL_r, int64_t *NN_a, int64_t *NL_a, float *Ri_r,
                                         float *Ri_a, int32_t *status, void *stream) {
    MATPL_CHECK_ARG(n >= 0, "nep_neighbor_ri32: bad size");
    if (n == 0) return OK;
    MATPL_CHECK_ARG(box && box_orig && num_cell && frame_start && atom_frame && pos && NN_r && NN_a && NL_r && NL_a &&
                        Ri_r && Ri_a && status,
                    "nep_neighbor_ri32: null pointer");
    NepGeo g{box, box_orig, num_cell};
    hipLaunchKernelGGL((nep_neighbor_kernel<0, float>), dim3(ceil_div(n, 4)), dim3(256), 0, (hipStream_t)stream, g,
                       frame_start, atom_frame, nullptr, pos, rc_r, rc_a, maxr, maxa, n, 0, rij4, NN_r, NL_r, NN_a,
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
    if (f64)
        return scatter_launch<double, 3>(NL, (const double *)dE, (const double *)Ri_d, nullptr, (double *)F, N, S, st);
    return scatter_launch<float, 3>(NL, (const float *)dE, (const float *)Ri_d, nullptr, (float *)F, N, S, st);
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

// LDS copies of F (blocks of FV_CH atoms) when the blocks fill the GPU and every copy spans at most
// the largest span of frames a block can touch (two partial frames and whole ones in between) within
// 64 KB; otherwise blocks of FV_CH_DIRECT atoms and global atomics (a few large frames: a block's
// slots name few of a frame's atoms, and few blocks would share many copies per atom).
struct FvPlan {
    int ch;
    int64_t rcap;
};
inline FvPlan fv_plan(int64_t N, int64_t max_frame, size_t tsize) {
    const int64_t cap = (int64_t)(65536 / (3 * tsize));
    const int64_t want = 2 * max_frame + FV_CH;
    if (N >= 256 * (int64_t)FV_CH && max_frame <= 2048 && want <= cap) return {FV_CH, want};
    return {FV_CH_DIRECT, 0};
}

MATPL_EXPORT int64_t matpl_nep_force_virial_ri_work(int f64, int64_t N, int64_t max_frame) {
    if (N <= 0) return 0;
    const FvPlan pl = fv_plan(N, max_frame, f64 ? 8 : 4);
    return 9 * N + (int64_t)ceil_div(N, pl.ch) * 3 * pl.rcap;
}

template <typename T>
int fv_launch(const int64_t *NL, const T *dE, const T *Ri, const int64_t *frame_start, const int32_t *atom_frame,
              int64_t nframes, int64_t N, int S, int64_t max_frame, T *F, T *virial, T *work, hipStream_t st) {
    int rc = fill_async<T>(F, T(0), 3 * N, st);
    if (rc) return rc;
    if (N == 0 || S == 0) return fill_async<T>(virial, T(0), 9 * nframes, st);
    const FvPlan pl = fv_plan(N, max_frame, sizeof(T));
    T *rows = work, *part = work + 9 * N;
    const int64_t blocks = ceil_div(N, pl.ch);
    hipLaunchKernelGGL(nep_fv_kernel<T>, dim3((unsigned)blocks), dim3(FV_WAVES * WAVE), 3 * pl.rcap * sizeof(T), st,
                       NL, dE, Ri, frame_start, atom_frame, N, S, pl.ch, pl.rcap, F, part, rows);
    if ((rc = check_launch("nep_fv_kernel"))) return rc;
    if (pl.rcap > 0) {
        hipLaunchKernelGGL(nep_fv_reduce_kernel<T>, dim3(grid_cap(N)), dim3(256), 0, st, frame_start, atom_frame, N,
                           pl.ch, pl.rcap, part, F);
        if ((rc = check_launch("nep_fv_reduce_kernel"))) return rc;
    }
    hipLaunchKernelGGL(nep_virial_reduce_kernel<T>, dim3(9, (unsigned)nframes), dim3(256), 0, st, rows, frame_start,
                       virial);
    return check_launch("nep_virial_reduce_kernel");
}

MATPL_EXPORT int matpl_nep_force_virial_ri(int f64, const int64_t *NL, const void *dE, const void *Ri,
                                           const int64_t *frame_start, const int32_t *atom_frame, int64_t nframes,
                                           int64_t N, int S, int64_t max_frame, void *F, void *virial, void *work,
                                           int64_t work_elems, void *stream) {
    MATPL_CHECK_ARG(N >= 0 && S >= 0 && nframes >= 1 && max_frame >= 1, "nep_force_virial_ri: bad sizes");
    MATPL_CHECK_ARG(frame_start && F && virial, "nep_force_virial_ri: null pointer");
    MATPL_CHECK_ARG(N * S == 0 || (NL && dE && Ri && atom_frame && work), "nep_force_virial_ri: null input");
    MATPL_CHECK_ARG(aligned16(dE) && aligned16(Ri), "nep_force_virial_ri: dE and Ri must be 16-byte aligned");
    MATPL_CHECK_ARG(work_elems >= matpl_nep_force_virial_ri_work(f64, N, max_frame),
                    "nep_force_virial_ri: work buffer too small");
    hipStream_t st = (hipStream_t)stream;
    if (f64)
        return fv_launch<double>(NL, (const double *)dE, (const double *)Ri, frame_start, atom_frame, nframes, N, S,
                                 max_frame, (double *)F, (double *)virial, (double *)work, st);
    return fv_launch<float>(NL, (const float *)dE, (const float *)Ri, frame_start, atom_frame, nframes, N, S,
                            max_frame, (float *)F, (float *)virial, (float *)work, st);
}

MATPL_EXPORT int matpl_nep_force_virial_ri_grad(int f64, const int64_t *NL, const void *Ri, const void *gF,
                                                const void *gV, const int32_t *atom_frame, int64_t N, int S, void *g,
                                                void *stream) {
    if (N * S == 0) return OK;
    MATPL_CHECK_ARG(NL && Ri && g && (gF || gV) && (!gV || atom_frame), "nep_force_virial_ri_grad: null pointer");
    MATPL_CHECK_ARG(aligned16(Ri) && aligned16(g), "nep_force_virial_ri_grad: Ri and g must be 16-byte aligned");
    hipStream_t st = (hipStream_t)stream;
    const unsigned blocks = (unsigned)(ceil_div(N, 4) < 8192 ? ceil_div(N, 4) : 8192);
    if (f64)
        hipLaunchKernelGGL(nep_fv_grad_kernel<double>, dim3(blocks), dim3(256), 0, st, NL, (const double *)Ri,
                           (const double *)gF, (const double *)gV, atom_frame, N, S, (double *)g);
    else
        hipLaunchKernelGGL(nep_fv_grad_kernel<float>, dim3(blocks), dim3(256), 0, st, NL, (const float *)Ri,
                           (const float *)gF, (const float *)gV, atom_frame, N, S, (float *)g);
    return check_launch("nep_fv_grad_kernel");
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
