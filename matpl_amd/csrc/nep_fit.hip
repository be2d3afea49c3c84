// NEP fitting net on gfx950: per-type one-hidden-layer nets
//   x = q_scaler * q,  z = x W0[t] + b0[t],  h = tanh z,  Ei = h . w1[t] + b1[t]
// (src/model/nep_fitting.py:40-163 with network_size [H, 1], dispatched per atom
// type at src/model/nep_net.py:470-501), evaluated for every type in one launch.
//
// The forward also writes g = dEi/dq (the reference obtains it by autograd through
// the net); the backward takes the upstream gradients of Ei and of g and returns
// dL/dq and the parameter gradients, i.e. it is the net's first and second
// derivative in one kernel (the force loss reaches the weights through g):
//   u = w1 (1 - h^2),  g_x = W0 u,  gamma = q_scaler * gg,  v = W0^T gamma,
//   t = gE u - 2 v w1 h (1 - h^2)
//   dL/dq  = q_scaler * (W0 t)
//   dL/dW0 = sum_atoms x t^T + gamma u^T,   dL/db0 = sum t,
//   dL/dw1 = sum gE h + v (1 - h^2),        dL/db1 = sum gE.
//
// Layout: atoms grouped by type (perm) and cut into tiles of at most TA atoms of
// one type; one 256-thread block per tile. The tile's rows of x (and gamma) are
// staged in LDS with coalesced loads; thread (a, group) then runs a strided
// subset of hidden units (or features) for atom a, with the weights read as the
// same address across the atom lanes. Parameter gradients are summed per tile
// over its atoms in LDS (threads over weight entries) and written as one partial
// row per tile; a second kernel adds each type's rows in a fixed order, so results
// do not depend on scheduling.
#include "common.h"

namespace matpl {
namespace {

constexpr int FIT_THREADS = 256;

struct FitArgs {
    const void *q;           // [N][D] unscaled features
    const void *qs;          // [D]
    const int32_t *perm;     // [N] atoms grouped by type
    const int32_t *tiles;    // [ntiles][3]: type, first position in perm, count (<= TA)
    const void *W0;          // [T][D][H]
    const void *b0;          // [T][H]
    const void *w1;          // [T][H]
    const void *b1;          // [T]
    void *Ei;                // [N]              forward
    void *g;                 // [N][D] dEi/dq    forward (may be null)
    const void *gE;          // [N]              backward (may be null)
    const void *gg;          // [N][D]           backward (may be null)
    void *dq;                // [N][D]           backward (may be null)
    void *part;              // [ntiles][PW]     backward: per-tile parameter-gradient rows
    int D, H, TA;
};

__host__ __device__ inline int odd(int n) { return n | 1; }     // LDS row stride (odd: fewer bank conflicts)

template <typename T>
__device__ __forceinline__ T tanh_t(T z);
template <>
__device__ __forceinline__ float tanh_t<float>(float z) { return tanhf(z); }
template <>
__device__ __forceinline__ double tanh_t<double>(double z) { return tanh(z); }

template <typename T>
__global__ void __launch_bounds__(FIT_THREADS) nep_fit_fwd_kernel(FitArgs a) {
    extern __shared__ unsigned char smem_raw[];
    const int D = a.D, H = a.H, TA = a.TA;
    const int sd = odd(D), sh = odd(H);
    const int NG = FIT_THREADS / TA;                      // thread groups per atom
    T *xs = reinterpret_cast<T *>(smem_raw);              // [TA][sd]  x, then g
    T *us = xs + TA * sd;                                 // [TA][sh]  u
    T *ep = us + TA * sh;                                 // [NG][TA]  energy partials
    const int *tile = a.tiles + 3 * blockIdx.x;
    const int ty = tile[0], first = tile[1], cnt = tile[2];
    const T *q = reinterpret_cast<const T *>(a.q);
    const T *qs = reinterpret_cast<const T *>(a.qs);
    const T *W0 = reinterpret_cast<const T *>(a.W0) + (size_t)ty * D * H;
    const T *b0 = reinterpret_cast<const T *>(a.b0) + (size_t)ty * H;
    const T *w1 = reinterpret_cast<const T *>(a.w1) + (size_t)ty * H;
    for (int e = threadIdx.x; e < cnt * D; e += FIT_THREADS) {
        const int at = e / D, d = e - at * D;
        xs[at * sd + d] = q[(size_t)a.perm[first + at] * D + d] * qs[d];
    }
    __syncthreads();
    const int at = threadIdx.x % TA, grp = threadIdx.x / TA;
    T epart = T(0);
    if (at < cnt) {
        const T *x = xs + at * sd;
        for (int j = grp; j < H; j += NG) {
            T z = b0[j];
            for (int d = 0; d < D; d++) z += x[d] * W0[d * H + j];
            const T h = tanh_t(z);
            epart += w1[j] * h;
            us[at * sh + j] = w1[j] * (T(1) - h * h);
        }
    }
    ep[grp * TA + at] = epart;
    __syncthreads();
    // g_x[d] = sum_j W0[d][j] u[j], written over x (no longer needed)
    if (at < cnt) {
        const T *u = us + at * sh;
        for (int d = grp; d < D; d += NG) {
            T s = T(0);
            for (int j = 0; j < H; j++) s += W0[d * H + j] * u[j];
            xs[at * sd + d] = s * qs[d];
        }
    }
    __syncthreads();
    if (threadIdx.x < cnt) {
        T e = reinterpret_cast<const T *>(a.b1)[ty];
        for (int k = 0; k < NG; k++) e += ep[k * TA + threadIdx.x];
        reinterpret_cast<T *>(a.Ei)[a.perm[first + threadIdx.x]] = e;
    }
    if (a.g) {
        T *g = reinterpret_cast<T *>(a.g);
        for (int e = threadIdx.x; e < cnt * D; e += FIT_THREADS) {
            const int r = e / D, d = e - r * D;
            g[(size_t)a.perm[first + r] * D + d] = xs[r * sd + d];
        }
    }
}

// One partial row per tile: [D*H] dW0, [H] db0, [H] dw1, [1] db1.
__host__ __device__ inline int part_width(int D, int H) { return D * H + 2 * H + 1; }

template <typename T>
__global__ void __launch_bounds__(FIT_THREADS) nep_fit_bwd_kernel(FitArgs a) {
    extern __shared__ unsigned char smem_raw[];
    const int D = a.D, H = a.H, TA = a.TA;
    const int sd = odd(D), sh = odd(H);
    const int NG = FIT_THREADS / TA;
    T *xs = reinterpret_cast<T *>(smem_raw);              // [TA][sd] x
    T *gs = xs + TA * sd;                                 // [TA][sd] gamma
    T *ts = gs + TA * sd;                                 // [TA][sh] t
    T *us = ts + TA * sh;                                 // [TA][sh] u
    T *rs = us + TA * sh;                                 // [TA][sh] gE h + v (1 - h^2)
    T *es = rs + TA * sh;                                 // [TA]     gE
    const int *tile = a.tiles + 3 * blockIdx.x;
    const int ty = tile[0], first = tile[1], cnt = tile[2];
    const T *q = reinterpret_cast<const T *>(a.q);
    const T *qs = reinterpret_cast<const T *>(a.qs);
    const T *gg = reinterpret_cast<const T *>(a.gg);
    const T *gE = reinterpret_cast<const T *>(a.gE);
    const T *W0 = reinterpret_cast<const T *>(a.W0) + (size_t)ty * D * H;
    const T *b0 = reinterpret_cast<const T *>(a.b0) + (size_t)ty * H;
    const T *w1 = reinterpret_cast<const T *>(a.w1) + (size_t)ty * H;
    for (int e = threadIdx.x; e < cnt * D; e += FIT_THREADS) {
        const int at = e / D, d = e - at * D;
        const size_t src = (size_t)a.perm[first + at] * D + d;
        xs[at * sd + d] = q[src] * qs[d];
        gs[at * sd + d] = gg ? gg[src] * qs[d] : T(0);
    }
    for (int r = threadIdx.x; r < TA; r += FIT_THREADS) es[r] = (gE && r < cnt) ? gE[a.perm[first + r]] : T(0);
    __syncthreads();
    const int at = threadIdx.x % TA, grp = threadIdx.x / TA;
    if (at < cnt) {
        const T *x = xs + at * sd, *gm = gs + at * sd;
        const T ge = es[at];
        for (int j = grp; j < H; j += NG) {
            T z = b0[j], v = T(0);
            for (int d = 0; d < D; d++) {
                const T w = W0[d * H + j];
                z += x[d] * w;
                v += gm[d] * w;
            }
            const T h = tanh_t(z), s = T(1) - h * h;
            const T u = w1[j] * s;
            us[at * sh + j] = u;
            ts[at * sh + j] = ge * u - T(2) * v * w1[j] * h * s;
            rs[at * sh + j] = ge * h + v * s;
        }
    }
    __syncthreads();
    if (a.dq && at < cnt) {
        T *dq = reinterpret_cast<T *>(a.dq);
        const T *t = ts + at * sh;
        for (int d = grp; d < D; d += NG) {
            T s = T(0);
            for (int j = 0; j < H; j++) s += W0[d * H + j] * t[j];
            dq[(size_t)a.perm[first + at] * D + d] = s * qs[d];
        }
    }
    // this tile's parameter-gradient row
    const int PW = part_width(D, H);
    T *row = reinterpret_cast<T *>(a.part) + (size_t)blockIdx.x * PW;
    for (int e = threadIdx.x; e < PW; e += FIT_THREADS) {
        T s = T(0);
        if (e < D * H) {
            const int d = e / H, j = e - d * H;
            for (int r = 0; r < cnt; r++) s += xs[r * sd + d] * ts[r * sh + j] + gs[r * sd + d] * us[r * sh + j];
        } else if (e < D * H + H) {
            const int j = e - D * H;
            for (int r = 0; r < cnt; r++) s += ts[r * sh + j];
        } else if (e < D * H + 2 * H) {
            const int j = e - D * H - H;
            for (int r = 0; r < cnt; r++) s += rs[r * sh + j];
        } else {
            for (int r = 0; r < cnt; r++) s += es[r];
        }
        row[e] = s;
    }
}

// Parameter gradients of type t = sum of its tiles' rows [tstart[t], tstart[t+1]): a block takes 32
// entries, its 8 thread groups sum every 8th tile each, and the 8 partial sums are added in group
// order (deterministic).
constexpr int RED_ENTRIES = 32, RED_GROUPS = 8;
template <typename T>
__global__ void __launch_bounds__(RED_ENTRIES * RED_GROUPS)
nep_fit_reduce_kernel(const T *part, const int32_t *tstart, int PW, int D, int H, T *dW0, T *db0, T *dw1, T *db1) {
    __shared__ T red[RED_GROUPS][RED_ENTRIES];
    const int t = blockIdx.y;
    const int el = threadIdx.x % RED_ENTRIES, g = threadIdx.x / RED_ENTRIES;
    const int e = blockIdx.x * RED_ENTRIES + el;
    T s = T(0);
    if (e < PW)
        for (int k = tstart[t] + g; k < tstart[t + 1]; k += RED_GROUPS) s += part[(size_t)k * PW + e];
    red[g][el] = s;
    __syncthreads();
    if (g != 0 || e >= PW) return;
    s = T(0);
    for (int q = 0; q < RED_GROUPS; q++) s += red[q][el];
    if (e < D * H) dW0[(size_t)t * D * H + e] = s;
    else if (e < D * H + H) db0[(size_t)t * H + e - D * H] = s;
    else if (e < D * H + 2 * H) dw1[(size_t)t * H + e - D * H - H] = s;
    else db1[t] = s;
}

size_t fwd_lds(int D, int H, int TA, size_t es) {
    return es * ((size_t)TA * odd(D) + (size_t)TA * odd(H) + (size_t)FIT_THREADS);
}
size_t bwd_lds(int D, int H, int TA, size_t es) {
    return es * (2 * (size_t)TA * odd(D) + 3 * (size_t)TA * odd(H) + (size_t)TA);
}

}  // namespace
}  // namespace matpl

using namespace matpl;

// Atoms per tile for a net of D features and H hidden units: the largest power of two in
// [8, 64] whose backward LDS fits 64 KiB (the forward needs less).
MATPL_EXPORT int matpl_nep_fit_tile_atoms(int f64, int D, int H) {
    if (D < 1 || H < 1) return -1;
    const size_t es = f64 ? 8 : 4;
    for (int ta = 64; ta >= 8; ta >>= 1)
        if (bwd_lds(D, H, ta, es) <= 64 * 1024) return ta;
    return -1;
}

MATPL_EXPORT int matpl_nep_fit_part_width(int D, int H) { return part_width(D, H); }

MATPL_EXPORT int matpl_nep_fit_forward(int f64, const void *q, const void *qs, const int32_t *perm,
                                       const int32_t *tiles, int ntiles, int tile_atoms, const void *W0,
                                       const void *b0, const void *w1, const void *b1, int D, int H, void *Ei,
                                       void *g, void *stream) {
    MATPL_CHECK_ARG(D >= 1 && H >= 1 && ntiles >= 0, "nep_fit_forward: bad sizes D=%d H=%d", D, H);
    MATPL_CHECK_ARG(tile_atoms == matpl_nep_fit_tile_atoms(f64, D, H),
                    "nep_fit_forward: tile_atoms %d is not matpl_nep_fit_tile_atoms (%d)", tile_atoms,
                    matpl_nep_fit_tile_atoms(f64, D, H));
    if (ntiles == 0) return OK;
    MATPL_CHECK_ARG(q && qs && perm && tiles && W0 && b0 && w1 && b1 && Ei, "nep_fit_forward: null pointer");
    FitArgs a{};
    a.q = q; a.qs = qs; a.perm = perm; a.tiles = tiles; a.W0 = W0; a.b0 = b0; a.w1 = w1; a.b1 = b1;
    a.Ei = Ei; a.g = g; a.D = D; a.H = H; a.TA = tile_atoms;
    hipStream_t st = (hipStream_t)stream;
    if (f64)
        hipLaunchKernelGGL(nep_fit_fwd_kernel<double>, dim3(ntiles), dim3(FIT_THREADS), fwd_lds(D, H, a.TA, 8), st, a);
    else
        hipLaunchKernelGGL(nep_fit_fwd_kernel<float>, dim3(ntiles), dim3(FIT_THREADS), fwd_lds(D, H, a.TA, 4), st, a);
    return check_launch("nep_fit_fwd_kernel");
}

MATPL_EXPORT int matpl_nep_fit_backward(int f64, const void *q, const void *qs, const int32_t *perm,
                                        const int32_t *tiles, int ntiles, int tile_atoms, const int32_t *tstart,
                                        int ntypes, const void *W0, const void *b0, const void *w1, const void *b1,
                                        int D, int H, const void *gE, const void *gg, void *dq, void *part,
                                        void *dW0, void *db0, void *dw1, void *db1, void *stream) {
    MATPL_CHECK_ARG(D >= 1 && H >= 1 && ntiles >= 0 && ntypes >= 1, "nep_fit_backward: bad sizes D=%d H=%d", D, H);
    MATPL_CHECK_ARG(tile_atoms == matpl_nep_fit_tile_atoms(f64, D, H),
                    "nep_fit_backward: tile_atoms %d is not matpl_nep_fit_tile_atoms (%d)", tile_atoms,
                    matpl_nep_fit_tile_atoms(f64, D, H));
    MATPL_CHECK_ARG(tstart && dW0 && db0 && dw1 && db1, "nep_fit_backward: null pointer");
    hipStream_t st = (hipStream_t)stream;
    const int PW = part_width(D, H);
    if (ntiles > 0) {
        MATPL_CHECK_ARG(q && qs && perm && tiles && W0 && b0 && w1 && b1 && part, "nep_fit_backward: null pointer");
        FitArgs a{};
        a.q = q; a.qs = qs; a.perm = perm; a.tiles = tiles; a.W0 = W0; a.b0 = b0; a.w1 = w1; a.b1 = b1;
        a.gE = gE; a.gg = gg; a.dq = dq; a.part = part; a.D = D; a.H = H; a.TA = tile_atoms;
        if (f64)
            hipLaunchKernelGGL(nep_fit_bwd_kernel<double>, dim3(ntiles), dim3(FIT_THREADS), bwd_lds(D, H, a.TA, 8), st,
                               a);
        else
            hipLaunchKernelGGL(nep_fit_bwd_kernel<float>, dim3(ntiles), dim3(FIT_THREADS), bwd_lds(D, H, a.TA, 4), st,
                               a);
        int rc = check_launch("nep_fit_bwd_kernel");
        if (rc) return rc;
    }
    const dim3 grid((PW + RED_ENTRIES - 1) / RED_ENTRIES, ntypes);
    const dim3 block(RED_ENTRIES * RED_GROUPS);
    if (f64)
        hipLaunchKernelGGL(nep_fit_reduce_kernel<double>, grid, block, 0, st, static_cast<const double *>(part),
                           tstart, PW, D, H, static_cast<double *>(dW0), static_cast<double *>(db0),
                           static_cast<double *>(dw1), static_cast<double *>(db1));
    else
        hipLaunchKernelGGL(nep_fit_reduce_kernel<float>, grid, block, 0, st, static_cast<const float *>(part),
                           tstart, PW, D, H, static_cast<float *>(dW0), static_cast<float *>(db0),
                           static_cast<float *>(dw1), static_cast<float *>(db1));
    return check_launch("nep_fit_reduce_kernel");
}
