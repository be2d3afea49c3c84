// DP force / virial scatter from dE = dE/dRi and their backward (the
// calculateForce / calculateVirial operator contract).
//
// Layouts (B frames, N local atoms, S slots per atom, NA = N + nghost rows):
//   nlist [B][N][S] int32, 0-based neighbor row, -1 = empty slot
//   dE    [B][N][S][4]      Ri_d [B][N][S][4][3]     Rij [B][N][S][3]
//   force [B][NA][3]        atom_virial [B][NA][9]   virial [B][9]
//
// One lane per slot. A lane reads its slot's 4 dE values and 12 Ri_d values
// as 16-byte vectors, so a wavefront streams 64 consecutive slots of each
// array; empty slots read only the 4-byte index. The neighbor-side sum is a
// hardware FP atomic (global_atomic_add_f32/f64) because any row may receive
// contributions from any slot. Everything is enqueued on the caller's stream.
#include "common.h"

#include <mutex>

namespace matpl {
namespace {

template <typename T> struct V4;
template <> struct V4<float> {
    static __device__ __forceinline__ void load(const float *p, float o[4]) {
        float4 v = *reinterpret_cast<const float4 *>(p);
        o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
    }
};
template <> struct V4<double> {
    static __device__ __forceinline__ void load(const double *p, double o[4]) {
        double2 a = reinterpret_cast<const double2 *>(p)[0];
        double2 b = reinterpret_cast<const double2 *>(p)[1];
        o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
    }
};

// f[a] = sum_w dE[w] * Ri_d[w][a] for one slot.
template <typename T>
__device__ __forceinline__ void slot_force(const T *__restrict__ dE, const T *__restrict__ Ri_d,
                                           int64_t s, T f[3]) {
    T e[4], d[12];
    V4<T>::load(dE + 4 * s, e);
    V4<T>::load(Ri_d + 12 * s, d);
    V4<T>::load(Ri_d + 12 * s + 4, d + 4);
    V4<T>::load(Ri_d + 12 * s + 8, d + 8);
#pragma unroll
    for (int a = 0; a < 3; a++) f[a] = e[0] * d[a] + e[1] * d[3 + a] + e[2] * d[6 + a] + e[3] * d[9 + a];
}

template <typename T>
__global__ void force_init_kernel(const T *__restrict__ Fc, T *__restrict__ force, int64_t N,
                                  int64_t NA, int64_t total) {
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = idx / (NA * 3), r = idx - b * NA * 3;
        force[idx] = (r < N * 3) ? Fc[b * N * 3 + r] : T(0);
    }
}

template <typename T>
__global__ void force_scatter_kernel(const int32_t *__restrict__ nlist, const T *__restrict__ dE,
                                     const T *__restrict__ Ri_d, T *__restrict__ force, int64_t NS,
                                     int64_t NA, int64_t nslots) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < nslots;
         s += (int64_t)gridDim.x * blockDim.x) {
        const int j = nlist[s];
        if (j < 0) continue;
        T f[3];
        slot_force(dE, Ri_d, s, f);
        T *out = force + ((s / NS) * NA + j) * 3;
        atomic_add(out + 0, f[0]);
        atomic_add(out + 1, f[1]);
        atomic_add(out + 2, f[2]);
    }
}

template <typename T>
__global__ void virial_scatter_kernel(const int32_t *__restrict__ nlist, const T *__restrict__ dE,
                                      const T *__restrict__ Rij, const T *__restrict__ Ri_d,
                                      T *__restrict__ atom_virial, int64_t NS, int64_t NA,
                                      int64_t nslots) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < nslots;
         s += (int64_t)gridDim.x * blockDim.x) {
        const int j = nlist[s];
        if (j < 0) continue;
        T f[3];
        slot_force(dE, Ri_d, s, f);
        const T r[3] = {Rij[3 * s], Rij[3 * s + 1], Rij[3 * s + 2]};
        T *out = atom_virial + ((s / NS) * NA + j) * 9;
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int c = 0; c < 3; c++) atomic_add(out + 3 * a + c, f[a] * r[c]);
    }
}

// Partial sums of atom_virial rows: block (chunk, b) -> partial[b][chunk][9].
constexpr int RED_THREADS = 256;
template <typename T>
__global__ void __launch_bounds__(RED_THREADS)
virial_partial_kernel(const T *__restrict__ atom_virial, T *__restrict__ partial, int64_t NA,
                      int64_t rows_per_chunk) {
    const int64_t b = blockIdx.y, chunk = blockIdx.x;
    const int64_t r0 = chunk * rows_per_chunk;
    const int64_t r1 = min(NA, r0 + rows_per_chunk);
    T acc[9];
#pragma unroll
    for (int v = 0; v < 9; v++) acc[v] = T(0);
    for (int64_t r = r0 + threadIdx.x; r < r1; r += RED_THREADS) {
        const T *row = atom_virial + (b * NA + r) * 9;
#pragma unroll
        for (int v = 0; v < 9; v++) acc[v] += row[v];
    }
    __shared__ T red[RED_THREADS / 64][9];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int v = 0; v < 9; v++) {
        T x = wave_sum(acc[v]);
        if (lane == 0) red[wid][v] = x;
    }
    __syncthreads();
    if (threadIdx.x < 9) {
        T x = T(0);
        for (int w = 0; w < RED_THREADS / 64; w++) x += red[w][threadIdx.x];
        partial[(b * gridDim.x + chunk) * 9 + threadIdx.x] = x;
    }
}

template <typename T>
__global__ void virial_final_kernel(const T *__restrict__ partial, T *__restrict__ virial, int nchunk,
                                    int64_t B) {
    const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (idx >= B * 9) return;
    const int64_t b = idx / 9, v = idx % 9;
    T x = T(0);
    for (int c = 0; c < nchunk; c++) x += partial[(b * nchunk + c) * 9 + v];
    virial[idx] = x;
}

// g[s][w] = -gF_i . Ri_d[s][w] + gF_{j mod N} . Ri_d[s][w]  (j >= 0)
template <typename T>
__global__ void force_grad_kernel(const int32_t *__restrict__ nlist, const T *__restrict__ Ri_d,
                                  const T *__restrict__ gF, T *__restrict__ g, int64_t N, int64_t S,
                                  int64_t NA, int64_t nslots) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < nslots;
         s += (int64_t)gridDim.x * blockDim.x) {
        const int64_t bi = s / S;              // b*N + i
        const int64_t b = bi / N, i = bi - b * N;
        int j = nlist[s];
        T d[12];
        V4<T>::load(Ri_d + 12 * s, d);
        V4<T>::load(Ri_d + 12 * s + 4, d + 4);
        V4<T>::load(Ri_d + 12 * s + 8, d + 8);
        const T *gi = gF + (b * NA + i) * 3;
        T q[3] = {-gi[0], -gi[1], -gi[2]};
        if (j >= 0) {
            if (j >= N) j = j % N;
            const T *gj = gF + (b * NA + j) * 3;
            q[0] += gj[0]; q[1] += gj[1]; q[2] += gj[2];
        }
        T o[4];
#pragma unroll
        for (int w = 0; w < 4; w++) o[w] = q[0] * d[3 * w] + q[1] * d[3 * w + 1] + q[2] * d[3 * w + 2];
#pragma unroll
        for (int w = 0; w < 4; w++) g[4 * s + w] = o[w];
    }
}

// g[s][w] = sum_{a,c} gV[3a+c] Rij[s][c] Ri_d[s][w][a]  (0 for empty slots)
template <typename T>
__global__ void virial_grad_kernel(const int32_t *__restrict__ nlist, const T *__restrict__ Rij,
                                   const T *__restrict__ Ri_d, const T *__restrict__ gV,
                                   T *__restrict__ g, int64_t NS, int64_t nslots) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < nslots;
         s += (int64_t)gridDim.x * blockDim.x) {
        T o[4] = {T(0), T(0), T(0), T(0)};
        if (nlist[s] >= 0) {
            const T *gv = gV + (s / NS) * 9;
            const T r[3] = {Rij[3 * s], Rij[3 * s + 1], Rij[3 * s + 2]};
            T q[3];   // q[a] = sum_c gV[3a+c] r[c]
#pragma unroll
            for (int a = 0; a < 3; a++) q[a] = gv[3 * a] * r[0] + gv[3 * a + 1] * r[1] + gv[3 * a + 2] * r[2];
            T d[12];
            V4<T>::load(Ri_d + 12 * s, d);
            V4<T>::load(Ri_d + 12 * s + 4, d + 4);
            V4<T>::load(Ri_d + 12 * s + 8, d + 8);
#pragma unroll
            for (int w = 0; w < 4; w++) o[w] = q[0] * d[3 * w] + q[1] * d[3 * w + 1] + q[2] * d[3 * w + 2];
        }
#pragma unroll
        for (int w = 0; w < 4; w++) g[4 * s + w] = o[w];
    }
}

constexpr int THREADS = 256;
inline unsigned grid_for(int64_t n) {
    // Enough waves to cover ~8 blocks per CU on 256 CUs; grid-stride beyond.
    const int64_t blocks = (n + THREADS - 1) / THREADS;
    return (unsigned)(blocks < 2048 ? (blocks > 0 ? blocks : 1) : 2048);
}

template <typename T>
int force_impl(const int32_t *nlist, const T *dE, const T *Ri_d, const T *Fc, int64_t B, int64_t N,
               int64_t S, int64_t nghost, T *force, hipStream_t st) {
    MATPL_CHECK_ARG(B >= 0 && N >= 0 && S >= 0 && nghost >= 0, "dp_force: negative size");
    const int64_t NA = N + nghost;
    if (B * NA == 0) return OK;
    MATPL_CHECK_ARG(force && (N == 0 || Fc), "dp_force: null pointer");
    hipLaunchKernelGGL(force_init_kernel<T>, dim3(grid_for(B * NA * 3)), dim3(THREADS), 0, st, Fc,
                       force, N, NA, B * NA * 3);
    int rc = check_launch("force_init_kernel");
    if (rc || B * N * S == 0) return rc;
    MATPL_CHECK_ARG(nlist && dE && Ri_d, "dp_force: null pointer");
    MATPL_CHECK_ARG(aligned16(dE) && aligned16(Ri_d), "dp_force: dE/Ri_d must be 16-byte aligned");
    hipLaunchKernelGGL(force_scatter_kernel<T>, dim3(grid_for(B * N * S)), dim3(THREADS), 0, st,
                       nlist, dE, Ri_d, force, N * S, NA, B * N * S);
    return check_launch("force_scatter_kernel");
}

template <typename T>
int virial_impl(const int32_t *nlist, const T *dE, const T *Rij, const T *Ri_d, int64_t B, int64_t N,
                int64_t S, int64_t nghost, T *virial, T *atom_virial, T *workspace,
                int64_t workspace_elems, hipStream_t st) {
    MATPL_CHECK_ARG(B >= 0 && N >= 0 && S >= 0 && nghost >= 0, "dp_virial: negative size");
    const int64_t NA = N + nghost;
    if (B == 0) return OK;
    MATPL_CHECK_ARG(virial && (NA == 0 || atom_virial), "dp_virial: null pointer");
    int rc = OK;
    if (NA > 0) {
        if ((rc = fill_async<T>(atom_virial, T(0), B * NA * 9, st))) return rc;
    }
    if (B * N * S > 0) {
        MATPL_CHECK_ARG(nlist && dE && Rij && Ri_d, "dp_virial: null pointer");
        MATPL_CHECK_ARG(aligned16(dE) && aligned16(Ri_d), "dp_virial: dE/Ri_d must be 16-byte aligned");
        hipLaunchKernelGGL(virial_scatter_kernel<T>, dim3(grid_for(B * N * S)), dim3(THREADS), 0, st,
                           nlist, dE, Rij, Ri_d, atom_virial, N * S, NA, B * N * S);
        if ((rc = check_launch("virial_scatter_kernel"))) return rc;
    }
    // Deterministic two-stage row reduction.
    int nchunk = (int)((NA + 4095) / 4096);
    if (nchunk < 1) nchunk = 1;
    if (nchunk > 256) nchunk = 256;
    MATPL_CHECK_ARG(workspace && workspace_elems >= B * nchunk * 9,
                    "dp_virial: workspace too small (%lld < %lld)", (long long)workspace_elems,
                    (long long)(B * nchunk * 9));
    const int64_t rows_per_chunk = (NA + nchunk - 1) / nchunk;
    if (NA > 0) {
        hipLaunchKernelGGL(virial_partial_kernel<T>, dim3(nchunk, (unsigned)B), dim3(RED_THREADS), 0, st,
                           atom_virial, workspace, NA, rows_per_chunk > 0 ? rows_per_chunk : 1);
        if ((rc = check_launch("virial_partial_kernel"))) return rc;
    } else {
        if ((rc = fill_async<T>(workspace, T(0), B * nchunk * 9, st))) return rc;
    }
    hipLaunchKernelGGL(virial_final_kernel<T>, dim3(ceil_div(B * 9, 64)), dim3(64), 0, st, workspace,
                       virial, nchunk, B);
    return check_launch("virial_final_kernel");
}

template <typename T>
int force_grad_impl(const int32_t *nlist, const T *Ri_d, const T *gF, int64_t B, int64_t N, int64_t S,
                    int64_t nghost, T *g, hipStream_t st) {
    MATPL_CHECK_ARG(B >= 0 && N >= 0 && S >= 0 && nghost >= 0, "dp_force_grad: negative size");
    if (B * N * S == 0) return OK;
    MATPL_CHECK_ARG(nlist && Ri_d && gF && g, "dp_force_grad: null pointer");
    MATPL_CHECK_ARG(aligned16(Ri_d), "dp_force_grad: Ri_d must be 16-byte aligned");
    hipLaunchKernelGGL(force_grad_kernel<T>, dim3(grid_for(B * N * S)), dim3(THREADS), 0, st, nlist,
                       Ri_d, gF, g, N, S, N + nghost, B * N * S);
    return check_launch("force_grad_kernel");
}

template <typename T>
int virial_grad_impl(const int32_t *nlist, const T *Rij, const T *Ri_d, const T *gV, int64_t B,
                     int64_t N, int64_t S, T *g, hipStream_t st) {
    MATPL_CHECK_ARG(B >= 0 && N >= 0 && S >= 0, "dp_virial_grad: negative size");
    if (B * N * S == 0) return OK;
    MATPL_CHECK_ARG(nlist && Rij && Ri_d && gV && g, "dp_virial_grad: null pointer");
    MATPL_CHECK_ARG(aligned16(Ri_d), "dp_virial_grad: Ri_d must be 16-byte aligned");
    hipLaunchKernelGGL(virial_grad_kernel<T>, dim3(grid_for(B * N * S)), dim3(THREADS), 0, st, nlist,
                       Rij, Ri_d, gV, g, N * S, B * N * S);
    return check_launch("virial_grad_kernel");
}

}  // namespace
}  // namespace matpl

// Scratch of matpl_dp_virial_noalloc: the partial sums of the virial's two-stage row reduction
// (at most 256 chunks x 9 per frame) for up to NOALLOC_MAX_FRAMES frames, 1.18 MB. Each GPU the
// library runs on holds its own copy until the process ends. Host code never passes the array's
// name to a kernel: it uses the device address hipGetSymbolAddress returns for the current device.
constexpr int64_t NOALLOC_MAX_FRAMES = 64;
constexpr int64_t NOALLOC_SCRATCH_ELEMS = NOALLOC_MAX_FRAMES * 256 * 9;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_ELEMS];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// The device address of `scratch` on the current device, looked up once per device.
int scratch_address(void **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        matpl::set_error("dp_virial_noalloc: hipGetDevice: %s", hipGetErrorString(e));
        return matpl::ERR_HIP;
    }
    MATPL_CHECK_ARG(dev >= 0 && dev < NOALLOC_MAX_DEVICES, "dp_virial_noalloc: device %d beyond %d", dev,
                    NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            matpl::set_error("dp_virial_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
            return matpl::ERR_HIP;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return matpl::OK;
}
}  // namespace

using matpl::force_grad_impl;
using matpl::force_impl;
using matpl::virial_grad_impl;
using matpl::virial_impl;

MATPL_EXPORT int matpl_dp_virial_workspace_elems(int64_t B, int64_t N, int64_t nghost) {
    (void)N; (void)nghost;
    return (int)(B * 256 * 9);
}

MATPL_EXPORT int matpl_dp_force_f32(const int32_t *nlist, const float *dE, const float *Ri_d,
                                    const float *F_center, int64_t B, int64_t N, int64_t S,
                                    int64_t nghost, float *force, void *stream) {
    return force_impl<float>(nlist, dE, Ri_d, F_center, B, N, S, nghost, force, (hipStream_t)stream);
}
MATPL_EXPORT int matpl_dp_force_f64(const int32_t *nlist, const double *dE, const double *Ri_d,
                                    const double *F_center, int64_t B, int64_t N, int64_t S,
                                    int64_t nghost, double *force, void *stream) {
    return force_impl<double>(nlist, dE, Ri_d, F_center, B, N, S, nghost, force, (hipStream_t)stream);
}
MATPL_EXPORT int matpl_dp_virial_f32(const int32_t *nlist, const float *dE, const float *Rij,
                                     const float *Ri_d, int64_t B, int64_t N, int64_t S, int64_t nghost,
                                     float *virial, float *atom_virial, float *workspace,
                                     int64_t workspace_elems, void *stream) {
    return virial_impl<float>(nlist, dE, Rij, Ri_d, B, N, S, nghost, virial, atom_virial, workspace,
                              workspace_elems, (hipStream_t)stream);
}
MATPL_EXPORT int matpl_dp_virial_f64(const int32_t *nlist, const double *dE, const double *Rij,
                                     const double *Ri_d, int64_t B, int64_t N, int64_t S, int64_t nghost,
                                     double *virial, double *atom_virial, double *workspace,
                                     int64_t workspace_elems, void *stream) {
    return virial_impl<double>(nlist, dE, Rij, Ri_d, B, N, S, nghost, virial, atom_virial, workspace,
                               workspace_elems, (hipStream_t)stream);
}
MATPL_EXPORT int matpl_dp_force_grad_f32(const int32_t *nlist, const float *Ri_d, const float *gF,
                                         int64_t B, int64_t N, int64_t S, int64_t nghost, float *g,
                                         void *stream) {
    return force_grad_impl<float>(nlist, Ri_d, gF, B, N, S, nghost, g, (hipStream_t)stream);
}
MATPL_EXPORT int matpl_dp_force_grad_f64(const int32_t *nlist, const double *Ri_d, const double *gF,
                                         int64_t B, int64_t N, int64_t S, int64_t nghost, double *g,
                                         void *stream) {
    return force_grad_impl<double>(nlist, Ri_d, gF, B, N, S, nghost, g, (hipStream_t)stream);
}
MATPL_EXPORT int matpl_dp_virial_grad_f32(const int32_t *nlist, const float *Rij, const float *Ri_d,
                                          const float *gV, int64_t B, int64_t N, int64_t S, float *g,
                                          void *stream) {
    return virial_grad_impl<float>(nlist, Rij, Ri_d, gV, B, N, S, g, (hipStream_t)stream);
}
MATPL_EXPORT int matpl_dp_virial_grad_f64(const int32_t *nlist, const double *Rij, const double *Ri_d,
                                          const double *gV, int64_t B, int64_t N, int64_t S, double *g,
                                          void *stream) {
    return virial_grad_impl<double>(nlist, Rij, Ri_d, gV, B, N, S, g, (hipStream_t)stream);
}

// calculateVirial without a workspace argument and without allocating: the reduction's partial sums
// go to this GPU's `scratch` (B <= NOALLOC_MAX_FRAMES). Same kernels, launch order and results as
// matpl_dp_virial_f32/_f64. Calls that overlap on one GPU share the scratch and race.
MATPL_EXPORT int matpl_dp_virial_noalloc(int f64, const int32_t *nlist, const void *dE, const void *Rij,
                                         const void *Ri_d, int64_t B, int64_t N, int64_t S, int64_t nghost,
                                         void *virial, void *atom_virial, void *stream) {
    MATPL_CHECK_ARG(B >= 0 && B <= NOALLOC_MAX_FRAMES, "dp_virial_noalloc: %lld frames, at most %lld",
                    (long long)B, (long long)NOALLOC_MAX_FRAMES);
    void *ws = nullptr;
    int rc = scratch_address(&ws);
    if (rc) return rc;
    if (f64)
        return virial_impl<double>(nlist, static_cast<const double *>(dE), static_cast<const double *>(Rij),
                                   static_cast<const double *>(Ri_d), B, N, S, nghost, static_cast<double *>(virial),
                                   static_cast<double *>(atom_virial), static_cast<double *>(ws),
                                   NOALLOC_SCRATCH_ELEMS, (hipStream_t)stream);
    return virial_impl<float>(nlist, static_cast<const float *>(dE), static_cast<const float *>(Rij),
                              static_cast<const float *>(Ri_d), B, N, S, nghost, static_cast<float *>(virial),
                              static_cast<float *>(atom_virial), static_cast<float *>(ws),
                              2 * NOALLOC_SCRATCH_ELEMS, (hipStream_t)stream);
}
