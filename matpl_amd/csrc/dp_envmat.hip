// DP se_e2_a environment matrix and its statistics.
//
// env_mat: ImageDR [B][N][S][4] = (r, dx, dy, dz) -> Ri [B][N][S][4] and
//   Ri_d [B][N][S][4][3], the tensors DP.calculate_Ri returns
//   (src/model/dp_dp.py:217-302):
//     s(r) = 1/r (0 < r < rmin); (1/r)(u^3(-6u^2+15u-10)+1), u=(r-rmin)/(rmax-rmin)
//            (rmin < r < rmax); 0 otherwise
//     Ri   = [s, s x/r, s y/r, s z/r]   (xyz terms only where |r| > 1e-5)
//     dfeat = d Ri / d(dx,dy,dz), analytic (dp_dp.py:266-288)
//     Ri <- (Ri - davg[t_i]) / dstd[t_i];  Ri_d <- -dfeat / dstd[t_i]
//   One lane per slot; davg/dstd are in the reference layout [T][S*4].
//
// calc_stat: davg/dstd [T][4] from the un-normalized Ri of every slot of every
//   atom, empty slots included (src/pre_data/dpuni_data_loader.py:422-560):
//   davg = [mean s, 0, 0, 0]; dstd = [std s, std_xyz x3] with the xyz moments
//   averaged over the three components and a 1e-2 floor. Per-block partial
//   sums in FP64 (LDS atomics inside a block), then a per-type final pass.
#include "common.h"

namespace matpl {
namespace {

template <typename T>
__device__ __forceinline__ void smooth_s(T r, T rmin, T rmax, T &s, T &ds) {
    if (r > T(0) && r < rmin) {
        s = T(1) / r;
        ds = -T(1) / (r * r);
    } else if (r > rmin && r < rmax) {
        const T inv = T(1) / (rmax - rmin);
        const T u = (r - rmin) * inv;
        const T vv = u * u * u * (T(-6) * u * u + T(15) * u - T(10)) + T(1);
        const T dvv = T(-30) * u * u * (u - T(1)) * (u - T(1)) * inv;
        s = vv / r;
        ds = dvv / r - vv / (r * r);
    } else {
        s = T(0);
        ds = T(0);
    }
}

template <typename T>
__global__ void env_mat_kernel(const T *__restrict__ idr, const int32_t *__restrict__ type_map,
                               const T *__restrict__ davg, const T *__restrict__ dstd, int64_t N, int64_t S,
                               int ntypes, int64_t nslots, T rmin, T rmax, T *__restrict__ Ri,
                               T *__restrict__ Ri_d) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nslots;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = q % S;
        const int64_t i = (q / S) % N;
        int ti = type_map[i];
        if (ti < 0) ti += ntypes;   // python-style negative index, as davg[Imagetype_map]
        const T *d = idr + 4 * q;
        const T r = d[0];
        T s, ds;
        smooth_s(r, rmin, rmax, s, ds);
        T ri[4] = {s, T(0), T(0), T(0)};
        T df[12];
#pragma unroll
        for (int x = 0; x < 12; x++) df[x] = T(0);
        if (fabs(r) > T(1e-5)) {
            const T x3[3] = {d[1], d[2], d[3]};
            ri[1] = s * x3[0] / r;
            ri[2] = s * x3[1] / r;
            ri[3] = s * x3[2] / r;
            const T common = (ds - s / r) / (r * r);
#pragma unroll
            for (int a = 0; a < 3; a++) {
                df[a] = ds * x3[a] / r;
#pragma unroll
                for (int b = 0; b < 3; b++) df[3 * (b + 1) + a] = common * x3[b] * x3[a] + (a == b ? s / r : T(0));
            }
        }
        const T *av = davg + ((int64_t)ti * S + k) * 4;
        const T *sd = dstd + ((int64_t)ti * S + k) * 4;
        T *ro = Ri + 4 * q;
        T *rdo = Ri_d + 12 * q;
#pragma unroll
        for (int w = 0; w < 4; w++) {
            const T inv = T(1) / sd[w];
            ro[w] = (ri[w] - av[w]) * inv;
#pragma unroll
            for (int a = 0; a < 3; a++) rdo[3 * w + a] = -df[3 * w + a] * inv;
        }
    }
}

constexpr int STAT_THREADS = 256;
constexpr int STAT_MAXT = 16;

// partial[blk][t][9] = {sum s, sum x, sum y, sum z, sum s^2, sum x^2, sum y^2, sum z^2, nslots}
__global__ void __launch_bounds__(STAT_THREADS)
stat_partial_kernel(const int32_t *__restrict__ list, const double *__restrict__ dR,
                    const int32_t *__restrict__ type_map, int64_t N, int64_t S, int ntypes, double rmin,
                    double rmax, double *__restrict__ partial) {
    __shared__ double acc[STAT_MAXT][9];
    for (int x = threadIdx.x; x < ntypes * 9; x += blockDim.x) acc[x / 9][x % 9] = 0.0;
    __syncthreads();
    const int64_t nslots = N * S;
    double loc[8];
    int cur_t = -1;
#pragma unroll
    for (int x = 0; x < 8; x++) loc[x] = 0.0;
    double locn = 0.0;
    auto flush = [&](int t) {
        if (t < 0) return;
#pragma unroll
        for (int x = 0; x < 8; x++) atomicAdd(&acc[t][x], loc[x]);
        atomicAdd(&acc[t][8], locn);
#pragma unroll
        for (int x = 0; x < 8; x++) loc[x] = 0.0;
        locn = 0.0;
    };
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nslots;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int t = type_map[q / S];
        if (t < 0 || t >= ntypes) continue;
        if (t != cur_t) { flush(cur_t); cur_t = t; }
        locn += 1.0;
        if (list[q] <= 0) continue;
        const double *d = dR + 3 * q;
        const double dr2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        const double r = sqrt(dr2);
        const double x = dr2 / r;
        double vv;
        if (x < rmin) vv = 1.0;
        else if (x < rmax) {
            const double u = (x - rmin) / (rmax - rmin);
            vv = u * u * u * (-6.0 * u * u + 15.0 * u - 10.0) + 1.0;
        } else vv = 0.0;
        const double ri[4] = {vv / x, vv * d[0] / dr2, vv * d[1] / dr2, vv * d[2] / dr2};
#pragma unroll
        for (int w = 0; w < 4; w++) { loc[w] += ri[w]; loc[4 + w] += ri[w] * ri[w]; }
    }
    flush(cur_t);
    __syncthreads();
    for (int x = threadIdx.x; x < ntypes * 9; x += blockDim.x)
        partial[(int64_t)blockIdx.x * ntypes * 9 + x] = acc[x / 9][x % 9];
}

__device__ double compute_std(double sum2, double sum, double n) {
    if (n == 0.0) return 1e-2;
    double val = sqrt(sum2 / n - (sum / n) * (sum / n));
    if (fabs(val) < 1e-2) val = 1e-2;
    return val;
}

__global__ void stat_final_kernel(const double *__restrict__ partial, int nblk, int ntypes,
                                  double *__restrict__ davg, double *__restrict__ dstd) {
    const int t = threadIdx.x;
    if (t >= ntypes) return;
    double a[9];
    for (int x = 0; x < 9; x++) a[x] = 0.0;
    for (int b = 0; b < nblk; b++)
        for (int x = 0; x < 9; x++) a[x] += partial[((int64_t)b * ntypes + t) * 9 + x];
    const double n = a[8];
    const double sa = (a[1] + a[2] + a[3]) / 3.0, sa2 = (a[5] + a[6] + a[7]) / 3.0;
    davg[4 * t] = a[0] / (n + 1e-15);
    davg[4 * t + 1] = davg[4 * t + 2] = davg[4 * t + 3] = 0.0;
    dstd[4 * t] = compute_std(a[4], a[0], n);
    dstd[4 * t + 1] = dstd[4 * t + 2] = dstd[4 * t + 3] = compute_std(sa2, sa, n);
}

template <typename T>
int env_mat_impl(const T *idr, const int32_t *type_map, const T *davg, const T *dstd, int64_t B, int64_t N,
                 int64_t S, int ntypes, double rmin, double rmax, T *Ri, T *Ri_d, hipStream_t st) {
    MATPL_CHECK_ARG(B >= 0 && N >= 0 && S >= 0 && ntypes >= 1, "dp_env_mat: bad sizes");
    MATPL_CHECK_ARG(rmax > rmin && rmin > 0, "dp_env_mat: need 0 < rmin < rmax");
    if (B * N * S == 0) return OK;
    MATPL_CHECK_ARG(idr && type_map && davg && dstd && Ri && Ri_d, "dp_env_mat: null pointer");
    hipLaunchKernelGGL(env_mat_kernel<T>, dim3(grid_cap(B * N * S)), dim3(256), 0, st, idr, type_map, davg, dstd,
                       N, S, ntypes, B * N * S, (T)rmin, (T)rmax, Ri, Ri_d);
    return check_launch("env_mat_kernel");
}

}  // namespace
}  // namespace matpl

using namespace matpl;

MATPL_EXPORT int matpl_dp_env_mat_f32(const float *image_dr, const int32_t *type_map, const float *davg,
                                      const float *dstd, int64_t B, int64_t N, int64_t S, int ntypes, double rmin,
                                      double rmax, float *Ri, float *Ri_d, void *stream) {
    return env_mat_impl<float>(image_dr, type_map, davg, dstd, B, N, S, ntypes, rmin, rmax, Ri, Ri_d,
                               (hipStream_t)stream);
}

MATPL_EXPORT int matpl_dp_env_mat_f64(const double *image_dr, const int32_t *type_map, const double *davg,
                                      const double *dstd, int64_t B, int64_t N, int64_t S, int ntypes,
                                      double rmin, double rmax, double *Ri, double *Ri_d, void *stream) {
    return env_mat_impl<double>(image_dr, type_map, davg, dstd, B, N, S, ntypes, rmin, rmax, Ri, Ri_d,
                                (hipStream_t)stream);
}

MATPL_EXPORT int64_t matpl_dp_calc_stat_workspace_elems(int ntypes) {
    return (int64_t)1024 * ntypes * 9;
}

MATPL_EXPORT int matpl_dp_calc_stat(const int32_t *list, const double *dR, const int32_t *type_map, int64_t N,
                                    int64_t S, int ntypes, double rmin, double rmax, double *davg, double *dstd,
                                    double *workspace, int64_t workspace_elems, void *stream) {
    hipStream_t st = (hipStream_t)stream;
    MATPL_CHECK_ARG(ntypes >= 1 && ntypes <= STAT_MAXT, "dp_calc_stat: ntypes must be in [1,%d]", STAT_MAXT);
    MATPL_CHECK_ARG(N >= 0 && S >= 0, "dp_calc_stat: bad sizes");
    MATPL_CHECK_ARG(davg && dstd && workspace, "dp_calc_stat: null pointer");
    int64_t nblk = (N * S + STAT_THREADS - 1) / STAT_THREADS;
    if (nblk < 1) nblk = 1;
    if (nblk > 1024) nblk = 1024;
    MATPL_CHECK_ARG(workspace_elems >= nblk * ntypes * 9, "dp_calc_stat: workspace too small");
    hipLaunchKernelGGL(stat_partial_kernel, dim3((unsigned)nblk), dim3(STAT_THREADS), 0, st, list, dR, type_map, N,
                       S, ntypes, rmin, rmax, workspace);
    int rc = check_launch("stat_partial_kernel");
    if (rc) return rc;
    hipLaunchKernelGGL(stat_final_kernel, dim3(1), dim3(64), 0, st, workspace, (int)nblk, ntypes, davg, dstd);
    return check_launch("stat_final_kernel");
}
