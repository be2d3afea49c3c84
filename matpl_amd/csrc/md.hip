// Velocity-Verlet integration and coordinate conversion for MD driven by
// the fused DP path. Positions/velocities are FP64 [N][3]; forces FP32.
//   kick      : v += (dt/2) F / m
//   kick_drift: v += (dt/2) F / m; x += dt v
//   cart->frac: f = x . Linv (row vectors), for neighbor-list rebuilds
//   frac->cart: x = f . L
// Units: Angstrom, fs, eV, amu (accel = F / (m * 103.642696) A/fs^2).
#include "common.h"

namespace matpl {
namespace {

constexpr double MVV2E = 103.642696;   // amu A^2 / fs^2 in eV

__global__ void kick_kernel(double *__restrict__ v, double *__restrict__ x, const float *__restrict__ F,
                            const double *__restrict__ inv_mass, double dt, int64_t n, int drift) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n * 3;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = q / 3;
        double vv = v[q] + 0.5 * dt * (double)F[q] * inv_mass[i] / MVV2E;
        v[q] = vv;
        if (drift) x[q] += dt * vv;
    }
}

struct Mat3 { double a[9]; };

__global__ void affine_kernel(const double *__restrict__ in, double *__restrict__ out, Mat3 M, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double x = in[3 * i], y = in[3 * i + 1], z = in[3 * i + 2];
        out[3 * i] = x * M.a[0] + y * M.a[3] + z * M.a[6];
        out[3 * i + 1] = x * M.a[1] + y * M.a[4] + z * M.a[7];
        out[3 * i + 2] = x * M.a[2] + y * M.a[5] + z * M.a[8];
    }
}

}  // namespace
}  // namespace matpl

using namespace matpl;

MATPL_EXPORT int matpl_md_kick(double *v, double *x, const float *F, const double *inv_mass, double dt, int64_t n,
                               int drift, void *stream) {
    if (n <= 0) return OK;
    MATPL_CHECK_ARG(v && F && inv_mass && (!drift || x), "md_kick: null pointer");
    hipLaunchKernelGGL(kick_kernel, dim3(grid_cap(n * 3)), dim3(256), 0, (hipStream_t)stream, v, x, F, inv_mass, dt,
                       n, drift);
    return check_launch("kick_kernel");
}

// out[i] = in[i] . M (row vector times 3x3 matrix M, host array [9])
MATPL_EXPORT int matpl_md_affine(const double *in, double *out, const double *M_host, int64_t n, void *stream) {
    if (n <= 0) return OK;
    MATPL_CHECK_ARG(in && out && M_host, "md_affine: null pointer");
    Mat3 M;
    for (int k = 0; k < 9; k++) M.a[k] = M_host[k];
    hipLaunchKernelGGL(affine_kernel, dim3(grid_cap(n)), dim3(256), 0, (hipStream_t)stream, in, out, M, n);
    return check_launch("affine_kernel");
}
