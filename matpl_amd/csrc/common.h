// Shared helpers for the gfx950 kernels and their C-ABI entry points.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

namespace matpl {

// Per-thread last error message returned by matpl_last_error().
void set_error(const char *fmt, ...);

// Status codes shared by every entry point (also listed in include/matpl_hip.h).
enum Status : int {
    OK = 0,
    ERR_ARGUMENT = -1,     // bad size / null pointer / unsupported shape
    ERR_HIP = -2,          // a HIP runtime call failed
    ERR_OVERFLOW = -3,     // a neighbor bucket overflowed its capacity
};

inline int check_hip(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        set_error("%s: %s", what, hipGetErrorString(e));
        return ERR_HIP;
    }
    return OK;
}

// Checks the launch just issued on the current thread.
inline int check_launch(const char *kernel) {
    return check_hip(hipGetLastError(), kernel);
}

inline unsigned ceil_div(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

// Grid for a grid-stride loop over n items with `threads` per block: one item
// per lane up to 8192 blocks (32 per CU on 256 CUs), strided beyond.
inline unsigned grid_cap(int64_t n, int threads = 256) {
    int64_t blocks = (n + threads - 1) / threads;
    if (blocks < 1) blocks = 1;
    if (blocks > 8192) blocks = 8192;
    return (unsigned)blocks;
}

// Hardware FP atomics on coarse-grained device memory (hipMalloc / torch).
__device__ __forceinline__ void atomic_add(float *p, float v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_add(double *p, double v) { unsafeAtomicAdd(p, v); }

// Wavefront (64-lane) sum via cross-lane shuffles.
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Zero-fill used instead of hipMemsetAsync so that a caller may capture the
// enqueued work in a HIP graph (memset nodes are avoided on this platform).
template <typename T>
__global__ void fill_kernel(T *__restrict__ p, T v, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

template <typename T>
inline int fill_async(T *p, T v, int64_t n, hipStream_t st) {
    if (n <= 0) return OK;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(fill_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, st, p, v, n);
    return check_launch("fill_kernel");
}

inline bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

}  // namespace matpl

#define MATPL_CHECK_ARG(cond, ...)                 \
    do {                                           \
        if (!(cond)) {                             \
            matpl::set_error(__VA_ARGS__);         \
            return matpl::ERR_ARGUMENT;            \
        }                                          \
    } while (0)

#define MATPL_EXPORT extern "C" __attribute__((visibility("default")))
