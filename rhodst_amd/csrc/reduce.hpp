// reduce.hpp -- deterministic two-level FP64 reductions shared by the HIP
// units: a 64-lane wavefront shuffle tree, then the wave sums in a fixed
// order; a second kernel (or every block of the next kernel) re-reduces the
// G block partials in the same order, so results do not depend on timing.
#pragma once
#include <hip/hip_runtime.h>

namespace rdst {

constexpr int RED_BS = 256;     // threads per reduction block
constexpr int RED_MAXG = 1024;  // max blocks of a reduction kernel

__device__ inline double block_sum(double v, double* sh) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) sh[wv] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += sh[k];
    __syncthreads();
    return t;  // valid in thread 0
}

__device__ inline double block_min(double v, double* sh) {
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_down(v, off, 64));
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) sh[wv] = v;
    __syncthreads();
    double t = 1.0e300;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t = fmin(t, sh[k]);
    __syncthreads();
    return t;  // valid in thread 0
}

// every block sums the G partials of the previous kernel in the same order
__device__ inline double reduce_partials(const double* part, int G, double* sh) {
    double v = 0.0;
    for (int i = threadIdx.x; i < G; i += blockDim.x) v += part[i];
    double t = block_sum(v, sh);
    if (threadIdx.x == 0) sh[8] = t;
    __syncthreads();
    double r = sh[8];
    __syncthreads();
    return r;
}

inline int reduce_grid(long n) {
    long g = (n + RED_BS - 1) / RED_BS;
    return (int)(g < 1 ? 1 : (g > RED_MAXG ? RED_MAXG : g));
}

}  // namespace rdst
