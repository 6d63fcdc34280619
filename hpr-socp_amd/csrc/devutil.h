// devutil.h — device helpers shared by the kernels: NaN-propagating max/clamp
// (the reference's Julia max/min/clamp semantics) and deterministic wave /
// workgroup reductions (wave64 butterflies, fixed-order combination).
#pragma once
#include <hip/hip_runtime.h>

namespace hpr {

__device__ __forceinline__ double nmax(double a, double b) { return (a > b || a != a) ? a : b; }
template <typename T>
__device__ __forceinline__ T jclamp(T z, T lo, T hi) {
    return (z != z) ? z : fmin(fmax(z, lo), hi);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = nmax(v, __shfl_xor(v, o, 64));
    return v;
}
// Deterministic workgroup reductions (256 threads); result valid in every thread.
__device__ __forceinline__ double block_sum(double v, double *sh) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    return (sh[0] + sh[1]) + (sh[2] + sh[3]);
}
__device__ __forceinline__ double block_max(double v, double *sh) {
    v = wave_max(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    return nmax(nmax(sh[0], sh[1]), nmax(sh[2], sh[3]));
}

}  // namespace hpr
