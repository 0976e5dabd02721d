// reduce_device.h -- workgroup reductions feeding one FP64 atomic per workgroup.
#pragma once

#include "common.h"

namespace eddy {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_down(v, off, 64));
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
    return v;
}

// Sum v over the workgroup and atomically add it to *dst (one atomic per workgroup).
// Every thread of the workgroup must call it (contains barriers).
__device__ __forceinline__ void block_sum_atomic(double v, double* dst) {
    __shared__ double part[16];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) part[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < nw; ++w) s += part[w];
        atomicAdd(dst, s);
    }
}

}  // namespace eddy
