// reduce_device.h -- workgroup reductions. A workgroup reduces its values and writes one
// partial per workgroup (no atomics: same-address device atomics from thousands of
// workgroups serialize); a one-workgroup kernel then folds the partials into the scalar.
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

__device__ __forceinline__ int linear_tid() {
    return threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
}

__device__ __forceinline__ int linear_bid() { return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z); }

enum ReduceOp { RED_OP_SUM = 0, RED_OP_MIN = 1, RED_OP_MAX = 2 };

template <int OP>
__device__ __forceinline__ double combine(double a, double b) {
    if (OP == RED_OP_SUM) return a + b;
    if (OP == RED_OP_MIN) return fmin(a, b);
    return fmax(a, b);
}

template <int OP>
__device__ __forceinline__ double wave_reduce(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = combine<OP>(v, __shfl_down(v, off, 64));
    return v;
}

// Reduce v over the workgroup; thread 0 stores the result to part[slot].
// Every thread must call it (contains barriers). Works for 1-D, 2-D and 3-D blocks.
template <int OP = RED_OP_SUM>
__device__ __forceinline__ void block_partial(double v, double* part, int slot) {
    __shared__ double wparts[16];
    const int t = linear_tid();
    const int nthr = blockDim.x * blockDim.y * blockDim.z;
    const int lane = t & 63, wid = t >> 6, nw = (nthr + 63) >> 6;
    v = wave_reduce<OP>(v);
    __syncthreads();
    if (lane == 0) wparts[wid] = v;
    __syncthreads();
    if (t == 0) {
        double s = wparts[0];
        for (int w = 1; w < nw; ++w) s = combine<OP>(s, wparts[w]);
        part[slot] = s;
    }
}

}  // namespace eddy
