// batch_device.h -- the device work queue of the batched (many-problem) launches.
//
// A batched stage is ONE persistent launch for all problems: workgroups pull work items (one
// tile of one problem) from device counters until the queue is empty. Items are numbered
// problem by problem, problems ordered by decreasing per-tile cost (longest first). The
// items are dealt round-robin to 8 queues, one per XCD (workgroup b serves queue b % 8, the
// XCD it is dispatched to), and a workgroup claims kQueueChunk items per atomic, so a few
// thousand atomics spread over 8 counters drive hundreds of thousands of tiles. A workgroup
// finds an item's problem by binary search over each problem's first item, and leaves the
// loop once its queue is exhausted.
#pragma once

#include "common.h"

namespace eddy {

constexpr int kQueues = 8;
constexpr int kQueueStride = 32;  // ints between counters: one 128-byte line each
constexpr int kQueueChunk = 16;

template <typename Args>
struct Job {
    Args a;
    int ntiles;  // work items of this problem in this stage
    int ntx;     // tiles per row (tile -> (x1 block, row))
    int first;   // first item number
    int aux;     // stage-specific selector (x2 transform kind, x1-solve chunk template)
};

template <typename Args, typename Body>
__device__ __forceinline__ void run_queue(const Job<Args>* __restrict__ jobs, int njobs, int total, int* counters,
                                          Body body) {
    __shared__ int s_base;
    const int q = blockIdx.x & (kQueues - 1);
    int* cnt = counters + q * kQueueStride;
    const int nq = (total > q) ? (total - q + kQueues - 1) / kQueues : 0;
    for (;;) {
        if (threadIdx.x == 0) s_base = atomicAdd(cnt, kQueueChunk);
        __syncthreads();
        // uniform across the workgroup: keep it (and the job search and the job's fields that
        // follow from it) in scalar registers
        const int base = __builtin_amdgcn_readfirstlane(s_base);
        __syncthreads();
        if (base >= nq) break;
        const int end = (base + kQueueChunk < nq) ? base + kQueueChunk : nq;
        for (int c = base; c < end; ++c) {
            const int it = q + kQueues * c;
            int lo = 0, hi = njobs - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (jobs[mid].first <= it) lo = mid;
                else hi = mid - 1;
            }
            body(jobs[lo], it - jobs[lo].first);
            __syncthreads();  // the next item reuses the LDS
        }
    }
}

}  // namespace eddy
