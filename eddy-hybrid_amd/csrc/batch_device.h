// batch_device.h -- the work items of the batched (many-problem) launches.
//
// A batched stage is ONE launch for all problems, one workgroup per work item (one tile of
// one problem). Items are numbered problem by problem, problems ordered by decreasing
// per-tile cost (longest first), so the dispatcher hands out the long tiles first; workgroup b
// takes item b, and consecutive items land on the 8 XCDs in turn. A workgroup finds its item's
// problem by binary search over each problem's first item. The stage's tile function is
// inlined with the job's arguments read through a uniform pointer (scalar loads), so a batch
// kernel holds the registers of the same tile in a plan kernel, not of a call.
#pragma once

#include "common.h"

namespace eddy {

template <typename Args>
struct Job {
    Args a;
    int ntiles;  // work items of this problem in this stage
    int ntx;     // tiles per row (tile -> (x1 block, row))
    int first;   // first item number
    int aux;     // stage-specific selector (x2 transform kind, x1-solve chunk template)
};

// The item of this workgroup (launches have exactly `total` workgroups).
template <typename Args, typename Body>
__device__ __forceinline__ void run_item(const Job<Args>* __restrict__ jobs, int njobs, int total, Body body) {
    const int it = __builtin_amdgcn_readfirstlane((int)blockIdx.x);
    if (it >= total) return;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].first <= it) lo = mid;
        else hi = mid - 1;
    }
    body(jobs[lo], it - jobs[lo].first);
}

}  // namespace eddy
