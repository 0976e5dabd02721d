// batch_device.h -- the work items of the batched (many-problem) launches.
//
// A batched stage is ONE launch for all problems, one workgroup per work item (one tile of
// one problem). Items are numbered problem by problem, problems ordered by decreasing
// per-tile cost (longest first), so the dispatcher hands out the long tiles first; runs of 64
// consecutive items go to one XCD (xcd_run). A workgroup finds its item's
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
    int aux;     // stage-specific: x2 transform kind, x1-solve rows-per-thread class, y blocks per plane
};

// Item of workgroup b when runs of C consecutive items share an XCD. Workgroups are dealt
// to the 8 XCDs in turn (XCD b % 8), so workgroup b takes item ((s / C) * 8 + b % 8) * C + s % C
// with s = b / 8: a tile's x1 and x2 neighbours (the items next to it) are read through the
// same L2. Bijective on the first total - total % (8 C) items; the rest keep item b.
__device__ __forceinline__ int xcd_run(int b, int total, int C) {
    const int full = total - total % (8 * C);
    if (b >= full) return b;
    const int s = b >> 3;
    return ((s / C) * 8 + (b & 7)) * C + s % C;
}

constexpr int kXcdRun = 64;

// The item of this workgroup (launches have exactly `total` workgroups).
template <typename Args, typename Body>
__device__ __forceinline__ void run_item(const Job<Args>* __restrict__ jobs, int njobs, int total, Body body) {
    const int b = __builtin_amdgcn_readfirstlane((int)blockIdx.x);
    if (b >= total) return;
    const int it = xcd_run(b, total, kXcdRun);
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].first <= it) lo = mid;
        else hi = mid - 1;
    }
    body(jobs[lo], it - jobs[lo].first);
}

}  // namespace eddy
