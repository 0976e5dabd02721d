// noalloc.hip -- eddy_direct_noalloc_dev: FP64 DIRECT whose work buffers live in a fixed
// device array instead of plan-owned allocations, so that the call itself allocates nothing
// (for callers under hipGraph capture or in loops that must not allocate).
//
// The array `scratch` is a file-scope __device__ variable: every GPU the library runs on holds
// its own copy, at its own address, from module load until the process ends. Host code never
// uses the array's name as an address (in host code it is only a stand-in symbol): the device
// address is looked up once per device with hipGetSymbolAddress and kept per device.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "plan.h"

namespace eddy {

// 64 MiB: grids up to about 8.3 million interior cells (e.g. 256 x 128 x 254)
constexpr size_t kScratchDoubles = (size_t)64 * 1024 * 1024 / sizeof(double);

static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (looked up on the first call there).
static int scratch_on_device(double** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return device_error(e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device index beyond the scratch table");
        return EDDY_ERR_DEVICE;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return device_error(e, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

}  // namespace eddy

using namespace eddy;

extern "C" {

int eddy_direct_noalloc_dev(eddy_plan_t plan, double* d_dp) {
    if (!plan || !d_dp) {
        set_error("null plan or DP");
        return EDDY_ERR_ARG;
    }
    const size_t need = direct64_work_doubles(plan);
    if (need > kScratchDoubles) {
        set_error("eddy_direct_noalloc_dev: the grid needs " + std::to_string(need * sizeof(double)) +
                  " bytes of work buffers, beyond the fixed 64 MiB scratch array");
        return EDDY_ERR_UNSUPPORTED;
    }
    double* sp = nullptr;
    int rc = scratch_on_device(&sp);
    if (rc) return rc;
    const size_t M = plan->d.M, L = plan->d.L, N = plan->d.N;
    WorkBuffers wb;
    wb.S = sp;
    wb.linesum = wb.S + M * L * N;
    wb.part = wb.linesum + (N > 0 ? N : 1);
    wb.red = wb.part + 2 * M * N;
    return direct64_with_buffers(plan, d_dp, wb);
}

}  // extern "C"
