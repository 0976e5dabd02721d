// plan.h -- the device plan: the explicit form of the state DIRECT SAVEs on its first call
// (direct.f:24-51, 121-585): coefficients, wavenumbers, transform tables and work buffers.
#pragma once

#include <string>
#include <vector>

#include "../../include/eddyproj.h"
#include "common.h"
#include "kernels.h"

namespace eddy {

template <typename T>
struct DevTables {
    std::vector<void*> owned;
    Tables<T> t{};
};

struct TimerSlot {
    long launches = 0;
    double total_ms = 0.0;
};

struct PendingTiming {
    int kernel;
    hipEvent_t start, stop;
};

}  // namespace eddy

struct eddy_plan_s {
    eddy::Dims d{};
    int mp = 1, lp = 0, np = 1;
    double delysq = 0.0;
    int device = 0;
    hipStream_t stream = nullptr;
    eddy_tuning tune{};
    int dense = 0;  // 1: dense x2 eigen-transform, 0: DCT
    // corner couplings of periodic lines (MP=0: alpha = CM(M), beta = AM(1) of the x1 lines;
    // NP=0: alpha = CN(N), beta = AN(1) of the x2 lines of the x1 form), 0 with walls
    double xalpha = 0.0, xbeta = 0.0, zalpha = 0.0, zbeta = 0.0;
    eddy::FftPlan fpy{}, fpz{};

    // FP64 operator (device)
    eddy::Op64 op{};
    eddy::Metrics mt{};
    eddy::DevTables<double> t64;
    eddy::DevTables<float> t32;
    std::vector<void*> owned;  // device allocations freed with the plan

    double* S = nullptr;        // spectral work array, M*L*N doubles (or 2x floats)
    double* S2 = nullptr;       // second spectral array of the GEMM transforms (zgemm or xgemm)
    void* blas = nullptr;       // rocBLAS handle of the GEMM transforms (zgemm or xgemm)
    int zgemm = 0;              // 1: dense x2 transform as batched GEMMs S <-> S2
    // xgemm: the stretched-x2 case diagonalised along x1 instead (x1 shorter than x2): per-slot
    // x1 eigenvector products S <-> S2 (S2 x2-fastest), then tridiagonal x2 lines solved by
    // the line-solve kernel with the tables tz (x2 operator, x1 eigenvalues per line) and
    // dims dz (line length N, M lines per y-mode)
    int xgemm = 0;
    eddy::Dims dz{};
    eddy::DevTables<double> tz64;
    eddy::DevTables<float> tz32;
    double* linesum = nullptr;  // N doubles (also used as N floats)
    double* red = nullptr;      // reduction scalars
    double* part = nullptr;     // per-workgroup reduction partials
    double* rhs = nullptr;      // RHS copy for the defect correction (NX*NY*NZ), lazily allocated
    double* dp_dev = nullptr;   // device DP for the host-pointer entry point, lazily allocated
    const double* vol_total_dev = nullptr;  // mt.vol_total on the device (batched folds read it)
    // pressure update (eddy_plan_set_pressure_bc): implicit masks (device, NX) and BC types
    const double* rvimplx = nullptr;
    const double* rvimply = nullptr;
    double ru1 = 0.0;
    int implx = 0, imply = 0, itype1 = 0, itype5 = 1, itype6 = 1;
    size_t bytes = 0;

    // timing
    int timing = 0;
    eddy::TimerSlot slots[EDDY_K_COUNT];
    std::vector<eddy::PendingTiming> pending;
    std::vector<hipEvent_t> event_pool;
};

namespace eddy {
void set_error(const std::string& s);
int device_error(hipError_t e, const char* what);

// Work buffers of one FP64 DIRECT (device memory): the spectral array (M*L*N), the line sums
// (N), the reduction partials (2*M*N) and scalars (RED_COUNT).
struct WorkBuffers {
    double *S, *linesum, *part, *red;
};
// FP64 DIRECT on d_dp with the given work buffers instead of the plan's (no allocation);
// EDDY_ERR_UNSUPPORTED for plans that take the GEMM forms.
int direct64_with_buffers(eddy_plan_s* P, double* dp, const WorkBuffers& wb);
// Doubles of work buffers direct64_with_buffers needs for the plan.
size_t direct64_work_doubles(const eddy_plan_s* P);
}  // namespace eddy
