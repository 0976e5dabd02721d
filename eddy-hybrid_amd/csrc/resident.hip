// resident.hip -- kernels of the resident time loop: a substep's predicted velocity made
// on the device from the previous corrected one plus seeded noise, and the per-step log of
// the post-projection divergence range (what SCRINFO prints, info.f:5-165).
//
// Noise: three N(0,1) draws per cell from Philox4x32-10 (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11) with counter {cell lo, cell hi,
// step, 0} and key {seed, 0}, Box-Muller on words (0,1) and (2,3). `cell` is the 0-based
// offset of the cell in the DP(NX,NY,NZ) array and `step` is read from device memory, so a
// captured graph draws fresh noise on every replay.
#include "kernels.h"
#include "reduce_device.h"
#include "stencil_device.h"

namespace eddy {

struct Words {
    uint32_t w0, w1, w2, w3;
};

__device__ __forceinline__ Words philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                               uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        // one 32 x 32 -> 64-bit product per multiplier (a single v_mad_u64_u32) gives both halves
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c1 = lo1;
        c3 = lo0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return Words{c0, c1, c2, c3};
}

// Both deviates of a pair (one sincos of the angle). The angle 2 pi u1 is taken as pi times
// the exact 2 u1 (sincospi): no rounding of the argument and no general-argument reduction,
// which made the FP64 trigonometry the larger part of this kernel's instructions. A libm
// sin/cos of the rounded 2 pi u1 differs in the last bits of the deviate.
__device__ __forceinline__ void box_muller(uint32_t a, uint32_t b, double& g0, double& g1) {
    const double u0 = ((double)a + 0.5) * 2.3283064365386962890625e-10;  // (a + 1/2) 2^-32
    const double u1 = ((double)b + 0.5) * 2.3283064365386962890625e-10;
    const double r = sqrt(-2.0 * log(u0));
    double sn, cs;
    sincospi(2.0 * u1, &sn, &cs);
    g0 = r * cs;
    g1 = r * sn;
}

// The first deviate of a pair only.
__device__ __forceinline__ double box_muller_first(uint32_t a, uint32_t b) {
    const double u0 = ((double)a + 0.5) * 2.3283064365386962890625e-10;
    const double u1 = ((double)b + 0.5) * 2.3283064365386962890625e-10;
    return sqrt(-2.0 * log(u0)) * cospi(2.0 * u1);
}

// One thread per (I, J, K) of the full array, I fastest. Faces CORRECT updates get
// sigma * N(0,1) (U: I = 1..NX-3, V: all, W: K = 1..NZ-3, 0-based, over interior cells);
// periodic ghost rows J = 0, NY-1 take their images' values; the rest is copied.
__global__ void __launch_bounds__(256) k_perturb(Dims d, const double* __restrict__ u, const double* __restrict__ v,
                                                 const double* __restrict__ w, double* __restrict__ us,
                                                 double* __restrict__ vs, double* __restrict__ ws, double sigma,
                                                 uint32_t seed, const int* __restrict__ step) {
    const int I = blockIdx.x * 64 + threadIdx.x;
    const int J = blockIdx.y * 4 + threadIdx.y;
    const int K = blockIdx.z;
    if (I >= d.nx || J >= d.ny) return;
    const size_t t = sidx(d, I, J, K);
    const bool inner = I >= 1 && I <= d.M && K >= 1 && K <= d.N;
    if (!inner) {
        us[t] = u[t];
        vs[t] = v[t];
        ws[t] = w[t];
        return;
    }
    const int Js = (J == 0) ? d.L : ((J == d.L + 1) ? 1 : J);
    const size_t o = sidx(d, I, Js, K);
    const Words r = philox4x32_10((uint32_t)o, (uint32_t)(o >> 32), (uint32_t)step[0], 0u, seed, 0u);
    double g0, g1;
    box_muller(r.w0, r.w1, g0, g1);
    const double g2 = box_muller_first(r.w2, r.w3);  // the pair's second deviate is not used
    us[t] = (I <= d.nx - 3) ? u[o] + sigma * g0 : u[t];
    vs[t] = v[o] + sigma * g1;
    ws[t] = (K <= d.nz - 3) ? w[o] + sigma * g2 : w[t];
}

// Step epilogue: log[2 s], log[2 s + 1] = red[RED_DMIN], red[RED_DMAX] for s = *step < cap,
// then *step = s + 1. One thread.
__global__ void k_log_step(const double* __restrict__ red, double* __restrict__ log, int* __restrict__ step, int cap) {
    if (threadIdx.x != 0) return;
    const int s = step[0];
    if (s >= 0 && s < cap) {
        log[2 * s] = red[RED_DMIN];
        log[2 * s + 1] = red[RED_DMAX];
    }
    step[0] = s + 1;
}

void launch_perturb(hipStream_t st, const Dims& d, const double* u, const double* v, const double* w, double* us,
                    double* vs, double* ws, double sigma, uint32_t seed, const int* step) {
    const dim3 grid(ceil_div(d.nx, 64), ceil_div(d.ny, 4), d.nz);
    hipLaunchKernelGGL(k_perturb, grid, dim3(64, 4, 1), 0, st, d, u, v, w, us, vs, ws, sigma, seed, step);
}

void launch_log_step(hipStream_t st, const double* red, double* log, int* step, int cap) {
    hipLaunchKernelGGL(k_log_step, dim3(1), dim3(64), 0, st, red, log, step, cap);
}

}  // namespace eddy
