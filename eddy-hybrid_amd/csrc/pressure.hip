// pressure.hip -- the pressure update that closes the substep after CORRECT
// (Eddy6.f:2772-2846):
//   P = P + DP, minus the implicit-viscous corrections when IMPLY / IMPLX are on
//     (:2777-2796, RU1 and the RVIMPLY / RVIMPLX masks per x1 index);
//   PREF = sum(P(IX1, JY1:JY2, KZ1)) / (NY-2) subtracted everywhere (:2801-2806);
//   ghost boundary values: x (:2813-2824, MP and the ITYPE(1) = 300 axis with JSYM), y
//   (:2829-2835, LP), z (:2843-2847, ITYPE(5) / ITYPE(6): 500 periodic, else Neumann).
//
// Three launches: the reference value (one workgroup), the interior update, and the ghost
// shell. The reference applies its boundary copies one axis after the other over whole
// planes; every ghost value therefore equals +-P at one interior cell, found by mapping the
// z index, then the y index, then the x index (with its sign and the axis reflection), and
// the ghost kernel gathers it from the already updated interior.
#include <algorithm>

#include "kernels.h"
#include "reduce_device.h"
#include "stencil_device.h"

namespace eddy {

// P + DP minus the implicit corrections at interior cell (I, J, K) (0-based with ghosts).
__device__ __forceinline__ double p_updated(const PressArgs& a, int I, int J, int K) {
    const Dims d = a.d;
    const size_t o = sidx(d, I, J, K);
    double v = a.p[o] + a.dp[o];
    if (a.imply) {
        const double r = a.dely * a.mt.rp[I];
        v = v - 0.5 * a.alfxdt * a.ru1 / (r * r) * (a.dp[o - d.nx] - 2. * a.dp[o] + a.dp[o + d.nx]) * a.rvy[I];
    }
    if (a.implx) {
        v = v - 0.5 * a.alfxdt * a.ru1 * a.mt.ap[I] / a.mt.rp[I] *
                    ((a.dp[o + 1] - a.dp[o]) * a.mt.au[I] * a.mt.ru[I] -
                     (a.dp[o] - a.dp[o - 1]) * a.mt.au[I - 1] * a.mt.ru[I - 1]) *
                    a.rvx[I];
    }
    return v;
}

// PREF = mean over the interior y rows of the updated P at (IX1, KZ1) -> red[RED_PREF].
__global__ void __launch_bounds__(256) k_pref(PressArgs a) {
    double s = 0.0;
    for (int J = 1 + threadIdx.x; J <= a.d.L; J += blockDim.x) s += p_updated(a, 1, J, 1);
    __shared__ double part[4];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) t += part[w];
        a.red[RED_PREF] = t / (double)a.d.L;
    }
}

__global__ void __launch_bounds__(256) k_p_interior(PressArgs a) {
    const int I = blockIdx.x * 64 + threadIdx.x + 1;
    const int J = blockIdx.y * 4 + threadIdx.y + 1;
    const int K = blockIdx.z + 1;
    if (I > a.d.M || J > a.d.L) return;
    const double v = p_updated(a, I, J, K);
    a.p[sidx(a.d, I, J, K)] = v - a.red[RED_PREF];
}

// Ghost cell (I, J, K) <- sign * P(interior source), the boundary copies composed z, y, x.
__device__ __forceinline__ void p_ghost(const PressArgs& a, int I, int J, int K) {
    const Dims d = a.d;
    int k = K;
    if (K == 0) k = (a.itype5 == 500) ? d.nz - 2 : 1;
    else if (K == d.nz - 1) k = (a.itype6 == 500) ? 1 : d.nz - 2;
    int j = J;
    if (J == 0) j = (a.lp == 0) ? d.ny - 2 : 1;
    else if (J == d.ny - 1) j = (a.lp == 0) ? 1 : d.ny - 2;
    int i = I;
    double s = 1.0;
    if (I == 0) {
        if (a.itype1 == 300) {  // axis: P(1,J,:) = P(IX1, JSYM(J), :), JSYM(J) = J + (NY-2)/2 mod the period
            i = 1;
            j += d.L / 2;
            if (j > d.L) j -= d.L;
        } else if (a.mp == 0) {
            i = d.nx - 2;
        } else {
            i = 1;
            s = -1.0;
        }
    } else if (I == d.nx - 1) {
        if (a.mp == 0) {
            i = 1;
        } else {
            i = d.nx - 2;
            s = -1.0;
        }
    }
    a.p[sidx(d, I, J, K)] = s * a.p[sidx(d, i, j, k)];
}

// The ghost shell as three pairs of planes: blockIdx.z selects the face (0/1: x, 2/3: y,
// 4/5: z); edges are written by two faces with the same value.
__global__ void __launch_bounds__(256) k_p_ghosts(PressArgs a) {
    const Dims d = a.d;
    const int face = blockIdx.z;
    const int u = blockIdx.x * 256 + threadIdx.x;
    const int v = blockIdx.y;
    int I, J, K;
    if (face < 2) {  // x faces: u over J, v over K
        if (u >= d.ny || v >= d.nz) return;
        I = face ? d.nx - 1 : 0;
        J = u;
        K = v;
    } else if (face < 4) {  // y faces: u over I, v over K
        if (u >= d.nx || v >= d.nz) return;
        I = u;
        J = (face & 1) ? d.ny - 1 : 0;
        K = v;
    } else {  // z faces: u over I, v over J
        if (u >= d.nx || v >= d.ny) return;
        I = u;
        J = v;
        K = (face & 1) ? d.nz - 1 : 0;
    }
    p_ghost(a, I, J, K);
}

void launch_pressure_update(hipStream_t st, const PressArgs& a) {
    const Dims& d = a.d;
    hipLaunchKernelGGL(k_pref, dim3(1), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_p_interior, dim3(ceil_div(d.M, 64), ceil_div(d.L, 4), d.N), dim3(64, 4, 1), 0, st, a);
    const int umax = std::max(d.nx, d.ny);
    const int vmax = std::max(d.ny, d.nz);
    hipLaunchKernelGGL(k_p_ghosts, dim3(ceil_div(umax, 256), vmax, 6), dim3(256), 0, st, a);
}

}  // namespace eddy
