// stencil_device.h -- device helpers shared by the stencil kernels and the fused y-FFT:
// the reference's 7-point divergence and the operator DIRECT inverts, at one interior cell.
#pragma once

#include "kernels.h"

namespace eddy {

// Load at a 32-bit byte offset from a uniform base: global_load's scalar-base form, one VGPR
// per address and no 64-bit address arithmetic. Every field and work array here is far
// below 4 GiB (the plan rejects larger grids).
template <typename V>
__device__ __forceinline__ V ld_off(const V* base, unsigned byte_off) {
    return *reinterpret_cast<const V*>(reinterpret_cast<const char*>(base) + byte_off);
}

// Store at a 32-bit byte offset from a base (the store form of ld_off).
template <typename V>
__device__ __forceinline__ void st_off(V* base, unsigned byte_off, V v) {
    *reinterpret_cast<V*>(reinterpret_cast<char*>(base) + byte_off) = v;
}

__device__ __forceinline__ size_t sidx(const Dims& d, int I, int J, int K) {
    return (size_t)I + (size_t)d.nx * ((size_t)J + (size_t)d.ny * (size_t)K);
}

// DIVERGENCE at interior cell (i, j, k), 0-based interior indices (timeindeg.f:944-946).
__device__ __forceinline__ double div_at(const Dims& d, const Metrics& mt, const double* __restrict__ us,
                                         const double* __restrict__ vs, const double* __restrict__ ws, int i, int j,
                                         int k) {
    const int I = i + 1, J = j + 1, K = k + 1;
    const size_t o = sidx(d, I, J, K);
    const double c = mt.cp[K] * (ws[o] - ws[o - (size_t)d.nx * d.ny]);
    if (mt.cart) {  // RU = RP = 1: the same arithmetic without the unit factors
        const double a = mt.ap[I] * (us[o] - us[o - 1]);
        const double b = mt.bp[J] * (vs[o] - vs[o - d.nx]);
        return a + b + c;
    }
    const double rpi = mt.rp[I];
    const double a = mt.ap[I] * (mt.ru[I] * us[o] - mt.ru[I - 1] * us[o - 1]) / rpi;
    const double b = mt.bp[J] * (vs[o] - vs[o - d.nx]) / rpi;
    return a + b + c;
}

// The operator DIRECT inverts (direct.f:344-352, 370-378, 536), applied at (i, j, k); the
// walls enter through AM(1) = CM(M) = AN(1) = CN(N) = 0, the y direction is periodic, and a
// periodic x1 (x2) couples cells 1 and M (1 and N) through AM(1), CM(M) (AN(1), CN(N)).
// Only interior values of phi are read.
__device__ __forceinline__ double lap_at(const Dims& d, const Op64& op, const double* __restrict__ phi, int i,
                                         int j, int k) {
    const int I = i + 1, J = j + 1, K = k + 1;
    const size_t o = sidx(d, I, J, K);
    const double p0 = phi[o];
    double sx = op.bm[i] * p0;
    if (i > 0) sx += op.am[i] * phi[o - 1];
    else if (d.cyc_x) sx += op.am[0] * phi[o + (d.M - 1)];
    if (i < d.M - 1) sx += op.cm[i] * phi[o + 1];
    else if (d.cyc_x) sx += op.cm[i] * phi[o - (d.M - 1)];
    const int jm = (j == 0) ? d.L : J - 1;
    const int jp = (j == d.L - 1) ? 1 : J + 1;
    const double sy = (phi[sidx(d, I, jp, K)] - 2.0 * p0 + phi[sidx(d, I, jm, K)]) * op.inv_delysq * op.r2inv[i];
    double sz = op.bn[k] * p0;
    const size_t pl = (size_t)d.nx * d.ny;
    if (k > 0) sz += op.an[k] * phi[o - pl];
    else if (d.cyc_z) sz += op.an[0] * phi[o + pl * (size_t)(d.N - 1)];
    if (k < d.N - 1) sz += op.cn[k] * phi[o + pl];
    else if (d.cyc_z) sz += op.cn[k] * phi[o - pl * (size_t)(d.N - 1)];
    return sx + sy + sz;
}

// Volume RP/(AP*BP*CP) of interior cell (i, j, k) from its separable factors.
__device__ __forceinline__ double cell_vol(const Metrics& mt, int i, int j, int k) {
    return mt.volx[i + 1] * mt.voly[j + 1] * mt.volz[k + 1];
}

}  // namespace eddy
