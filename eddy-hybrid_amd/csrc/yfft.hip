// yfft.hip -- batched real FFTs along the periodic axis (the reference's Y, 2nd index).
//
// Forward: direct.f:620-629 (RFFTF of every interior pencil DP(I+1,2:L+1,K+1));
// inverse: direct.f:1035-1106 (DUMMY/L, RFFTB, ghost fill). Output packing follows
// fftpack (fftpack.f:140-240): [Re X0, Re X1, Im X1, ..., Re X_{L/2}] (L even).
//
// Periodic lengths the radix FFT cannot take (L odd, or L/2 with a prime factor >= 7) go through
// the same tiles with the transform done as a dense real DFT (the DENSE kernels): the tile holds
// TI real rows of length L in LDS and every output is its O(L) sum against the table
// exp(-2 pi i t/L), t < L, the packing and scaling as above.
//
// A workgroup owns TI (a power of two) consecutive x1 pencils at one x2 index k. The pencils
// are read row by row (TI contiguous doubles per y index: coalesced), with several loads in
// flight per thread, packed as TI complex sequences z[n] = x[2n] + i x[2n+1] of length
// H = L/2 in LDS, transformed by a Stockham FFT and unpacked to the L real coefficients. The
// spectral array S keeps x1 fastest: S[i + M*(k + N*j)], so the x2 transform reads rows of
// contiguous values too.
#include <algorithm>
#include <numeric>
#include <type_traits>

#include "batch_device.h"
#include "fft_device.h"
#include "kernels.h"
#include "reduce_device.h"
#include "stencil_device.h"

namespace eddy {

constexpr int kLoadBatch = 8;
constexpr int kInvTasks = 2;  // packing tasks (4 coefficient rows each) per thread and load batch
constexpr int kDivRows = 4;  // rows per load sub-batch of the fused divergence and residual
// Kernel-side forms of YF_DIVERGENCE with the metric kind fixed at compile time, so that the
// Cartesian kernel does not carry the registers of the cylindrical one's divisions by RP.
constexpr int YF_DIV_CART = 3, YF_DIV_CYL = 4;

template <typename T>
struct YArgs {
    const double* src;
    const double* phi;
    const double *us, *vs, *ws;  // YF_DIVERGENCE inputs
    Metrics mt;
    double scale;
    const double* dscale;  // device {ALFXDT, 1/ALFXDT}: when non-null, scale = dscale[1]
    Op64 op;
    double* part;
    double* red;
    T* S;
    const T* Sin;
    double* dp;
    double* rhs_out;  // FP32 YF_DIVERGENCE: the FP64 divergence, DP layout
    Dims d;
    FftPlan fp;
    Tables<T> tb;
    int TI, lti;  // pencils per workgroup and its log2
    int iblk;     // forward: tiles per x1 block of the tile order (iblock_tile), 0 = whole rows
    int ld;
    int shift;
};

__device__ __forceinline__ size_t dpidx(const Dims& d, int I, int J, int K) { return sidx(d, I, J, K); }

// The metric factors of the thread's x1 index and the tile's x2 index (loaded once).
struct DivFactors {
    double cpk, apI, vx, vz, ruI, ruIm, rpI;
};

template <bool CART>
__device__ __forceinline__ DivFactors div_factors(const Metrics& mt, int gi, int k) {
    const int I = gi + 1, K = k + 1;
    DivFactors f;
    f.cpk = mt.cp[K];
    f.apI = mt.ap[I];
    f.vx = mt.volx[I];
    f.vz = mt.volz[K];
    f.ruI = CART ? 1.0 : mt.ru[I];
    f.ruIm = CART ? 1.0 : mt.ru[I - 1];
    f.rpI = CART ? 1.0 : mt.rp[I];
    return f;
}

// B cells (gi, j0 + u*jstep, k) of the forward transform's input, DIVERGENCE * scale
// (div_at's arithmetic), with sum vol*b accumulated into vsum. Every field value is loaded
// before any is used: arithmetic on a value inside the guard of its own load makes the
// compiler wait for it before the next load is issued, one load in flight per wave. Rows
// past L read row L-1 and are zeroed afterwards, as is every row when !iok.
template <int B, bool CART>
__device__ __forceinline__ void divergence_rows(const Dims& d, const Metrics& mt, const DivFactors& f,
                                                const double* __restrict__ us, const double* __restrict__ vs,
                                                const double* __restrict__ ws, double scale, int gi, bool iok,
                                                int j0, int jstep, int k, double* v, double& vsum) {
    const int I = gi + 1, K = k + 1;
    // 32-bit byte offsets from the uniform field bases (one VGPR per address; the fields are
    // far below 4 GiB): the same offset serves u, v and w
    const unsigned pl8 = 8u * (unsigned)d.nx * (unsigned)d.ny, nx8 = 8u * (unsigned)d.nx;
    double wc[B], wm[B], uc[B], um[B], vc[B], vm[B], bpJ[B], vyJ[B];
#pragma unroll
    for (int u = 0; u < B; ++u) {
        const int J = min(j0 + u * jstep, d.L - 1) + 1;
        const unsigned o = 8u * ((unsigned)I + (unsigned)d.nx * ((unsigned)J + (unsigned)d.ny * (unsigned)K));
        wc[u] = ld_off(ws, o);
        wm[u] = ld_off(ws, o - pl8);
        uc[u] = ld_off(us, o);
        um[u] = ld_off(us, o - 8u);
        vc[u] = ld_off(vs, o);
        vm[u] = ld_off(vs, o - nx8);
    }
    // the y metric: a uniform y (DIRECT's FFT axis) takes the plan's scalars, no loads
    if (mt.bp0 != 0.0) {
#pragma unroll
        for (int u = 0; u < B; ++u) {
            bpJ[u] = mt.bp0;
            vyJ[u] = mt.vy0;
        }
    } else {
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const int J = min(j0 + u * jstep, d.L - 1) + 1;
            bpJ[u] = ld_off(mt.bp, 8u * (unsigned)J);
            vyJ[u] = ld_off(mt.voly, 8u * (unsigned)J);
        }
    }
    const double cpk = f.cpk, apI = f.apI, vx = f.vx, vz = f.vz, ruI = f.ruI, ruIm = f.ruIm, rpI = f.rpI;
#pragma unroll
    for (int u = 0; u < B; ++u) {
        const double c = cpk * (wc[u] - wm[u]);
        double x;
        if (CART) {
            x = apI * (uc[u] - um[u]) + bpJ[u] * (vc[u] - vm[u]) + c;
        } else {
            const double ax = apI * (ruI * uc[u] - ruIm * um[u]) / rpI;
            const double bx = bpJ[u] * (vc[u] - vm[u]) / rpI;
            x = ax + bx + c;
        }
        const bool ok = iok && j0 + u * jstep < d.L;
        v[u] = ok ? x * scale : 0.0;
        vsum += ok ? vx * vyJ[u] * vz * v[u] : 0.0;  // cell_vol's product order
    }
}

// B cells of the residual r = src - L(phi) (lap_at's arithmetic), sum r^2 into rsum; loads
// first, as in divergence_rows. The x and x2 neighbours at the walls are ghosts (in range)
// whose terms are dropped, as lap_at drops them.
template <int B>
__device__ __forceinline__ void residual_rows(const Dims& d, const Op64& op, const double* __restrict__ src,
                                              const double* __restrict__ phi, int gi, bool iok, int j0, int jstep,
                                              int k, double* v, double& rsum, double shift) {
    const int I = gi + 1, K = k + 1;
    // 32-bit byte offsets from the field bases (fields below 4 GiB); a neighbour step is added
    // modulo 2^32, so a negative step is its two's complement. Across a wall the neighbour is
    // the ghost (in range, its term dropped below), across a periodic x1 / x2 boundary the
    // interior cell at the other end (lap_at's coupling).
    const unsigned nx8 = 8u * (unsigned)d.nx, pl8 = nx8 * (unsigned)d.ny;
    const int i = gi;
    const unsigned oxm = (i == 0 && d.cyc_x) ? 8u * (unsigned)(d.M - 1) : 0u - 8u;
    const unsigned oxp = (i == d.M - 1 && d.cyc_x) ? 0u - 8u * (unsigned)(d.M - 1) : 8u;
    const unsigned ozm = (k == 0 && d.cyc_z) ? pl8 * (unsigned)(d.N - 1) : 0u - pl8;
    const unsigned ozp = (k == d.N - 1 && d.cyc_z) ? 0u - pl8 * (unsigned)(d.N - 1) : pl8;
    const unsigned oik = 8u * (unsigned)I + pl8 * (unsigned)K;
    double s[B], p0[B], pxm[B], pxp[B], pym[B], pyp[B], pzm[B], pzp[B];
#pragma unroll
    for (int u = 0; u < B; ++u) {
        const int j = min(j0 + u * jstep, d.L - 1);
        const int J = j + 1;
        const int jm = (j == 0) ? d.L : J - 1;
        const int jp = (j == d.L - 1) ? 1 : J + 1;
        const unsigned o = oik + nx8 * (unsigned)J;
        s[u] = ld_off(src, o);
        p0[u] = ld_off(phi, o);
        pxm[u] = ld_off(phi, o + oxm);
        pxp[u] = ld_off(phi, o + oxp);
        pym[u] = ld_off(phi, oik + nx8 * (unsigned)jm);
        pyp[u] = ld_off(phi, oik + nx8 * (unsigned)jp);
        pzm[u] = ld_off(phi, o + ozm);
        pzp[u] = ld_off(phi, o + ozp);
    }
    const double am = op.am[i], bm = op.bm[i], cm = op.cm[i], r2 = op.r2inv[i];
    const double an = op.an[k], bn = op.bn[k], cn = op.cn[k];
    const bool xlo = i > 0 || d.cyc_x, xhi = i < d.M - 1 || d.cyc_x;
    const bool zlo = k > 0 || d.cyc_z, zhi = k < d.N - 1 || d.cyc_z;
#pragma unroll
    for (int u = 0; u < B; ++u) {
        double sx = bm * p0[u];
        if (xlo) sx += am * pxm[u];
        if (xhi) sx += cm * pxp[u];
        const double sy = (pyp[u] - 2.0 * p0[u] + pym[u]) * op.inv_delysq * r2;
        double sz = bn * p0[u];
        if (zlo) sz += an * pzm[u];
        if (zhi) sz += cn * pzp[u];
        const bool ok = iok && j0 + u * jstep < d.L;
        v[u] = ok ? (s[u] - shift) - (sx + sy + sz) : 0.0;
        rsum += v[u] * v[u];
    }
}

// One tile (TI pencils at one x2 index) of the forward transform; `tile` is the logical tile,
// ntx the tiles per x2 row, `slot`/`nslots` the reduction partial index and count.
template <typename T, int MODE, bool DENSE = false>
__device__ __forceinline__ void yfft_fwd_tile(const YArgs<T>& a, int tile, int ntx, int slot, int nslots) {
    extern __shared__ unsigned char smem_raw[];
    T* smem = reinterpret_cast<T*>(smem_raw);
    const Dims d = a.d;
    const int TI = a.TI, lti = a.lti, ld = a.ld, H = d.L / 2, L = d.L;
    // FFT: TI complex rows (split re/im) transformed in place, then the two twiddle tables;
    // DENSE: TI real rows, then the table exp(-2 pi i t/L), t < L
    T* Ar = smem;
    T* Ai = Ar + TI * ld;
    T* twc_r = Ai + TI * ld;  // exp(-2 pi i t/H), t < H
    T* twc_i = twc_r + H;
    T* twr_r = DENSE ? Ar + TI * ld : twc_i + H;  // exp(-2 pi i k/L), k <= H (DENSE: k < L)
    T* twr_i = twr_r + (DENSE ? L : H + 1);
    const int i0 = (tile % ntx) * TI;
    const int k = tile / ntx;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int total = TI * L;
    if (DENSE) {
        for (int t = tid; t < L; t += nthr) {
            twr_r[t] = a.tb.ytwr_re[t];
            twr_i[t] = a.tb.ytwr_im[t];
        }
    } else {
        for (int t = tid; t <= H; t += nthr) {
            if (t < H) {
                twc_r[t] = a.tb.ytwc_re[t];
                twc_i[t] = a.tb.ytwc_im[t];
            }
            twr_r[t] = a.tb.ytwr_re[t];
            twr_i[t] = a.tb.ytwr_im[t];
        }
    }

    double shift = 0.0;
    if ((MODE == YF_DIRECT || MODE == YF_RESIDUAL) && a.shift) shift = a.red[RED_VOLB] / a.red[RED_VOL];
    double rsq = 0.0;
    // The thread's x1 offset is fixed (blockDim is a multiple of TI); its rows are
    // j0, j0 + jstep, ... Rows past L and pencils past M read a clamped in-range cell and are
    // zeroed after the loads, so each batch issues all its loads before using any.
    const int ii = tid & (TI - 1);
    const bool iok = i0 + ii < d.M;
    const int gi = iok ? i0 + ii : d.M - 1;
    const int jstep = nthr >> lti;
    auto put = [&](double x, int j) {  // row j of pencil ii into the packed complex buffer
        if (j < L) {
            if (DENSE) Ar[ii * ld + pidx(j)] = T(x);
            else if (j & 1) Ai[ii * ld + pidx((j >> 1))] = T(x);
            else Ar[ii * ld + pidx((j >> 1))] = T(x);
        }
    };
    for (int e0 = tid; e0 < total; e0 += nthr * kLoadBatch) {
        const int j0 = e0 >> lti;
        if (MODE == YF_DIRECT) {
            // 32-bit byte offsets from the field base (DP < 4 GiB)
            const unsigned nx8 = 8u * (unsigned)d.nx;
            const unsigned ok0 = nx8 * (unsigned)d.ny * (unsigned)(k + 1) + 8u * (unsigned)(gi + 1);
            double x[kLoadBatch];
#pragma unroll
            for (int u = 0; u < kLoadBatch; ++u) {
                const int j = min(j0 + u * jstep, L - 1);
                x[u] = ld_off(a.src, ok0 + nx8 * (unsigned)(j + 1));
            }
#pragma unroll
            for (int u = 0; u < kLoadBatch; ++u) put((iok && j0 + u * jstep < L) ? x[u] - shift : 0.0, j0 + u * jstep);
        } else if (MODE == YF_RESIDUAL) {
#pragma unroll
            for (int h = 0; h < kLoadBatch / kDivRows; ++h) {  // sub-batches as for the divergence below
                const int jh = j0 + h * kDivRows * jstep;
                double v[kDivRows];
                residual_rows<kDivRows>(d, a.op, a.src, a.phi, gi, iok, jh, jstep, k, v, rsq, shift);
#pragma unroll
                for (int r = 0; r < kDivRows; ++r) put(v[r], jh + r * jstep);
                __builtin_amdgcn_sched_barrier(0);
            }
        } else {
            const bool cart = MODE == YF_DIV_CART || (MODE == YF_DIVERGENCE && a.mt.cart);
            const DivFactors f = cart ? div_factors<true>(a.mt, gi, k) : div_factors<false>(a.mt, gi, k);
            const double scale = a.dscale ? a.dscale[1] : a.scale;
            // sub-batches of kDivRows rows: the loads of one sub-batch are in flight together;
            // the scheduling barrier keeps the compiler from hoisting the next sub-batch's loads
            // above this one's arithmetic, which would need registers for all of them at once
#pragma unroll
            for (int h = 0; h < kLoadBatch / kDivRows; ++h) {
                const int jh = j0 + h * kDivRows * jstep;
                double v[kDivRows];
                if (cart)
                    divergence_rows<kDivRows, true>(d, a.mt, f, a.us, a.vs, a.ws, scale, gi, iok, jh, jstep, k, v,
                                                    rsq);
                else
                    divergence_rows<kDivRows, false>(d, a.mt, f, a.us, a.vs, a.ws, scale, gi, iok, jh, jstep, k, v,
                                                     rsq);
#pragma unroll
                for (int r = 0; r < kDivRows; ++r) put(v[r], jh + r * jstep);
                if constexpr (std::is_same<T, float>::value) {
                    // the FP64 RHS of the defect correction, stored beside the FP32 transform
                    // input (the refined projection forms it nowhere else)
                    if (iok && a.rhs_out) {
                        const unsigned nx8 = 8u * (unsigned)d.nx;
                        const unsigned ok0 = nx8 * (unsigned)d.ny * (unsigned)(k + 1) + 8u * (unsigned)(gi + 1);
#pragma unroll
                        for (int r = 0; r < kDivRows; ++r)
                            if (jh + r * jstep < L) st_off(a.rhs_out, ok0 + nx8 * (unsigned)(jh + r * jstep + 1), v[r]);
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    }
    if (MODE == YF_RESIDUAL) block_partial<RED_OP_SUM>(rsq, a.part, slot);
    if (MODE == YF_DIVERGENCE || MODE == YF_DIV_CART || MODE == YF_DIV_CYL)
        block_partial<RED_OP_SUM>(rsq, a.part, slot);
    __syncthreads();
    const size_t js = (size_t)d.M * (size_t)d.N;
    T* outK = a.S + (size_t)d.M * (size_t)k;
    if (DENSE) {
        // packed slot s of pencil ii: sum_j x_j W^{q j}, q = (s+1)/2, its real part for s = 0
        // and odd s (the Nyquist slot L-1 of an even L included), its imaginary part for even s
        for (int e = tid; e < total; e += nthr) {
            const int ii = e & (TI - 1);
            const int sl = e >> lti;
            const int gi = i0 + ii;
            if (gi >= d.M) continue;
            const int q = (sl + 1) >> 1;
            const T* w = (sl == 0 || (sl & 1)) ? twr_r : twr_i;
            const T* row = Ar + ii * ld;
            T acc = T(0);
            int t = 0;
            for (int j = 0; j < L; ++j) {
                acc += row[pidx(j)] * w[t];
                t += q;
                if (t >= L) t -= L;
            }
            outK[js * (size_t)sl + gi] = acc;
        }
        return;
    }
    if (H > 1) cfft_lds<T>(Ar, Ai, ld, TI, a.fp, twc_r, twc_i, T(-1), tid, nthr);
    const T* Zr = Ar;
    const T* Zi = Ai;
    // X_k = (Z_k + conj Z_{H-k})/2 - i W^k (Z_k - conj Z_{H-k})/2, W = exp(-2 pi i/L). The
    // thread's pencil is fixed (blockDim is a multiple of TI); its coefficients k step by
    // blockDim/TI. S addresses as 32-bit byte offsets from the tile's column (S < 4 GiB).
    {
        const int ii = tid & (TI - 1);
        const int gi = i0 + ii;
        if (gi < d.M) {
            const T* zr = Zr + ii * ld;
            const T* zi = Zi + ii * ld;
            T* out = outK + gi;
            const unsigned jsb = (unsigned)(js * sizeof(T));
            const int kstep = nthr >> lti;
            for (int kk = tid >> lti; kk <= H; kk += kstep) {
                const int k1 = (kk == H) ? 0 : kk;
                const int k2 = (kk == 0) ? 0 : H - kk;
                const T ar = zr[pidx(k1)], ai = zi[pidx(k1)];
                const T br = zr[pidx(k2)], bi = -zi[pidx(k2)];
                const T er = T(0.5) * (ar + br), ei = T(0.5) * (ai + bi);
                const T dr = T(0.5) * (ar - br), di = T(0.5) * (ai - bi);
                const T wr = twr_r[kk], wi = twr_i[kk];
                const T wdr = wr * dr - wi * di, wdi = wr * di + wi * dr;
                const T xr = er + wdi, xi = ei - wdr;
                if (kk == 0) {
                    out[0] = xr;
                } else if (kk == H) {
                    st_off(out, jsb * (unsigned)(L - 1), xr);
                } else {
                    st_off(out, jsb * (unsigned)(2 * kk - 1), xr);
                    st_off(out, jsb * (unsigned)(2 * kk), xi);
                }
            }
        }
    }
}

template <typename T, int MODE, bool DENSE>
__global__ void __launch_bounds__(kFftThreads) k_yfft_fwd(YArgs<T> a) {
    const int nb = gridDim.x * gridDim.y;
    const int t = iblock_tile(xcd_tile(blockIdx.x + gridDim.x * blockIdx.y, nb), gridDim.x, gridDim.y, a.iblk);
    yfft_fwd_tile<T, MODE, DENSE>(a, t, gridDim.x, linear_bid(), nb);
}

// The forward transform of a stored RHS at no more than 80 VGPRs: with its 42 KB LDS tile
// (L = 256, 16 pencils) three workgroups of 8 waves then share a CU (6 waves per SIMD).
template <typename T, bool DENSE>
__global__ void __launch_bounds__(kFftThreads) __attribute__((amdgpu_waves_per_eu(6))) k_yfft_fwd_rhs(YArgs<T> a) {
    const int nb = gridDim.x * gridDim.y;
    const int t = iblock_tile(xcd_tile(blockIdx.x + gridDim.x * blockIdx.y, nb), gridDim.x, gridDim.y, a.iblk);
    yfft_fwd_tile<T, YF_DIRECT, DENSE>(a, t, gridDim.x, linear_bid(), nb);
}

template <typename T, int MODE, bool DENSE = false>
__device__ __forceinline__ void yfft_inv_tile(const YArgs<T>& a, int tile, int ntx, int slot) {
    extern __shared__ unsigned char smem_raw[];
    T* smem = reinterpret_cast<T*>(smem_raw);
    const Dims d = a.d;
    const int TI = a.TI, lti = a.lti, ld = a.ld, H = d.L / 2, L = d.L;
    T* Ar = smem;  // FFT: TI complex rows, transformed in place; DENSE: TI real rows
    T* Ai = Ar + TI * ld;
    T* twc_r = Ai + TI * ld;
    T* twc_i = twc_r + H;
    T* twr_r = DENSE ? Ar + TI * ld : twc_i + H;
    T* twr_i = twr_r + (DENSE ? L : H + 1);
    const int i0 = (tile % ntx) * TI;
    const int k = tile / ntx;
    const int tid = threadIdx.x, nthr = blockDim.x;
    if (DENSE) {
        for (int t = tid; t < L; t += nthr) {
            twr_r[t] = a.tb.ytwr_re[t];
            twr_i[t] = a.tb.ytwr_im[t];
        }
    } else {
        for (int t = tid; t <= H; t += nthr) {
            if (t < H) {
                twc_r[t] = a.tb.ytwc_re[t];
                twc_i[t] = a.tb.ytwc_im[t];
            }
            twr_r[t] = a.tb.ytwr_re[t];
            twr_i[t] = a.tb.ytwr_im[t];
        }
    }
    const size_t js = (size_t)d.M * (size_t)d.N;
    const T* inK = a.Sin + (size_t)d.M * (size_t)k;
    __syncthreads();  // the twiddle tables are read below by other threads than wrote them
    if (DENSE) {
        // the packed coefficients of the tile's pencils into LDS rows, scaled by 1/L (rows
        // of contiguous x1 values: coalesced); pencils past M read a clamped in-range row
        const T invL = T(1) / T(L);
        const int ii = tid & (TI - 1);
        const bool iok = i0 + ii < d.M;
        const T* col = inK + (iok ? i0 + ii : d.M - 1);
        const int sstep = nthr >> lti;
        for (int s0 = tid >> lti; s0 < L; s0 += sstep * kLoadBatch) {
            T x[kLoadBatch];
#pragma unroll
            for (int u = 0; u < kLoadBatch; ++u) x[u] = col[js * (size_t)min(s0 + u * sstep, L - 1)];
#pragma unroll
            for (int u = 0; u < kLoadBatch; ++u)
                if (s0 + u * sstep < L) Ar[ii * ld + pidx(s0 + u * sstep)] = iok ? x[u] * invL : T(0);
        }
        __syncthreads();
    }

    // Packed spectrum -> Z'_q = [(X_q + conj X_{H-q}) + i conj(W^q) (X_q - conj X_{H-q})] / L,
    // q < H. Task (ii, q) for q <= H/2 reads the coefficient rows of X_q and X_{H-q} straight
    // from S (X_0 = row 0, X_H = row L-1, X_p = rows 2p-1, 2p otherwise) and forms Z'_q and
    // Z'_{H-q}: every row is read once and nothing is staged.
    const T invL = T(1) / T(L);
    const int half = DENSE ? -1 : H / 2;  // DENSE: no packing tasks
    constexpr int UB = kInvTasks;
    // The thread's x1 offset is fixed (blockDim is a multiple of TI). The four coefficient
    // rows of a task are loaded before any arithmetic (arithmetic inside a load's own guard
    // makes the compiler wait for it before the next load): rows 2q-1, 2q, 2(H-q)-1, 2(H-q),
    // with q = 0 reading rows 0 and L-1 (X_0, X_H real). Tasks past H/2 and pencils past M
    // read clamped in-range rows whose values are zeroed afterwards.
    const int ii = tid & (TI - 1);
    const bool iok = i0 + ii < d.M;
    const T* col = inK + (iok ? i0 + ii : d.M - 1);
    const int qstep = nthr >> lti;
    // Row addresses as 32-bit byte offsets from the column (S < 4 GiB), stepped with the task
    // index instead of multiplied: with rs the bytes between coefficient rows and
    // qo = 2 q rs, rows 2q-1, 2q, 2(H-q)-1 and 2(H-q) are qo - rs, qo, (L-1) rs - qo and
    // L rs - qo (q = 0: rows 0, 1, L-1 and L-1 again). Tasks past H/2 take task H/2's rows.
    const unsigned rs = (unsigned)sizeof(T) * (unsigned)d.M * (unsigned)d.N;
    const unsigned lm1 = (unsigned)(L - 1) * rs;
    const unsigned q2 = 2u * rs, qcap = (unsigned)half * q2;
    unsigned qo0 = (unsigned)(tid >> lti) * q2;
    for (int q0 = tid >> lti; q0 <= half; q0 += qstep * UB, qo0 += (unsigned)(qstep * UB) * q2) {
        T xa_r[UB], xa_i[UB], xb_r[UB], xb_i[UB];
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const unsigned qo = (q0 + u * qstep <= half) ? qo0 + (unsigned)(u * qstep) * q2 : qcap;
            const bool q0row = qo == 0;
            const unsigned ob = lm1 - qo;
            xa_r[u] = ld_off(col, q0row ? 0u : qo - rs);
            xa_i[u] = ld_off(col, q0row ? rs : qo);
            xb_r[u] = ld_off(col, ob);
            xb_i[u] = ld_off(col, q0row ? lm1 : ob + rs);
        }
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const int q = q0 + u * qstep;
            if (q > half) continue;
            const bool im = iok && q > 0;
            xa_r[u] = iok ? xa_r[u] : T(0);
            xb_r[u] = iok ? xb_r[u] : T(0);
            xa_i[u] = im ? xa_i[u] : T(0);
            xb_i[u] = im ? xb_i[u] : T(0);
            const int qb = H - q;
            T* rr = Ar + ii * ld;
            T* ri = Ai + ii * ld;
            {
                const T yr = xb_r[u], yi = -xb_i[u];
                const T sr = xa_r[u] + yr, si = xa_i[u] + yi;
                const T dr = xa_r[u] - yr, di = xa_i[u] - yi;
                const T wr = twr_r[q], wi = -twr_i[q];
                const T vr = wr * dr - wi * di, vi = wr * di + wi * dr;
                rr[pidx(q)] = (sr - vi) * invL;
                ri[pidx(q)] = (si + vr) * invL;
            }
            if (qb < H && qb != q) {
                const T yr = xa_r[u], yi = -xa_i[u];
                const T sr = xb_r[u] + yr, si = xb_i[u] + yi;
                const T dr = xb_r[u] - yr, di = xb_i[u] - yi;
                const T wr = twr_r[qb], wi = -twr_i[qb];
                const T vr = wr * dr - wi * di, vi = wr * di + wi * dr;
                rr[pidx(qb)] = (sr - vi) * invL;
                ri[pidx(qb)] = (si + vr) * invL;
            }
        }
    }
    if (!DENSE) {
        __syncthreads();
        if (H > 1) cfft_lds<T>(Ar, Ai, ld, TI, a.fp, twc_r, twc_i, T(1), tid, nthr);
    }
    const T* Zr = Ar;
    const T* Zi = Ai;
    double psum = 0.0;
    const int K = k + 1;
    const double mean =
        MODE == YI_UPDATE_FINAL ? a.red[RED_PHISUM] / ((double)d.M * (double)d.L * (double)d.N) : 0.0;
    const int total = TI * L;
    // DP addresses as 32-bit byte offsets from the field base (DP < 4 GiB)
    const unsigned nx8 = 8u * (unsigned)d.nx, pl8K = nx8 * (unsigned)d.ny * (unsigned)K;
    for (int e = tid; e < total; e += nthr) {
        const int ii = e & (TI - 1);
        const int j = e >> lti;
        const int gi = i0 + ii;
        if (gi >= d.M) continue;
        const unsigned oj = pl8K + nx8 * (unsigned)(j + 1) + 8u * (unsigned)(gi + 1);
        double v;
        if (DENSE) {
            // x_j = X_0 + 2 sum_q (Re X_q cos(2 pi q j/L) - Im X_q sin(2 pi q j/L)) + X_{L/2} (-1)^j,
            // the coefficients already scaled by 1/L (fftpack's RFFTB after direct.f:1036)
            const T* row = Ar + ii * ld;
            T acc = T(0);
            int t = j;
            for (int q = 1; 2 * q < L; ++q) {
                acc += row[pidx(2 * q - 1)] * twr_r[t] + row[pidx(2 * q)] * twr_i[t];
                t += j;
                if (t >= L) t -= L;
            }
            T x = row[pidx(0)] + T(2) * acc;
            if ((L & 1) == 0) x += (j & 1) ? -row[pidx(L - 1)] : row[pidx(L - 1)];
            v = (double)x;
        } else {
            v = (double)((j & 1) ? Zi[ii * ld + pidx((j >> 1))] : Zr[ii * ld + pidx((j >> 1))]);
        }
        const int I = gi + 1, J = j + 1;
        if (MODE == YI_UPDATE_FINAL) v = ld_off(a.dp, oj) + v - mean;
        if (MODE == YI_DIRECT || MODE == YI_UPDATE_FINAL) {
            // interior plus the x ghosts (MP=1: DP(1)=DP(2), DP(NX)=DP(NX-1); MP=0: DP(1)=DP(NX-1),
            // DP(NX)=DP(2)) and the y ghosts (LP=0: DP(:,1)=DP(:,NY-1), DP(:,NY)=DP(:,2); LP=1:
            // DP(:,1)=DP(:,2), DP(:,NY)=DP(:,NY-1)), direct.f:1090-1106
            st_off(a.dp, oj, v);
            const int Ilo = (gi == (d.cyc_x ? d.M - 1 : 0)) ? 0 : -1;
            const int Ihi = (gi == (d.cyc_x ? 0 : d.M - 1)) ? d.nx - 1 : -1;
            const int Jg = d.ycopy ? ((j == 0) ? 0 : ((j == L - 1) ? d.ny - 1 : -1))
                                   : ((j == L - 1) ? 0 : ((j == 0) ? d.ny - 1 : -1));
            if (Ilo >= 0) a.dp[dpidx(d, Ilo, J, K)] = v;
            if (Ihi >= 0) a.dp[dpidx(d, Ihi, J, K)] = v;
            if (Jg >= 0) {
                a.dp[dpidx(d, I, Jg, K)] = v;
                if (Ilo >= 0) a.dp[dpidx(d, Ilo, Jg, K)] = v;
                if (Ihi >= 0) a.dp[dpidx(d, Ihi, Jg, K)] = v;
            }
        } else if (MODE == YI_STORE) {
            st_off(a.dp, oj, v);
            psum += v;
        } else {
            const double nv = ld_off(a.dp, oj) + v;
            st_off(a.dp, oj, nv);
            psum += nv;
        }
    }
    if (MODE == YI_DIRECT || MODE == YI_UPDATE_FINAL) {
        // z ghost planes (K=1, K=NZ) are zero (direct.f:1035); the blocks at the two ends
        // of x2 clear them for their x1 range (with the x ghosts at the domain edges).
        if (k == 0 || k == d.N - 1) {
            const int Kg = (k == 0) ? 0 : d.nz - 1;
            const int ilo = (i0 == 0) ? 0 : i0 + 1;
            const int ihi = (i0 + TI >= d.M) ? d.nx - 1 : i0 + TI;
            const int w = ihi - ilo + 1;
            for (int e = tid; e < w * d.ny; e += nthr) {
                const int I = ilo + e % w;
                const int J = e / w;
                a.dp[dpidx(d, I, J, Kg)] = 0.0;
            }
        }
    } else {
        block_partial<RED_OP_SUM>(psum, a.part, slot);
    }
}

template <typename T, int MODE, bool DENSE>
// at most 80 VGPRs (the last defect-correction update needs 82 unbounded): with its 42 KB LDS
// tile (L = 256, 16 pencils) three workgroups of 8 waves share a CU (6 waves per SIMD)
__global__ void __launch_bounds__(kFftThreads) __attribute__((amdgpu_waves_per_eu(6))) k_yfft_inv(YArgs<T> a) {
    yfft_inv_tile<T, MODE, DENSE>(a, xcd_tile(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y), gridDim.x,
                                  linear_bid());
}

// TI complex rows (transformed in place) + the two twiddle tables
static inline size_t ylds_bytes(int TI, int ld, size_t tsz) { return (2 * (size_t)TI * ld + 4 * (size_t)ld) * tsz; }
// the dense DFT: TI real rows + the table of L complex roots
static inline size_t ylds_dense_bytes(int TI, int L, size_t tsz) { return ((size_t)TI * fft_ld(L) + 2 * (size_t)L) * tsz; }

static inline int ilog2(int v) {
    int l = 0;
    while ((1 << l) < v) ++l;
    return l;
}

template <typename T>
void launch_yfft_fwd(hipStream_t st, const Dims& d, const FftPlan& fp, const Tables<T>& tb, const FwdInput& in,
                     const Op64& op, const Reduce& rd, T* S, int TI, int iblk) {
    if constexpr (std::is_same<T, float>::value) {
        if (tb.ydft_f && (in.mode == YF_DIRECT || in.mode == YF_RESIDUAL)) {
            launch_ygemm_fwd(st, d, tb.ydft_f, in, op, rd, S);
            return;
        }
    }
    const double* src = in.src;
    const double* phi = in.phi;
    const int shift = in.shift;
    const int mode = in.mode;
    YArgs<T> a{};
    a.src = src;
    a.phi = phi;
    a.rhs_out = in.rhs_out;
    a.op = op;
    a.part = rd.part;
    a.red = rd.red;
    a.S = S;
    a.d = d;
    a.fp = fp;
    a.tb = tb;
    a.TI = TI;
    a.lti = ilog2(TI);
    a.iblk = iblk;
    const bool dense = fp.n == 0;  // the plan's dense-DFT marker (no radix plan for L)
    a.ld = dense ? fft_ld(d.L) : fft_ld(d.L / 2);
    a.shift = shift;
    dim3 grid(ceil_div(d.M, TI), d.N);
    size_t sh = dense ? ylds_dense_bytes(TI, d.L, sizeof(T)) : ylds_bytes(TI, a.ld, sizeof(T));
    auto go = [&](auto fn) {
        (void)hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
        hipLaunchKernelGGL(fn, grid, dim3(kFftThreads), sh, st, a);
    };
    if (mode == YF_DIRECT) {
        if (dense) go(k_yfft_fwd<T, YF_DIRECT, true>);
        else go(k_yfft_fwd_rhs<T, false>);
    } else if (mode == YF_DIVERGENCE) {
        a.us = in.us;
        a.vs = in.vs;
        a.ws = in.ws;
        a.mt = in.mt;
        a.scale = in.scale;
        a.dscale = in.dscale;
        if (dense) {
            if (in.mt.cart) go(k_yfft_fwd<T, YF_DIV_CART, true>);
            else go(k_yfft_fwd<T, YF_DIV_CYL, true>);
        } else {
            if (in.mt.cart) go(k_yfft_fwd<T, YF_DIV_CART, false>);
            else go(k_yfft_fwd<T, YF_DIV_CYL, false>);
        }
        const int nb = (int)(grid.x * grid.y);
        launch_fold(st, rd.part, nb, rd.red + RED_VOLB, RED_OP_SUM, 0);
        launch_fill(st, rd.red + RED_VOL, 1, in.mt.vol_total);
    } else {
        if (dense) go(k_yfft_fwd<T, YF_RESIDUAL, true>);
        else go(k_yfft_fwd<T, YF_RESIDUAL, false>);
        launch_fold(st, rd.part, (int)(grid.x * grid.y), rd.red + RED_RNORM2, RED_OP_SUM, 0);
    }
}

template <typename T>
void launch_yfft_inv(hipStream_t st, const Dims& d, const FftPlan& fp, const Tables<T>& tb,
                     const T* S, double* dp, const Reduce& rd, int TI, int mode) {
    if constexpr (std::is_same<T, float>::value) {
        if (tb.ydft_b && (mode == YI_STORE || mode == YI_UPDATE)) {
            launch_ygemm_inv(st, d, tb.ydft_b, S, dp, rd, mode);
            return;
        }
    }
    YArgs<T> a{};
    a.Sin = S;
    a.dp = dp;
    a.part = rd.part;
    a.red = rd.red;
    a.d = d;
    a.fp = fp;
    a.tb = tb;
    a.TI = TI;
    a.lti = ilog2(TI);
    const bool dense = fp.n == 0;
    a.ld = dense ? fft_ld(d.L) : fft_ld(d.L / 2);
    dim3 grid(ceil_div(d.M, TI), d.N);
    size_t sh = dense ? ylds_dense_bytes(TI, d.L, sizeof(T)) : ylds_bytes(TI, a.ld, sizeof(T));
    auto go = [&](auto fn) {
        (void)hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
        hipLaunchKernelGGL(fn, grid, dim3(kFftThreads), sh, st, a);
    };
    switch (mode) {
        case YI_DIRECT:
            if (dense) go(k_yfft_inv<T, YI_DIRECT, true>);
            else go(k_yfft_inv<T, YI_DIRECT, false>);
            break;
        case YI_STORE:
            if (dense) go(k_yfft_inv<T, YI_STORE, true>);
            else go(k_yfft_inv<T, YI_STORE, false>);
            launch_fold(st, rd.part, (int)(grid.x * grid.y), rd.red + RED_PHISUM, RED_OP_SUM, 0);
            break;
        case YI_UPDATE_FINAL:
            if (dense) go(k_yfft_inv<T, YI_UPDATE_FINAL, true>);
            else go(k_yfft_inv<T, YI_UPDATE_FINAL, false>);
            break;
        default:
            if (dense) go(k_yfft_inv<T, YI_UPDATE, true>);
            else go(k_yfft_inv<T, YI_UPDATE, false>);
            launch_fold(st, rd.part, (int)(grid.x * grid.y), rd.red + RED_PHISUM, RED_OP_SUM, 0);
            break;
    }
}

// ------------------------------------------------------------------ batched (one item per workgroup)
// The residual form at no more than 128 VGPRs (129 unbounded): two workgroups of 8 waves per
// CU instead of one, as its plan kernel runs.
template <typename T, int MODE>
__global__ void __launch_bounds__(kFftThreads) __attribute__((amdgpu_waves_per_eu(MODE == YF_RESIDUAL ? 4 : 1)))
k_batch_yfft_fwd(const Job<YArgs<T>>* jobs, int njobs, int total) {
    run_item(jobs, njobs, total,
             [](const Job<YArgs<T>>& J, int t) { yfft_fwd_tile<T, MODE>(J.a, t, J.ntx, t, J.ntiles); });
}

// the stored-RHS form at no more than 80 VGPRs, as k_yfft_fwd_rhs
template <typename T>
__global__ void __launch_bounds__(kFftThreads) __attribute__((amdgpu_waves_per_eu(6)))
k_batch_yfft_fwd_rhs(const Job<YArgs<T>>* jobs, int njobs, int total) {
    run_item(jobs, njobs, total,
             [](const Job<YArgs<T>>& J, int t) { yfft_fwd_tile<T, YF_DIRECT>(J.a, t, J.ntx, t, J.ntiles); });
}

template <typename T, int MODE>
__global__ void __launch_bounds__(kFftThreads) k_batch_yfft_inv(const Job<YArgs<T>>* jobs, int njobs, int total) {
    run_item(jobs, njobs, total, [](const Job<YArgs<T>>& J, int t) { yfft_inv_tile<T, MODE>(J.a, t, J.ntx, t); });
}

template <typename T>
static YArgs<T> yargs_of(const BatchProblem<T>& p, int fwd, int mode) {
    YArgs<T> a{};
    a.d = p.d;
    a.fp = p.fpy;
    a.tb = p.tb;
    a.TI = p.tiy;
    a.lti = ilog2(p.tiy);
    a.ld = fft_ld(p.d.L / 2);
    a.part = p.part;
    a.red = p.red;
    a.op = p.op;
    if (fwd) {
        a.S = p.S;
        if (mode == YF_DIRECT) {
            a.src = p.rhs ? p.rhs : p.dp;
        } else if (mode == YF_RESIDUAL) {
            a.src = p.rhs;
            a.phi = p.dp;
        } else {
            a.us = p.us;
            a.vs = p.vs;
            a.ws = p.ws;
            a.mt = p.mt;
            a.dscale = p.dscale;
        }
    } else {
        a.Sin = p.S;
        a.dp = p.dp;
    }
    return a;
}


template <typename T>
int prepare_yfft(const std::vector<BatchProblem<T>>& pr, int fwd, int mode, StageLaunch& out,
                 std::vector<void*>& owned) {
    auto lds_of = [&](int x) { return ylds_bytes(pr[x].tiy, fft_ld(pr[x].d.L / 2), sizeof(T)); };
    std::vector<int> order(pr.size());
    std::iota(order.begin(), order.end(), 0);
    // the tiles of most LDS (fewest workgroups per CU) first, then the longest
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        const int ox = lds_occupancy(lds_of(x)), oy = lds_occupancy(lds_of(y));
        return ox != oy ? ox < oy : pr[x].tiy * pr[x].d.L > pr[y].tiy * pr[y].d.L;
    });
    std::vector<Job<YArgs<T>>> jobs;
    out = StageLaunch{};
    for (int idx : order) {
        const BatchProblem<T>& p = pr[idx];
        Job<YArgs<T>> j{};
        j.a = yargs_of(p, fwd, mode);
        j.ntx = ceil_div(p.d.M, p.tiy);
        j.ntiles = j.ntx * p.d.N;
        if (!add_class_job(out, jobs, j, 0, lds_of(idx))) return -1;
    }
    if (upload_jobs(jobs, out, owned)) return -1;
    if (fwd && mode == YF_DIVERGENCE && !pr.empty()) {
        bool all_cart = true, all_cyl = true;
        for (const BatchProblem<T>& p : pr) {
            all_cart = all_cart && p.mt.cart;
            all_cyl = all_cyl && !p.mt.cart;
        }
        out.kind = all_cart ? 1 : (all_cyl ? 2 : 0);
    }
    return 0;
}

template <typename T>
void batch_yfft(hipStream_t st, const StageLaunch& sl, int fwd, int mode) {
    auto go = [&](auto fn) {
        (void)hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sl.lds);
        for (int c = 0; c < sl.ncls; ++c) {
            const Job<YArgs<T>>* jobs = (const Job<YArgs<T>>*)sl.jobs + sl.cls_job[c];
            hipLaunchKernelGGL(fn, dim3(std::max(sl.cls_total[c], 1)), dim3(kFftThreads), sl.cls_lds[c], st, jobs,
                               sl.cls_job[c + 1] - sl.cls_job[c], sl.cls_total[c]);
        }
    };
    if (fwd) {
        if (mode == YF_DIRECT) go(k_batch_yfft_fwd_rhs<T>);
        else if (mode == YF_RESIDUAL) go(k_batch_yfft_fwd<T, YF_RESIDUAL>);
        else if (sl.kind == 1) go(k_batch_yfft_fwd<T, YF_DIV_CART>);
        else if (sl.kind == 2) go(k_batch_yfft_fwd<T, YF_DIV_CYL>);
        else go(k_batch_yfft_fwd<T, YF_DIVERGENCE>);
    } else {
        if (mode == YI_DIRECT) go(k_batch_yfft_inv<T, YI_DIRECT>);
        else if (mode == YI_STORE) go(k_batch_yfft_inv<T, YI_STORE>);
        else go(k_batch_yfft_inv<T, YI_UPDATE>);
    }
}

template int prepare_yfft<double>(const std::vector<BatchProblem<double>>&, int, int, StageLaunch&,
                                  std::vector<void*>&);
template int prepare_yfft<float>(const std::vector<BatchProblem<float>>&, int, int, StageLaunch&,
                                 std::vector<void*>&);
template void batch_yfft<double>(hipStream_t, const StageLaunch&, int, int);
template void batch_yfft<float>(hipStream_t, const StageLaunch&, int, int);

template void launch_yfft_fwd<double>(hipStream_t, const Dims&, const FftPlan&, const Tables<double>&,
                                      const FwdInput&, const Op64&, const Reduce&, double*, int, int);
template void launch_yfft_fwd<float>(hipStream_t, const Dims&, const FftPlan&, const Tables<float>&,
                                     const FwdInput&, const Op64&, const Reduce&, float*, int, int);
template void launch_yfft_inv<double>(hipStream_t, const Dims&, const FftPlan&, const Tables<double>&,
                                      const double*, double*, const Reduce&, int, int);
template void launch_yfft_inv<float>(hipStream_t, const Dims&, const FftPlan&, const Tables<float>&,
                                     const float*, double*, const Reduce&, int, int);

}  // namespace eddy
