// yfft.hip -- batched real FFTs along the periodic axis (the reference's Y, 2nd index).
//
// Forward: direct.f:620-629 (RFFTF of every interior pencil DP(I+1,2:L+1,K+1));
// inverse: direct.f:1035-1106 (DUMMY/L, RFFTB, ghost fill). Output packing follows
// fftpack (fftpack.f:140-240): [Re X0, Re X1, Im X1, ..., Re X_{L/2}] (L even).
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

template <typename T>
struct YArgs {
    const double* src;
    const double* phi;
    const double *us, *vs, *ws;  // YF_DIVERGENCE inputs
    Metrics mt;
    double scale;
    Op64 op;
    double* part;
    double* red;
    T* S;
    const T* Sin;
    double* dp;
    Dims d;
    FftPlan fp;
    Tables<T> tb;
    int TI, lti;  // pencils per workgroup and its log2
    int ld;
    int shift;
};

__device__ __forceinline__ size_t dpidx(const Dims& d, int I, int J, int K) { return sidx(d, I, J, K); }

// One tile (TI pencils at one x2 index) of the forward transform; `tile` is the logical tile,
// ntx the tiles per x2 row, `slot`/`nslots` the reduction partial index and count.
template <typename T, int MODE>
__device__ void yfft_fwd_tile(const YArgs<T>& a, int tile, int ntx, int slot, int nslots) {
    extern __shared__ unsigned char smem_raw[];
    T* smem = reinterpret_cast<T*>(smem_raw);
    const Dims d = a.d;
    const int TI = a.TI, lti = a.lti, ld = a.ld, H = d.L / 2, L = d.L;
    T* Ar = smem;  // TI complex rows, transformed in place
    T* Ai = Ar + TI * ld;
    T* twc_r = Ai + TI * ld;  // exp(-2 pi i t/H), t < H
    T* twc_i = twc_r + H;
    T* twr_r = twc_i + H;     // exp(-2 pi i k/L), k <= H
    T* twr_i = twr_r + H + 1;
    const int i0 = (tile % ntx) * TI;
    const int k = tile / ntx;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int total = TI * L;
    for (int t = tid; t <= H; t += nthr) {
        if (t < H) {
            twc_r[t] = a.tb.ytwc_re[t];
            twc_i[t] = a.tb.ytwc_im[t];
        }
        twr_r[t] = a.tb.ytwr_re[t];
        twr_i[t] = a.tb.ytwr_im[t];
    }

    double shift = 0.0;
    if (MODE == YF_DIRECT && a.shift) shift = a.red[RED_VOLB] / a.red[RED_VOL];
    double rsq = 0.0;
    const double* srcK = a.src + (size_t)d.nx * (size_t)d.ny * (size_t)(k + 1) + 1;
    for (int e0 = tid; e0 < total; e0 += nthr * kLoadBatch) {
        double v[kLoadBatch];
#pragma unroll
        for (int u = 0; u < kLoadBatch; ++u) {
            const int e = e0 + u * nthr;
            const int ii = e & (TI - 1);
            const int j = e >> lti;
            const int gi = i0 + ii;
            v[u] = 0.0;
            if (e < total && gi < d.M) {
                if (MODE == YF_DIRECT) {
                    v[u] = srcK[(size_t)gi + (size_t)d.nx * (size_t)(j + 1)] - shift;
                } else if (MODE == YF_RESIDUAL) {
                    v[u] = a.src[sidx(d, gi + 1, j + 1, k + 1)] - lap_at(d, a.op, a.phi, gi, j, k);
                    rsq += v[u] * v[u];
                } else {
                    v[u] = div_at(d, a.mt, a.us, a.vs, a.ws, gi, j, k) * a.scale;
                    rsq += cell_vol(a.mt, gi, j, k) * v[u];  // sum vol*b (solvability)
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kLoadBatch; ++u) {
            const int e = e0 + u * nthr;
            if (e < total) {
                const int ii = e & (TI - 1);
                const int j = e >> lti;
                if (j & 1) Ai[ii * ld + pidx((j >> 1))] = T(v[u]);
                else Ar[ii * ld + pidx((j >> 1))] = T(v[u]);
            }
        }
    }
    if (MODE == YF_RESIDUAL) block_partial<RED_OP_SUM>(rsq, a.part, slot);
    if (MODE == YF_DIVERGENCE) block_partial<RED_OP_SUM>(rsq, a.part, slot);
    __syncthreads();
    if (H > 1) cfft_lds_ip<T>(Ar, Ai, ld, TI, a.fp, twc_r, twc_i, T(-1), tid, nthr);
    const T* Zr = Ar;
    const T* Zi = Ai;
    // X_k = (Z_k + conj Z_{H-k})/2 - i W^k (Z_k - conj Z_{H-k})/2, W = exp(-2 pi i/L)
    const size_t js = (size_t)d.M * (size_t)d.N;
    T* outK = a.S + (size_t)d.M * (size_t)k;
    for (int e = tid; e < TI * (H + 1); e += nthr) {
        const int ii = e & (TI - 1);
        const int kk = e >> lti;
        const int gi = i0 + ii;
        if (gi >= d.M) continue;
        const int k1 = (kk == H) ? 0 : kk;
        const int k2 = (kk == 0) ? 0 : H - kk;
        const T ar = Zr[ii * ld + pidx(k1)], ai = Zi[ii * ld + pidx(k1)];
        const T br = Zr[ii * ld + pidx(k2)], bi = -Zi[ii * ld + pidx(k2)];
        const T er = T(0.5) * (ar + br), ei = T(0.5) * (ai + bi);
        const T dr = T(0.5) * (ar - br), di = T(0.5) * (ai - bi);
        const T wr = twr_r[kk], wi = twr_i[kk];
        const T wdr = wr * dr - wi * di, wdi = wr * di + wi * dr;
        const T xr = er + wdi, xi = ei - wdr;
        T* out = outK + gi;
        if (kk == 0) {
            out[0] = xr;
        } else if (kk == H) {
            out[js * (size_t)(L - 1)] = xr;
        } else {
            out[js * (size_t)(2 * kk - 1)] = xr;
            out[js * (size_t)(2 * kk)] = xi;
        }
    }
}

template <typename T, int MODE>
__global__ void __launch_bounds__(kFftThreads) k_yfft_fwd(YArgs<T> a) {
    const int nb = gridDim.x * gridDim.y;
    yfft_fwd_tile<T, MODE>(a, xcd_tile(blockIdx.x + gridDim.x * blockIdx.y, nb), gridDim.x, linear_bid(), nb);
}

template <typename T, int MODE>
__device__ void yfft_inv_tile(const YArgs<T>& a, int tile, int ntx, int slot) {
    extern __shared__ unsigned char smem_raw[];
    T* smem = reinterpret_cast<T*>(smem_raw);
    const Dims d = a.d;
    const int TI = a.TI, lti = a.lti, ld = a.ld, H = d.L / 2, L = d.L;
    T* Ar = smem;  // TI complex rows, transformed in place
    T* Ai = Ar + TI * ld;
    T* twc_r = Ai + TI * ld;
    T* twc_i = twc_r + H;
    T* twr_r = twc_i + H;
    T* twr_i = twr_r + H + 1;
    const int i0 = (tile % ntx) * TI;
    const int k = tile / ntx;
    const int tid = threadIdx.x, nthr = blockDim.x;
    for (int t = tid; t <= H; t += nthr) {
        if (t < H) {
            twc_r[t] = a.tb.ytwc_re[t];
            twc_i[t] = a.tb.ytwc_im[t];
        }
        twr_r[t] = a.tb.ytwr_re[t];
        twr_i[t] = a.tb.ytwr_im[t];
    }
    const size_t js = (size_t)d.M * (size_t)d.N;
    const int total = TI * L;
    const T* inK = a.Sin + (size_t)d.M * (size_t)k;
    __syncthreads();  // the twiddle tables are read below by other threads than wrote them

    // Packed spectrum -> Z'_q = [(X_q + conj X_{H-q}) + i conj(W^q) (X_q - conj X_{H-q})] / L,
    // q < H. Task (ii, q) for q <= H/2 reads the coefficient rows of X_q and X_{H-q} straight
    // from S (X_0 = row 0, X_H = row L-1, X_p = rows 2p-1, 2p otherwise) and forms Z'_q and
    // Z'_{H-q}: every row is read once and nothing is staged.
    const T invL = T(1) / T(L);
    const int half = H / 2;
    const int ntask = TI * (half + 1);
    constexpr int UB = 2;
    for (int e0 = tid; e0 < ntask; e0 += nthr * UB) {
        T xa_r[UB], xa_i[UB], xb_r[UB], xb_i[UB];
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const int e = e0 + u * nthr;
            const int ii = e & (TI - 1);
            const int q = e >> lti;
            const int gi = i0 + ii;
            xa_r[u] = xa_i[u] = xb_r[u] = xb_i[u] = T(0);
            if (e < ntask && gi < d.M) {
                const T* col = inK + gi;
                const int qb = H - q;
                if (q == 0) {
                    xa_r[u] = col[0];
                } else {
                    xa_r[u] = col[js * (size_t)(2 * q - 1)];
                    xa_i[u] = col[js * (size_t)(2 * q)];
                }
                if (qb == H) {
                    xb_r[u] = col[js * (size_t)(L - 1)];
                } else {
                    xb_r[u] = col[js * (size_t)(2 * qb - 1)];
                    xb_i[u] = col[js * (size_t)(2 * qb)];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const int e = e0 + u * nthr;
            if (e >= ntask) continue;
            const int ii = e & (TI - 1);
            const int q = e >> lti;
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
    __syncthreads();
    if (H > 1) cfft_lds_ip<T>(Ar, Ai, ld, TI, a.fp, twc_r, twc_i, T(1), tid, nthr);
    const T* Zr = Ar;
    const T* Zi = Ai;
    double psum = 0.0;
    const int K = k + 1;
    for (int e = tid; e < total; e += nthr) {
        const int ii = e & (TI - 1);
        const int j = e >> lti;
        const int gi = i0 + ii;
        if (gi >= d.M) continue;
        const double v = (double)((j & 1) ? Zi[ii * ld + pidx((j >> 1))] : Zr[ii * ld + pidx((j >> 1))]);
        const int I = gi + 1, J = j + 1;
        if (MODE == YI_DIRECT) {
            // interior plus the x ghosts (MP=1: DP(1)=DP(2), DP(NX)=DP(NX-1)) and the periodic
            // y ghosts (LP=0: DP(:,1)=DP(:,NY-1), DP(:,NY)=DP(:,2)), direct.f:1090-1106
            a.dp[dpidx(d, I, J, K)] = v;
            const int Ilo = (gi == 0) ? 0 : -1, Ihi = (gi == d.M - 1) ? d.nx - 1 : -1;
            const int Jg = (j == L - 1) ? 0 : ((j == 0) ? d.ny - 1 : -1);
            if (Ilo >= 0) a.dp[dpidx(d, Ilo, J, K)] = v;
            if (Ihi >= 0) a.dp[dpidx(d, Ihi, J, K)] = v;
            if (Jg >= 0) {
                a.dp[dpidx(d, I, Jg, K)] = v;
                if (Ilo >= 0) a.dp[dpidx(d, Ilo, Jg, K)] = v;
                if (Ihi >= 0) a.dp[dpidx(d, Ihi, Jg, K)] = v;
            }
        } else if (MODE == YI_STORE) {
            a.dp[dpidx(d, I, J, K)] = v;
            psum += v;
        } else {
            const size_t o = dpidx(d, I, J, K);
            const double nv = a.dp[o] + v;
            a.dp[o] = nv;
            psum += nv;
        }
    }
    if (MODE == YI_DIRECT) {
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

template <typename T, int MODE>
__global__ void __launch_bounds__(kFftThreads) k_yfft_inv(YArgs<T> a) {
    yfft_inv_tile<T, MODE>(a, xcd_tile(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y), gridDim.x,
                           linear_bid());
}

// TI complex rows (transformed in place) + the two twiddle tables
static inline size_t ylds_bytes(int TI, int ld, size_t tsz) { return (2 * (size_t)TI * ld + 4 * (size_t)ld) * tsz; }

static inline int ilog2(int v) {
    int l = 0;
    while ((1 << l) < v) ++l;
    return l;
}

template <typename T>
void launch_yfft_fwd(hipStream_t st, const Dims& d, const FftPlan& fp, const Tables<T>& tb, const FwdInput& in,
                     const Op64& op, const Reduce& rd, T* S, int TI) {
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
    a.op = op;
    a.part = rd.part;
    a.red = rd.red;
    a.S = S;
    a.d = d;
    a.fp = fp;
    a.tb = tb;
    a.TI = TI;
    a.lti = ilog2(TI);
    a.ld = fft_ld(d.L / 2);
    a.shift = shift;
    dim3 grid(ceil_div(d.M, TI), d.N);
    size_t sh = ylds_bytes(TI, a.ld, sizeof(T));
    if (mode == YF_DIRECT) {
        (void)hipFuncSetAttribute((const void*)k_yfft_fwd<T, YF_DIRECT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)sh);
        hipLaunchKernelGGL((k_yfft_fwd<T, YF_DIRECT>), grid, dim3(kFftThreads), sh, st, a);
    } else if (mode == YF_DIVERGENCE) {
        a.us = in.us;
        a.vs = in.vs;
        a.ws = in.ws;
        a.mt = in.mt;
        a.scale = in.scale;
        (void)hipFuncSetAttribute((const void*)k_yfft_fwd<T, YF_DIVERGENCE>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
        hipLaunchKernelGGL((k_yfft_fwd<T, YF_DIVERGENCE>), grid, dim3(kFftThreads), sh, st, a);
        const int nb = (int)(grid.x * grid.y);
        launch_fold(st, rd.part, nb, rd.red + RED_VOLB, RED_OP_SUM, 0);
        launch_fill(st, rd.red + RED_VOL, 1, in.mt.vol_total);
    } else {
        (void)hipFuncSetAttribute((const void*)k_yfft_fwd<T, YF_RESIDUAL>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)sh);
        hipLaunchKernelGGL((k_yfft_fwd<T, YF_RESIDUAL>), grid, dim3(kFftThreads), sh, st, a);
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
    a.ld = fft_ld(d.L / 2);
    dim3 grid(ceil_div(d.M, TI), d.N);
    size_t sh = ylds_bytes(TI, a.ld, sizeof(T));
    switch (mode) {
        case YI_DIRECT:
            (void)hipFuncSetAttribute((const void*)k_yfft_inv<T, YI_DIRECT>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
            hipLaunchKernelGGL((k_yfft_inv<T, YI_DIRECT>), grid, dim3(kFftThreads), sh, st, a);
            break;
        case YI_STORE:
            (void)hipFuncSetAttribute((const void*)k_yfft_inv<T, YI_STORE>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
            hipLaunchKernelGGL((k_yfft_inv<T, YI_STORE>), grid, dim3(kFftThreads), sh, st, a);
            launch_fold(st, rd.part, (int)(grid.x * grid.y), rd.red + RED_PHISUM, RED_OP_SUM, 0);
            break;
        default:
            (void)hipFuncSetAttribute((const void*)k_yfft_inv<T, YI_UPDATE>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
            hipLaunchKernelGGL((k_yfft_inv<T, YI_UPDATE>), grid, dim3(kFftThreads), sh, st, a);
            launch_fold(st, rd.part, (int)(grid.x * grid.y), rd.red + RED_PHISUM, RED_OP_SUM, 0);
            break;
    }
}

// ------------------------------------------------------------------ batched (persistent)
template <typename T, int MODE>
__global__ void __launch_bounds__(kFftThreads) k_batch_yfft_fwd(const Job<YArgs<T>>* jobs, int njobs, int total,
                                                        int* counters) {
    run_queue(jobs, njobs, total, counters,
              [](const Job<YArgs<T>>& J, int t) { yfft_fwd_tile<T, MODE>(J.a, t, J.ntx, t, J.ntiles); });
}

template <typename T, int MODE>
__global__ void __launch_bounds__(kFftThreads) k_batch_yfft_inv(const Job<YArgs<T>>* jobs, int njobs, int total,
                                                        int* counters) {
    run_queue(jobs, njobs, total, counters,
              [](const Job<YArgs<T>>& J, int t) { yfft_inv_tile<T, MODE>(J.a, t, J.ntx, t); });
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
            a.scale = p.scale;
        }
    } else {
        a.Sin = p.S;
        a.dp = p.dp;
    }
    return a;
}

int device_cus();

template <typename T>
int prepare_yfft(const std::vector<BatchProblem<T>>& pr, int fwd, int mode, StageLaunch& out,
                 std::vector<void*>& owned) {
    std::vector<int> order(pr.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return pr[x].tiy * pr[x].d.L > pr[y].tiy * pr[y].d.L; });
    std::vector<Job<YArgs<T>>> jobs;
    int total = 0;
    size_t lds = 0;
    for (int idx : order) {
        const BatchProblem<T>& p = pr[idx];
        Job<YArgs<T>> j{};
        j.a = yargs_of(p, fwd, mode);
        j.ntx = ceil_div(p.d.M, p.tiy);
        j.ntiles = j.ntx * p.d.N;
        j.first = total;
        total += j.ntiles;
        jobs.push_back(j);
        lds = std::max(lds, ylds_bytes(p.tiy, fft_ld(p.d.L / 2), sizeof(T)));
    }
    void* buf = nullptr;
    if (hipMalloc(&buf, jobs.size() * sizeof(jobs[0])) != hipSuccess) return -1;
    owned.push_back(buf);
    if (hipMemcpy(buf, jobs.data(), jobs.size() * sizeof(jobs[0]), hipMemcpyHostToDevice) != hipSuccess) return -1;
    out.jobs = buf;
    out.njobs = (int)jobs.size();
    out.total = total;
    out.lds = lds;
    return 0;
}

template <typename T>
void batch_yfft(hipStream_t st, const StageLaunch& sl, int fwd, int mode, int* counters) {
    auto go = [&](auto fn) {
        (void)hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sl.lds);
        int occ = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, kFftThreads, sl.lds) != hipSuccess || occ < 1) occ = 1;
        int grid = std::min(occ * device_cus(), sl.total + kQueues);
        grid = ceil_div(std::max(grid, kQueues), kQueues) * kQueues;
        hipLaunchKernelGGL(fn, dim3(grid), dim3(kFftThreads), sl.lds, st, (const Job<YArgs<T>>*)sl.jobs, sl.njobs, sl.total,
                           counters);
    };
    if (fwd) {
        if (mode == YF_DIRECT) go(k_batch_yfft_fwd<T, YF_DIRECT>);
        else if (mode == YF_RESIDUAL) go(k_batch_yfft_fwd<T, YF_RESIDUAL>);
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
template void batch_yfft<double>(hipStream_t, const StageLaunch&, int, int, int*);
template void batch_yfft<float>(hipStream_t, const StageLaunch&, int, int, int*);

template void launch_yfft_fwd<double>(hipStream_t, const Dims&, const FftPlan&, const Tables<double>&,
                                      const FwdInput&, const Op64&, const Reduce&, double*, int);
template void launch_yfft_fwd<float>(hipStream_t, const Dims&, const FftPlan&, const Tables<float>&,
                                     const FwdInput&, const Op64&, const Reduce&, float*, int);
template void launch_yfft_inv<double>(hipStream_t, const Dims&, const FftPlan&, const Tables<double>&,
                                      const double*, double*, const Reduce&, int, int);
template void launch_yfft_inv<float>(hipStream_t, const Dims&, const FftPlan&, const Tables<float>&,
                                     const float*, double*, const Reduce&, int, int);

}  // namespace eddy
