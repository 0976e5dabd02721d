// ztrans.hip -- transform along x2 (the reference's Z, 3rd index) that diagonalizes the
// x2 part of DIRECT's operator (AN/BN/CN, direct.f:370-378) for every y-mode.
//
// Uniform x2 spacing: the Neumann second difference is diagonalized by the DCT-II,
//   forward  X[m] = sum_k x[k] cos(pi m (2k+1) / (2N)),
//   inverse  x[k] = (X[0] + 2 sum_{m>0} X[m] cos(pi m (2k+1)/(2N))) / N,
// evaluated with Makhoul's reordering through one complex FFT of length N/2.
// Stretched x2: dense eigenvector matrices (generalized symmetric eigenproblem of the x2
// operator, computed on the host), applied as a small GEMM per y-mode.
//
// A workgroup owns TI consecutive x1 lines of one y-mode j: it reads rows of TI contiguous
// values S[i + M*(k + N*j)] for all k into LDS, transforms, and writes back in place.
#include <algorithm>
#include <numeric>

#include "batch_device.h"
#include "fft_device.h"
#include "kernels.h"
#include "stencil_device.h"

namespace eddy {

template <typename T>
struct ZArgs {
    T* S;
    const T* linesum;
    const double* shift_red;  // forward: zero y-mode shift (L * VOLB/VOL) when non-null
    Dims d;
    FftPlan fp;  // complex FFT of length N/2 (DCT mode)
    Tables<T> tb;
    int TI, lti;  // lines per workgroup (power of two) and its log2
    int ld;
    int j0;       // first y-mode of this launch (launches may cover a chunk of the modes)
};

constexpr int kZBatch = 8;  // x2 rows per thread and load batch of the dense tile kernel
constexpr int kZFwdRows = 8;  // x2 rows per thread and load batch of the forward DCT
constexpr int kZInvTasks = 2;  // packing tasks (4 coefficient rows each) per thread and batch

// Plain-mean correction of the zero y-mode (direct.f:894-898): shift the null x2 mode of
// every x1 line by delta = -(sum_m wsum[m] S_m) / (wsum[null] * M), S_m = sum_i x(i, m).
template <typename T>
__device__ T mean_delta(const ZArgs<T>& a) {
    __shared__ double part[16];
    double s = 0.0;
    for (int m = threadIdx.x; m < a.d.N; m += blockDim.x) s += (double)a.tb.wsum[m] * (double)a.linesum[m];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    double tot = 0.0;
    for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) tot += part[w];
    __syncthreads();
    return T(-tot / ((double)a.tb.wsum[a.d.null_m] * (double)a.d.M));
}

// Copy the DCT tables into LDS (layout: twc H, twr H+1, dct N complex values, split re/im).
template <typename T>
__device__ __forceinline__ void load_ztables(const Tables<T>& tb, T* tw, int N, int tid, int nthr, T*& twc_r,
                                             T*& twc_i, T*& twr_r, T*& twr_i, T*& dct_r, T*& dct_i) {
    const int H = N / 2;
    twc_r = tw;
    twc_i = twc_r + H;
    twr_r = twc_i + H;
    twr_i = twr_r + H + 1;
    dct_r = twr_i + H + 1;
    dct_i = dct_r + N;
    for (int t = tid; t < N; t += nthr) {
        if (t < H) {
            twc_r[t] = tb.ztwc_re[t];
            twc_i[t] = tb.ztwc_im[t];
        }
        if (t <= H) {
            twr_r[t] = tb.ztwr_re[t];
            twr_i[t] = tb.ztwr_im[t];
        }
        dct_r[t] = tb.zdct_re[t];
        dct_i[t] = tb.zdct_im[t];
    }
    if (tid == 0) {  // t = H when N = 2 (loop above stops at N - 1)
        twr_r[H] = tb.ztwr_re[H];
        twr_i[H] = tb.ztwr_im[H];
    }
}

// ---------------------------------------------------------------- DCT, forward
template <typename T>
__device__ __forceinline__ void zdct_fwd_tile(const ZArgs<T>& a, int tile, int ntx) {
    extern __shared__ unsigned char smem_raw[];
    T* smem = reinterpret_cast<T*>(smem_raw);
    const Dims d = a.d;
    const int TI = a.TI, ld = a.ld, N = d.N, H = N / 2;
    T* Ar = smem;  // TI complex rows, transformed in place
    T* Ai = Ar + TI * ld;
    T* tw = Ai + TI * ld;  // twc (H), twr (H+1), dct (N) complex tables
    const int i0 = (tile % ntx) * TI;
    const int j = a.j0 + tile / ntx;
    const int tid = threadIdx.x, nthr = blockDim.x;
    T *twc_r, *twc_i, *twr_r, *twr_i, *dct_r, *dct_i;
    load_ztables(a.tb, tw, N, tid, nthr, twc_r, twc_i, twr_r, twr_i, dct_r, dct_i);
    T* base = a.S + (size_t)d.M * (size_t)N * (size_t)j;
    // Makhoul order: v[k/2] = x[k] (k even), v[N-1-(k-1)/2] = x[k] (k odd); z[q] = v[2q] + i v[2q+1]
    const int lti = a.lti;
    const T zshift = (j == 0 && a.shift_red)
                         ? T((double)d.L * a.shift_red[RED_VOLB] / a.shift_red[RED_VOL]) : T(0);
    // The thread's x1 offset is fixed (blockDim is a multiple of TI). Every load of a batch is
    // issued before any value is used (arithmetic inside a load's own guard makes the compiler
    // wait for it before the next load); rows past N and lines past M read a clamped
    // in-range value that is zeroed afterwards.
    const int ii = tid & (TI - 1);
    const bool iok = i0 + ii < d.M;
    const T* colp = base + (iok ? i0 + ii : d.M - 1);
    const int kstep = nthr >> lti;
    // x2 rows as 32-bit byte offsets (S < 4 GiB), stepped with the row index
    const unsigned rb = (unsigned)sizeof(T) * (unsigned)d.M, klast = (unsigned)(N - 1) * rb;
    unsigned ko = (unsigned)(tid >> lti) * rb;
    for (int k0 = tid >> lti; k0 < N; k0 += kstep * kZFwdRows, ko += (unsigned)(kstep * kZFwdRows) * rb) {
        T vv[kZFwdRows];
#pragma unroll
        for (int u = 0; u < kZFwdRows; ++u)
            vv[u] = ld_off(colp, (k0 + u * kstep < N) ? ko + (unsigned)(u * kstep) * rb : klast);
#pragma unroll
        for (int u = 0; u < kZFwdRows; ++u) {
            // rows past N store into the row's last padding element, never read (fft_ld leaves
            // it unused): no branch, so the compiler keeps every load of the batch ahead of the
            // stores instead of sinking one into a guarded block and waiting for it there
            const int k = k0 + u * kstep;
            const bool kok = k < N;
            const T x = (iok && kok) ? vv[u] - zshift : T(0);
            const int p = (k & 1) ? (N - 1 - (k >> 1)) : (k >> 1);
            T* const row = ((p & 1) ? Ai : Ar) + ii * ld;
            row[kok ? pidx((p >> 1)) : ld - 1] = x;
        }
    }
    __syncthreads();
    if (H > 1) cfft_lds<T>(Ar, Ai, ld, TI, a.fp, twc_r, twc_i, T(-1), tid, nthr);
    // V_q = (Z_q + conj Z_{H-q})/2 - i W^q (Z_q - conj Z_{H-q})/2 (Z_H = Z_0), q = 0..H, then
    // X[m] = Re(exp(-i pi m/(2N)) V_m) and X[N-m] = Re(exp(-i pi (N-m)/(2N)) conj V_m).
    // Task (ii, q <= H/2) forms V_q and V_{H-q} and writes their (up to) four outputs.
    const int half = H / 2;
    for (int e = tid; e < TI * (half + 1); e += nthr) {
        const int ii = e & (TI - 1);
        const int q = e >> lti;
        const int gi = i0 + ii;
        if (gi >= d.M) continue;
        const T* zr = Ar + ii * ld;
        const T* zi = Ai + ii * ld;
        const int qb = H - q;  // partner (q = 0 pairs with H)
        const int ka = q, kb = (q == 0) ? 0 : qb;
        const T ar = zr[pidx(ka)], ai = zi[pidx(ka)];
        const T br = zr[pidx(kb)], bi = zi[pidx(kb)];
        T* col = base + gi;
        // V_q from (Z_ka, conj Z_kb); V_qb from (Z_kb, conj Z_ka)
        for (int s = 0; s < 2; ++s) {
            const int m = s ? qb : q;
            if (s && qb == q) break;
            const T pr = s ? br : ar, pi = s ? bi : ai;
            const T cr = s ? ar : br, ci = s ? -ai : -bi;
            const T er = T(0.5) * (pr + cr), ei = T(0.5) * (pi + ci);
            const T dr = T(0.5) * (pr - cr), di = T(0.5) * (pi - ci);
            const T wr = twr_r[m], wi = twr_i[m];
            const T wdr = wr * dr - wi * di, wdi = wr * di + wi * dr;
            const T vr = er + wdi, vi = ei - wdr;
            const unsigned mo = (unsigned)m * rb;
            st_off(col, mo, dct_r[m] * vr - dct_i[m] * vi);
            if (m > 0 && m < H) st_off(col, (unsigned)N * rb - mo, dct_r[N - m] * vr + dct_i[N - m] * vi);
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(kFftThreads) k_zdct_fwd(ZArgs<T> a) {
    zdct_fwd_tile<T>(a, xcd_tile(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y), gridDim.x);
}

// ---------------------------------------------------------------- DCT, inverse

template <typename T>
__device__ __forceinline__ void zdct_inv_tile(const ZArgs<T>& a, int tile, int ntx) {
    extern __shared__ unsigned char smem_raw[];
    T* smem = reinterpret_cast<T*>(smem_raw);
    const Dims d = a.d;
    const int TI = a.TI, ld = a.ld, N = d.N, H = N / 2;
    T* Ar = smem;  // TI complex rows, transformed in place
    T* Ai = Ar + TI * ld;
    T* tw = Ai + TI * ld;
    const int i0 = (tile % ntx) * TI;
    const int j = a.j0 + tile / ntx;
    const int tid = threadIdx.x, nthr = blockDim.x;
    T *twc_r, *twc_i, *twr_r, *twr_i, *dct_r, *dct_i;
    load_ztables(a.tb, tw, N, tid, nthr, twc_r, twc_i, twr_r, twr_i, dct_r, dct_i);
    T* base = a.S + (size_t)d.M * (size_t)N * (size_t)j;
    T delta = T(0);
    if (j == 0) delta = mean_delta(a);
    __syncthreads();  // the DCT tables are read below by other threads than wrote them
    const int lti = a.lti;
    // V_m = exp(i pi m/(2N)) (X_m - i X_{N-m}) (X_N = 0, V_0 and V_H real), m = 0..H;
    // Z'_q = [(V_q + conj V_{H-q}) + i conj(W^q) (V_q - conj V_{H-q})] / N, q < H.
    // Task (ii, q <= H/2) reads X_q, X_{N-q}, X_{H-q}, X_{H+q} straight from S (the plain-mean
    // shift of direct.f:894-898 added to the null x2 mode of y-mode 0) and forms Z'_q, Z'_{H-q}.
    const T invN = T(1) / T(N);
    const int half = H / 2;
    const int nm = (j == 0) ? d.null_m : -1;
    constexpr int UB = kZInvTasks;
    // loads first, then arithmetic (as in the forward transform); task rows past H/2 and lines
    // past M read clamped in-range values, X_N (q = 0) is 0
    const int ii = tid & (TI - 1);
    const bool iok = i0 + ii < d.M;
    const T* col = base + (iok ? i0 + ii : d.M - 1);
    const int qstep = nthr >> lti;
    // x2 rows as 32-bit byte offsets (S < 4 GiB): rows qc, N - qc (N - 1 for qc = 0: X_N is 0
    // and its load only stays in range), H - qc and H + qc from one product qc * rb
    const unsigned rb = (unsigned)sizeof(T) * (unsigned)d.M, nrb = (unsigned)N * rb, hrb = (unsigned)H * rb;
    for (int q0 = tid >> lti; q0 <= half; q0 += qstep * UB) {
        T x[UB][4];
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const int qc = min(q0 + u * qstep, half);
            const unsigned qo = (unsigned)qc * rb;
            x[u][0] = ld_off(col, qo);
            x[u][1] = ld_off(col, qc == 0 ? nrb - rb : nrb - qo);
            x[u][2] = ld_off(col, hrb - qo);
            x[u][3] = ld_off(col, hrb + qo);
        }
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const int q = q0 + u * qstep;
            if (q > half) continue;
            const int ms[4] = {q, N - q, H - q, H + q};
#pragma unroll
            for (int c = 0; c < 4; ++c)
                x[u][c] = (iok && ms[c] < N) ? x[u][c] + ((ms[c] == nm) ? delta : T(0)) : T(0);
            const int qb = H - q;
            // V_q (from X_q, X_{N-q}) and V_qb (from X_{H-q}, X_{H+q} = X_{N-qb})
            T vr[2], vi[2];
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const int m = s ? qb : q;
                const T xr = x[u][2 * s], xi = -x[u][2 * s + 1];
                const T cr = dct_r[m], ci = -dct_i[m];
                vr[s] = cr * xr - ci * xi;
                vi[s] = (m == 0 || m == H) ? T(0) : cr * xi + ci * xr;
            }
            T* rr = Ar + ii * ld;
            T* ri = Ai + ii * ld;
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const int kk = s ? qb : q;
                if (kk >= H || (s && qb == q)) continue;
                const T xr = vr[s], xi = vi[s];
                const T yr = vr[1 - s], yi = -vi[1 - s];
                const T sr = xr + yr, si = xi + yi;
                const T dr = xr - yr, di = xi - yi;
                const T wr = twr_r[kk], wi = -twr_i[kk];
                const T wvr = wr * dr - wi * di, wvi = wr * di + wi * dr;
                rr[pidx(kk)] = (sr - wvi) * invN;
                ri[pidx(kk)] = (si + wvr) * invN;
            }
        }
    }
    __syncthreads();
    if (H > 1) cfft_lds<T>(Ar, Ai, ld, TI, a.fp, twc_r, twc_i, T(1), tid, nthr);
    const T* Zr = Ar;
    const T* Zi = Ai;
    for (int e = tid; e < TI * N; e += nthr) {
        const int ii = e & (TI - 1);
        const int k = e >> lti;
        const int gi = i0 + ii;
        if (gi >= d.M) continue;
        const int p = (k & 1) ? (N - 1 - (k >> 1)) : (k >> 1);
        const T v = (p & 1) ? Zi[ii * ld + pidx((p >> 1))] : Zr[ii * ld + pidx((p >> 1))];
        st_off(base, (unsigned)sizeof(T) * (unsigned)gi + (unsigned)k * rb, v);
    }
}

template <typename T>
__global__ void __launch_bounds__(kFftThreads) k_zdct_inv(ZArgs<T> a) {
    zdct_inv_tile<T>(a, xcd_tile(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y), gridDim.x);
}

// ---------------------------------------------------------------- dense transform
// out[o][i] = sum_s A[s*N + o] * in[s][i] for the TI lines of the tile, in place.
// Each thread accumulates up to 8 outputs (o = og + 64*q) for 4 consecutive x1 lines.

template <typename T, int DIR>
__device__ __forceinline__ void zdense_tile(const ZArgs<T>& a, int tl, int ntx) {
    extern __shared__ unsigned char smem_raw[];
    T* tile = reinterpret_cast<T*>(smem_raw);  // N * TI
    const Dims d = a.d;
    const int TI = a.TI, N = d.N;
    const int i0 = (tl % ntx) * TI;
    const int j = a.j0 + tl / ntx;
    const int tid = threadIdx.x, nthr = blockDim.x;
    T* base = a.S + (size_t)d.M * (size_t)N * (size_t)j;
    const T* A = (DIR == 0) ? a.tb.qf : a.tb.qiT;
    T delta = T(0);
    if (DIR == 1 && j == 0) delta = mean_delta(a);
    const T zshift = (DIR == 0 && j == 0 && a.shift_red)
                         ? T((double)d.L * a.shift_red[RED_VOLB] / a.shift_red[RED_VOL]) : T(0);
    const int lti = a.lti;
    for (int e0 = tid; e0 < TI * N; e0 += nthr * kZBatch) {
        T vv[kZBatch];
#pragma unroll
        for (int u = 0; u < kZBatch; ++u) {
            const int e = e0 + u * nthr;
            const int gi = i0 + (e & (TI - 1));
            vv[u] = (e < TI * N && gi < d.M) ? base[(size_t)gi + (size_t)d.M * (e >> lti)] : T(0);
        }
#pragma unroll
        for (int u = 0; u < kZBatch; ++u) {
            const int e = e0 + u * nthr;
            if (e >= TI * N) continue;
            const int s = e >> lti;
            tile[e] = vv[u] + ((DIR == 1 && j == 0 && s == d.null_m) ? delta : T(0)) -
                      ((i0 + (e & (TI - 1)) < d.M) ? zshift : T(0));
        }
    }
    __syncthreads();
    const int ngi = TI / 4;                 // groups of 4 lines
    const int og = tid / ngi;               // output group 0..(256/ngi - 1)
    const int ig = tid % ngi;
    const int ostep = nthr / ngi;
    constexpr int QO = 8;
    T acc[QO][4];
#pragma unroll
    for (int q = 0; q < QO; ++q)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[q][c] = T(0);
    for (int obase = 0; obase < N; obase += ostep * QO) {
        for (int s = 0; s < N; ++s) {
            const T* trow = tile + s * TI + 4 * ig;
            const T x0 = trow[0], x1 = trow[1], x2 = trow[2], x3 = trow[3];
            const T* arow = A + (size_t)s * N;
#pragma unroll
            for (int q = 0; q < QO; ++q) {
                const int o = obase + og + q * ostep;
                const T w = (o < N) ? arow[o] : T(0);
                acc[q][0] += w * x0;
                acc[q][1] += w * x1;
                acc[q][2] += w * x2;
                acc[q][3] += w * x3;
            }
        }
#pragma unroll
        for (int q = 0; q < QO; ++q) {
            const int o = obase + og + q * ostep;
            if (o < N) {
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int gi = i0 + 4 * ig + c;
                    if (gi < d.M) base[(size_t)gi + (size_t)d.M * o] = acc[q][c];
                    acc[q][c] = T(0);
                }
            }
        }
    }
}

template <typename T, int DIR>
__global__ void __launch_bounds__(256) k_zdense(ZArgs<T> a) {
    zdense_tile<T, DIR>(a, xcd_tile(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y), gridDim.x);
}

// TI complex rows (transformed in place) + the DCT tables (twc H, twr H+1, dct N complex)
static inline size_t zlds_bytes(int TI, int ld, int N, size_t tsz) {
    return (2 * (size_t)TI * ld + 4 * (size_t)N + 2) * tsz;
}

template <typename T>
void launch_ztrans_fwd(hipStream_t st, const Dims& d, const FftPlan& fpz, const Tables<T>& tb, T* S, int TI,
                       int dense, const double* shift_red, int j0, int nj) {
    ZArgs<T> a{};
    a.S = S;
    a.shift_red = shift_red;
    a.d = d;
    a.fp = fpz;
    a.tb = tb;
    a.TI = TI;
    a.lti = 0;
    while ((1 << a.lti) < TI) ++a.lti;
    a.ld = fft_ld(d.N / 2);
    a.j0 = j0;
    dim3 grid(ceil_div(d.M, TI), nj);
    if (dense) {
        size_t sh = (size_t)d.N * TI * sizeof(T);
        (void)hipFuncSetAttribute((const void*)k_zdense<T, 0>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
        hipLaunchKernelGGL((k_zdense<T, 0>), grid, dim3(256), sh, st, a);
    } else {
        size_t sh = zlds_bytes(TI, a.ld, d.N, sizeof(T));
        (void)hipFuncSetAttribute((const void*)k_zdct_fwd<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
        hipLaunchKernelGGL((k_zdct_fwd<T>), grid, dim3(kFftThreads), sh, st, a);
    }
}

template <typename T>
void launch_ztrans_inv(hipStream_t st, const Dims& d, const FftPlan& fpz, const Tables<T>& tb, T* S,
                       const T* linesum, int TI, int dense, int j0, int nj) {
    ZArgs<T> a{};
    a.S = S;
    a.linesum = linesum;
    a.d = d;
    a.fp = fpz;
    a.tb = tb;
    a.TI = TI;
    a.lti = 0;
    while ((1 << a.lti) < TI) ++a.lti;
    a.ld = fft_ld(d.N / 2);
    a.j0 = j0;
    dim3 grid(ceil_div(d.M, TI), nj);
    if (dense) {
        size_t sh = (size_t)d.N * TI * sizeof(T);
        (void)hipFuncSetAttribute((const void*)k_zdense<T, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
        hipLaunchKernelGGL((k_zdense<T, 1>), grid, dim3(256), sh, st, a);
    } else {
        size_t sh = zlds_bytes(TI, a.ld, d.N, sizeof(T));
        (void)hipFuncSetAttribute((const void*)k_zdct_inv<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
        hipLaunchKernelGGL((k_zdct_inv<T>), grid, dim3(kFftThreads), sh, st, a);
    }
}

// ------------------------------------------------------------------ batched (one item per workgroup)
// One kernel per x2 transform kind (dense or DCT): a kernel holding both would take the
// registers of the dense tile for the DCT problems too.
template <typename T, int DIR, int DENSE>
__global__ void __launch_bounds__(kFftThreads) k_batch_ztrans(const Job<ZArgs<T>>* jobs, int njobs, int total) {
    run_item(jobs, njobs, total, [](const Job<ZArgs<T>>& J, int t) {
        if (DENSE) zdense_tile<T, DIR>(J.a, t, J.ntx);
        else if (DIR == 0) zdct_fwd_tile<T>(J.a, t, J.ntx);
        else zdct_inv_tile<T>(J.a, t, J.ntx);
    });
}


template <typename T>
int prepare_ztrans(const std::vector<BatchProblem<T>>& pr, int dir, int shift, StageLaunch& out,
                   std::vector<void*>& owned) {
    auto lds_of = [&](int x) {
        const BatchProblem<T>& p = pr[x];
        return p.dense ? (size_t)p.d.N * p.tiz * sizeof(T) : zlds_bytes(p.tiz, fft_ld(p.d.N / 2), p.d.N, sizeof(T));
    };
    auto cost = [&](int x) { return (long)pr[x].tiz * pr[x].d.N * (pr[x].dense ? pr[x].d.N / 8 : 1); };
    std::vector<int> order(pr.size());
    std::iota(order.begin(), order.end(), 0);
    // grouped by transform kind (dense first); the tiles of most LDS first, then the costliest
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        if (pr[x].dense != pr[y].dense) return pr[x].dense > pr[y].dense;
        const int ox = lds_occupancy(lds_of(x)), oy = lds_occupancy(lds_of(y));
        return ox != oy ? ox < oy : cost(x) > cost(y);
    });
    std::vector<Job<ZArgs<T>>> jobs;
    out = StageLaunch{};
    for (int idx : order) {
        const BatchProblem<T>& p = pr[idx];
        const int kind = p.dense ? 1 : 0;
        Job<ZArgs<T>> j{};
        ZArgs<T>& a = j.a;
        a.S = p.S;
        a.linesum = p.linesum;
        a.shift_red = (dir == 0 && shift) ? p.red : nullptr;
        a.d = p.d;
        a.fp = p.fpz;
        a.tb = p.tb;
        a.TI = p.tiz;
        a.lti = 0;
        while ((1 << a.lti) < p.tiz) ++a.lti;
        a.ld = fft_ld(p.d.N / 2);
        a.j0 = 0;
        j.ntx = ceil_div(p.d.M, p.tiz);
        j.ntiles = j.ntx * p.d.L;
        j.aux = kind;
        if (!add_class_job(out, jobs, j, kind, lds_of(idx))) return -1;
    }
    return upload_jobs(jobs, out, owned);
}

template <typename T>
void batch_ztrans(hipStream_t st, const StageLaunch& sl, int dir) {
    for (int c = 0; c < sl.ncls; ++c) {
        const size_t lds = sl.cls_lds[c];
        const int total = sl.cls_total[c];
        const Job<ZArgs<T>>* jobs = (const Job<ZArgs<T>>*)sl.jobs + sl.cls_job[c];
        const int njobs = sl.cls_job[c + 1] - sl.cls_job[c];
        auto go = [&](auto fn) {
            (void)hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            hipLaunchKernelGGL(fn, dim3(std::max(total, 1)), dim3(kFftThreads), lds, st, jobs, njobs, total);
        };
        const bool dense = sl.cls_q[c] != 0;
        if (dir == 0) dense ? go(k_batch_ztrans<T, 0, 1>) : go(k_batch_ztrans<T, 0, 0>);
        else dense ? go(k_batch_ztrans<T, 1, 1>) : go(k_batch_ztrans<T, 1, 0>);
    }
}

template int prepare_ztrans<double>(const std::vector<BatchProblem<double>>&, int, int, StageLaunch&,
                                    std::vector<void*>&);
template int prepare_ztrans<float>(const std::vector<BatchProblem<float>>&, int, int, StageLaunch&,
                                   std::vector<void*>&);
template void batch_ztrans<double>(hipStream_t, const StageLaunch&, int);
template void batch_ztrans<float>(hipStream_t, const StageLaunch&, int);

template void launch_ztrans_fwd<double>(hipStream_t, const Dims&, const FftPlan&, const Tables<double>&, double*,
                                        int, int, const double*, int, int);
template void launch_ztrans_fwd<float>(hipStream_t, const Dims&, const FftPlan&, const Tables<float>&, float*,
                                       int, int, const double*, int, int);
template void launch_ztrans_inv<double>(hipStream_t, const Dims&, const FftPlan&, const Tables<double>&, double*,
                                        const double*, int, int, int, int);
template void launch_ztrans_inv<float>(hipStream_t, const Dims&, const FftPlan&, const Tables<float>&, float*,
                                       const float*, int, int, int, int);

}  // namespace eddy
