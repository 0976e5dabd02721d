// common.h -- shared host/device definitions of the projection library (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace eddy {

constexpr int kMaxRadices = 16;

// Radix plan of a complex FFT of length n = product of rad[0..count).
struct FftPlan {
    int n;
    int count;
    int rad[kMaxRadices];
};

// Operator and transform tables of one plan, in the precision T of the kernels that read
// them. All pointers are device memory.
template <typename T>
struct Tables {
    // x1 operator (direct.f:344-352, 536) in the x1 solve's thread order: entry r*P + p holds
    // row i = p*Q + r (P threads per line, Q rows each, rows past M padding)
    const T* am;     // x1 sub-diagonal   AM(i)
    const T* bm;     // x1 diagonal       BM(i) = -AM-CM
    const T* cm;     // x1 super-diagonal CM(i)
    const T* r2inv;  // 1/RP(i+1)^2 (multiplies AK in BML)
    const T* akj;    // AK(J/2+1) per fftpack slot j=0..L-1 (direct.f:321-331)
    const T* lam;    // x2 eigenvalues, m=0..N-1
    const T* lamj;   // per-line shift of the line solve instead of lam[m] (line j*N + m), or null
    const T* wsum;   // column sums of the inverse x2 transform (mean correction)
    const T* qf;     // dense forward x2 transform, row-major qf[m*N + k] (dense mode)
    const T* qiT;    // dense inverse x2 transform, transposed: qiT[m*N + k] = Qinv(k,m)
    // y real FFT of length L = 2H: twc[t] = exp(-2 pi i t/H) (t<H), twr[k] = exp(-2 pi i k/L)
    const T* ytwc_re; const T* ytwc_im; const T* ytwr_re; const T* ytwr_im;
    // x2 DCT of length N = 2H2: same tables for length N plus dct[m] = exp(-i pi m/(2N))
    const T* ztwc_re; const T* ztwc_im; const T* ztwr_re; const T* ztwr_im;
    const T* zdct_re; const T* zdct_im;
    // y real DFT as dense matrices for the matrix-core variant (FP32 tables only, else null):
    // ydft_f[j*L + p] = F(p, j) forward, ydft_b[p*L + j] = G(j, p) inverse with 1/L
    const T* ydft_f; const T* ydft_b;
};

struct Dims {
    int nx, ny, nz;   // DP(NX,NY,NZ)
    int M, L, N;      // interior counts (x1, y, x2)
    int null_m;       // index of the x2 null mode (eigenvalue 0)
    // boundary flags of DIRECT, zero = the walls + periodic-y default: cyc_x = 1 when MP = 0
    // (x1 periodic: AM(1), CM(M) couple cells 1 and M), cyc_z = 1 when NP = 0 (likewise AN(1),
    // CN(N) along x2), ycopy = 1 when LP = 1 (DIRECT's y ghosts copy J=2 and NY-1 instead of
    // taking the periodic images; the y transform is periodic either way, direct.f:1100-1106)
    int cyc_x, cyc_z, ycopy;
};

// Reduction scalars kept on the device (FP64, zeroed by a fill kernel each use).
enum RedSlot {
    RED_RNORM2 = 0,   // sum r^2 of the current residual
    RED_BNORM2 = 1,   // sum b^2
    RED_PHISUM = 2,   // sum of phi (plain mean)
    RED_VOLB = 3,     // sum vol*b (compatibility)
    RED_VOL = 4,      // sum vol
    RED_DMIN = 5,
    RED_DMAX = 6,
    RED_PREF = 7,     // reference pressure of the pressure update
    RED_CHK_VB = 8,   // RHS check: sum vol*b, sum vol*|b|, count of non-finite values
    RED_CHK_VABS = 9,
    RED_CHK_NF = 10,
    RED_COUNT = 11
};

__host__ __device__ inline int ceil_div(int a, int b) { return (a + b - 1) / b; }

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on XCD b % 8) and
// each XCD has its own L2. Give workgroups b, b+8, b+16, ... (one XCD) consecutive logical
// tiles, so neighbouring tiles -- which share cache lines at their edges and, for stencils,
// neighbouring planes -- meet in the same L2. Bijective on [0, n).
__device__ __forceinline__ int xcd_tile(int b, int n) {
    const int q = n >> 3, r = n & 7;
    const int x = b & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// Tiles (ix, k) of an ntx x nk grid taken in blocks of W consecutive ix: ix fastest within a
// block, then k, then the block (the last block may be narrower). Consecutive logical tiles
// then walk down x2 over a narrow x1 range, so a tile's x2 neighbour planes were touched a few
// tiles earlier and their working set stays in the XCD's L2. W <= 0 (or >= ntx): ix fastest
// over the whole row. Returns the row-major tile index ix + ntx * k; bijective on [0, ntx*nk).
__device__ __forceinline__ int iblock_tile(int t, int ntx, int nk, int W) {
    if (W <= 0 || W >= ntx) return t;
    const int per = W * nk;
    const int b = t / per;
    const int r = t - b * per;
    const int wb = min(W, ntx - b * W);
    const int k = r / wb;
    return b * W + (r - k * wb) + ntx * k;
}

}  // namespace eddy
