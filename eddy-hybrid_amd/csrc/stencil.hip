// stencil.hip -- the HBM-streaming FP64 kernels around the solve:
//   DIVERGENCE (timeindeg.f:944-946) fused with the DP/ALFXDT scaling (Eddy6.f:2715) and,
//     optionally, the volume-weighted solvability sums;
//   CORRECT (timeindeg.f:1002, 1012, 1022) over the faces of setup.f:17-72;
//   the 7-point operator of DIRECT (for residuals), the final mean removal + ghost fill of
//   the defect correction, and reductions.
// Element-wise kernels: one thread per interior cell, x1 fastest, 64 x1 cells by 4 y rows
// per block. Reducing kernels: a block owns 64 x1 cells of one x2 plane for every y (its
// threads stride over y), so a whole plane row reduces to one partial per block.
#include <algorithm>
#include <numeric>

#include "batch_device.h"
#include "kernels.h"
#include "reduce_device.h"
#include "stencil_device.h"

namespace eddy {

static inline dim3 cell_grid(const Dims& d) { return dim3(ceil_div(d.M, 64), ceil_div(d.L, 4), d.N); }
static inline dim3 red_grid(const Dims& d) { return dim3(ceil_div(d.M, 64), 1, d.N); }
static const dim3 kCellBlock(64, 4, 1);

int reduce_blocks(const Dims& d) { return ceil_div(d.M, 64) * d.N; }

// logical (x, y, z) block of an XCD-ordered 3-D grid
struct Blk {
    int x, y, z;
};
__device__ __forceinline__ Blk xcd_block() {
    const int gx = gridDim.x, gy = gridDim.y;
    const int t = xcd_tile(blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z), gx * gy * gridDim.z);
    return Blk{t % gx, (t / gx) % gy, t / (gx * gy)};
}

// The same blocks with the x2 planes taken in chunks of KC: x fastest, then the plane within
// the chunk, then y, then the chunk. A block's x2 neighbour plane is then read by a block
// dispatched gx blocks later on the same XCD, while it is still in that XCD's L2.
template <int KC>
__device__ __forceinline__ Blk xcd_block_kc() {
    const int gx = gridDim.x, gy = gridDim.y, gz = gridDim.z;
    const int t = xcd_tile(blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z), gx * gy * gz);
    const int cb = gx * gy * KC;
    const int c = t / cb;
    int r = t - c * cb;
    const int kc = min(KC, gz - c * KC);  // the last chunk may hold fewer planes
    const int x = r % gx;
    r /= gx;
    return Blk{x, r / kc, c * KC + r % kc};
}

__global__ void __launch_bounds__(256) k_divergence(Dims d, Metrics mt, const double* __restrict__ us,
                                                    const double* __restrict__ vs, const double* __restrict__ ws,
                                                    double* __restrict__ div, double scale, double* part,
                                                    int accumulate, const double* __restrict__ dscale) {
    const Blk bk = xcd_block();
    if (dscale) scale = dscale[1];
    const int i = bk.x * 64 + threadIdx.x;
    const int k = bk.z;
    double vb = 0.0;
    if (i < d.M) {
        const int I = i + 1, K = k + 1;
        for (int j = threadIdx.y; j < d.L; j += 4) {
            const double v = div_at(d, mt, us, vs, ws, i, j, k) * scale;
            div[sidx(d, I, j + 1, K)] = v;
            if (accumulate) vb += cell_vol(mt, i, j, k) * v;
        }
    }
    if (accumulate) block_partial<RED_OP_SUM>(vb, part, linear_bid());
}

__global__ void __launch_bounds__(256) k_div_minmax(Dims d, Metrics mt, const double* __restrict__ us,
                                                    const double* __restrict__ vs, const double* __restrict__ ws,
                                                    double* part) {
    const Blk bk = xcd_block();
    const int i = bk.x * 64 + threadIdx.x;
    const int k = bk.z;
    double mn = 1e300, mx = -1e300;
    if (i < d.M) {
        for (int j = threadIdx.y; j < d.L; j += 4) {
            const double v = div_at(d, mt, us, vs, ws, i, j, k);
            mn = fmin(mn, v);
            mx = fmax(mx, v);
        }
    }
    const int nb = gridDim.x * gridDim.y * gridDim.z;
    block_partial<RED_OP_MIN>(mn, part, linear_bid());
    block_partial<RED_OP_MAX>(mx, part + nb, linear_bid());
}

// Rows of CORRECT per thread: rows j0, j0 + 4, ... of a block of 64 x1 cells by 4 * kCorrectRows
// y rows; every load of a thread is issued before its first store (rows past L load a clamped
// in-range row and store nothing). Addresses are 32-bit byte offsets (fields below 4 GiB).
constexpr int kCorrectRows = 2;

template <int R>
__global__ void __launch_bounds__(256) k_correct(Dims d, Metrics mt, const double* __restrict__ us,
                                                 const double* __restrict__ vs, const double* __restrict__ ws,
                                                 double* __restrict__ uo, double* __restrict__ vo,
                                                 double* __restrict__ wo, const double* __restrict__ dp,
                                                 double alfxdt, int refresh, const double* __restrict__ dscale) {
    // dp(k+1) is this block's neighbour plane; plane chunks of 2 (chunks of 1, 4 and 8 measured
    // 2.6 %, 0.8 % and 1.4 % slower with two rows per thread)
    const Blk bk = xcd_block_kc<2>();
    if (dscale) alfxdt = dscale[0];
    const int i = bk.x * 64 + threadIdx.x;
    const int j0 = bk.y * (4 * R) + threadIdx.y;
    const int k = bk.z;
    if (i >= d.M || j0 >= d.L) return;
    const int I = i + 1, K = k + 1;
    const unsigned nx8 = 8u * (unsigned)d.nx, pl8 = nx8 * (unsigned)d.ny;
    const unsigned oik = 8u * (unsigned)I + pl8 * (unsigned)K;
    // U faces I = IBU..IEU: 2..NX-2 with x1 walls (MP=1), 2..NX-1 when x1 is periodic
    // (ITYPE(2)=500, setup.f:34-35; DP(NX) is then DIRECT's periodic ghost); W faces K =
    // KBW..KEW: 2..NZ-2 with x2 walls (NP=1), 2..NZ-1 when x2 is periodic (ITYPE(6)=500,
    // setup.f:70; DP(NZ) then holds REFRESHBC's periodic image)
    const bool uface = I <= d.nx - 3 + d.cyc_x, wface = K <= d.nz - 3 + d.cyc_z;
    const double auI = mt.au[I], cwK = mt.cw[K], rpI = mt.cart ? 1.0 : mt.rp[I];
    unsigned o[R];
    double p0[R], pe[R], pn[R], pt[R], u[R], v[R], w[R], bvJ[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int J = min(j0 + 4 * q, d.L - 1) + 1;
        o[q] = oik + nx8 * (unsigned)J;
        p0[q] = ld_off(dp, o[q]);
        pe[q] = ld_off(dp, o[q] + 8u);
        pn[q] = ld_off(dp, o[q] + nx8);
        pt[q] = ld_off(dp, o[q] + pl8);
        u[q] = ld_off(us, o[q]);
        v[q] = ld_off(vs, o[q]);
        w[q] = ld_off(ws, o[q]);
        bvJ[q] = mt.bv[J];
    }
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int j = j0 + 4 * q;
        if (j >= d.L) continue;
        const int J = j + 1;
        // periodic y ghost rows of the corrected faces (the velocity BC step after CORRECT,
        // Eddy6.f:2746-2761): row J=1 mirrors J=NY-1, row J=NY mirrors J=2
        const int gofs = (j == d.L - 1) ? -(int)(nx8 * (unsigned)(d.ny - 2))
                                        : ((j == 0) ? (int)(nx8 * (unsigned)(d.ny - 2)) : 0);
        const bool ghost = refresh && gofs != 0 && !d.ycopy;
        if (uface) {
            const double un = u[q] - auI * (pe[q] - p0[q]) * alfxdt;
            st_off(uo, o[q], un);
            if (ghost) st_off(uo, o[q] + (unsigned)gofs, un);
        }
        // V faces J = JBV..JEV: 2..NY-1 for the periodic y (ITYPE(4)=500, LP=0), 2..NY-2 for LP=1
        if (J <= d.ny - 2 - d.ycopy) {
            const double gv = bvJ[q] * (pn[q] - p0[q]) * alfxdt;
            const double vn = v[q] - (mt.cart ? gv : gv / rpI);
            st_off(vo, o[q], vn);
            if (ghost) st_off(vo, o[q] + (unsigned)gofs, vn);
        }
        if (wface) {
            const double wn = w[q] - cwK * (pt[q] - p0[q]) * alfxdt;
            st_off(wo, o[q], wn);
            if (ghost) st_off(wo, o[q] + (unsigned)gofs, wn);
        }
    }
}

__global__ void __launch_bounds__(256) k_laplacian(Dims d, Op64 op, const double* __restrict__ phi,
                                                   double* __restrict__ out) {
    const Blk bk = xcd_block();
    const int i = bk.x * 64 + threadIdx.x;
    const int j = bk.y * 4 + threadIdx.y;
    const int k = bk.z;
    if (i >= d.M || j >= d.L) return;
    out[sidx(d, i + 1, j + 1, k + 1)] = lap_at(d, op, phi, i, j, k);
}

// Partial sums over the interior of (rhs - L(phi))^2.
__global__ void __launch_bounds__(256) k_resid_sumsq(Dims d, Op64 op, const double* __restrict__ phi,
                                                     const double* __restrict__ rhs, double* part) {
    const Blk bk = xcd_block();
    const int i = bk.x * 64 + threadIdx.x;
    const int k = bk.z;
    double rr = 0.0;
    if (i < d.M) {
        for (int j = threadIdx.y; j < d.L; j += 4) {
            const double r = rhs[sidx(d, i + 1, j + 1, k + 1)] - lap_at(d, op, phi, i, j, k);
            rr += r * r;
        }
    }
    block_partial<RED_OP_SUM>(rr, part, linear_bid());
}

// phi <- phi - mean (mean = red[RED_PHISUM]/(M L N)) on the interior, then the ghosts as
// direct.f:1035-1106 leaves them: x ghosts copies (MP=1) or periodic images (MP=0), y ghosts
// periodic images (LP=0) or copies (LP=1), z ghost planes zero.
__global__ void __launch_bounds__(256) k_shift_ghosts(Dims d, double* __restrict__ phi, const double* red) {
    const Blk bk = xcd_block();
    const int i = bk.x * 64 + threadIdx.x;
    const int j = bk.y * 4 + threadIdx.y;
    const int k = bk.z;
    if (i >= d.M || j >= d.L) return;
    const double mean = red[RED_PHISUM] / ((double)d.M * (double)d.L * (double)d.N);
    const int I = i + 1, J = j + 1, K = k + 1;
    const double v = phi[sidx(d, I, J, K)] - mean;
    int Is[3], Js[3];
    int ni = 0, nj = 0;
    Is[ni++] = I;
    if (i == (d.cyc_x ? d.M - 1 : 0)) Is[ni++] = 0;
    if (i == (d.cyc_x ? 0 : d.M - 1)) Is[ni++] = d.nx - 1;
    Js[nj++] = J;
    if (j == (d.ycopy ? 0 : d.L - 1)) Js[nj++] = 0;
    if (j == (d.ycopy ? d.L - 1 : 0)) Js[nj++] = d.ny - 1;
    for (int p = 0; p < ni; ++p)
        for (int q = 0; q < nj; ++q) phi[sidx(d, Is[p], Js[q], K)] = v;
    if (k == 0 || k == d.N - 1) {
        const int Kg = (k == 0) ? 0 : d.nz - 1;
        for (int p = 0; p < ni; ++p)
            for (int q = 0; q < nj; ++q) phi[sidx(d, Is[p], Js[q], Kg)] = 0.0;
    }
}

__global__ void __launch_bounds__(256) k_sumsq(Dims d, const double* __restrict__ a, double* part) {
    const Blk bk = xcd_block();
    const int i = bk.x * 64 + threadIdx.x;
    const int k = bk.z;
    double s = 0.0;
    if (i < d.M) {
        for (int j = threadIdx.y; j < d.L; j += 4) {
            const double v = a[sidx(d, i + 1, j + 1, k + 1)];
            s += v * v;
        }
    }
    block_partial<RED_OP_SUM>(s, part, linear_bid());
}

// The RHS check: sum vol*b and sum vol*|b| over the finite interior values of b, and the count
// of non-finite ones, as three per-block partials (nb blocks each).
__global__ void __launch_bounds__(256) k_rhs_check(Dims d, Metrics mt, const double* __restrict__ b, double* part,
                                                   int nb) {
    const Blk bk = xcd_block();
    const int i = bk.x * 64 + threadIdx.x;
    const int k = bk.z;
    double s = 0.0, sa = 0.0, nf = 0.0;
    if (i < d.M) {
        for (int j = threadIdx.y; j < d.L; j += 4) {
            const double v = b[sidx(d, i + 1, j + 1, k + 1)];
            if (!isfinite(v)) {
                nf += 1.0;
                continue;
            }
            const double vol = cell_vol(mt, i, j, k);
            s += vol * v;
            sa += vol * fabs(v);
        }
    }
    block_partial<RED_OP_SUM>(s, part, linear_bid());
    block_partial<RED_OP_SUM>(sa, part + nb, linear_bid());
    block_partial<RED_OP_SUM>(nf, part + 2 * nb, linear_bid());
}

// dst <- src - red[RED_VOLB]/red[RED_VOL] on the interior (volume-weighted solvability
// shift of the RHS).
__global__ void __launch_bounds__(256) k_copy_shift(Dims d, const double* __restrict__ src,
                                                    double* __restrict__ dst, const double* red) {
    const Blk bk = xcd_block();
    const int i = bk.x * 64 + threadIdx.x;
    const int j = bk.y * 4 + threadIdx.y;
    const int k = bk.z;
    if (i >= d.M || j >= d.L) return;
    const double c = red[RED_VOLB] / red[RED_VOL];
    const size_t o = sidx(d, i + 1, j + 1, k + 1);
    dst[o] = src[o] - c;
}

// REFRESHBC (refreshbc.f:1-28) on one GPU: the z ghost planes exchanged with the
// neighbouring slabs. With x2 walls both neighbours are walls and the ghosts stay as DIRECT
// left them (zero); with a periodic x2 (ITYPE(5) = ITYPE(6) = 500) the one slab is its own
// neighbour: DP(:,:,1) = DP(:,:,NZ-1), DP(:,:,NZ) = DP(:,:,2).
__global__ void k_refresh_z(Dims d, double* __restrict__ dp) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t pl = (size_t)d.nx * d.ny;
    if (t >= pl) return;
    dp[t] = dp[t + pl * (size_t)(d.nz - 2)];
    dp[t + pl * (size_t)(d.nz - 1)] = dp[t + pl];
}

__global__ void k_fill(double* p, size_t n, double v) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

// out[0] (= or combined with) the reduction of part[0..n), one workgroup of 1024 threads.
template <int OP>
__global__ void __launch_bounds__(1024) k_fold(const double* __restrict__ part, int n, double* out, int add) {
    const double init = (OP == RED_OP_SUM) ? 0.0 : ((OP == RED_OP_MIN) ? 1e300 : -1e300);
    double v = init;
    for (int t = threadIdx.x; t < n; t += blockDim.x) v = combine<OP>(v, part[t]);
    __shared__ double w[16];
    v = wave_reduce<OP>(v);
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = w[0];
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) s = combine<OP>(s, w[q]);
        *out = add ? combine<OP>(*out, s) : s;
    }
}

// ------------------------------------------------------------------ batched (one item per workgroup)
struct CellArgs {
    Dims d;
    Metrics mt;
    const double *us, *vs, *ws;
    double *uo, *vo, *wo;
    double* dp;
    double* rhs;
    const double* red;
    const double* dscale;  // device {ALFXDT, 1/ALFXDT} of the projection
};

constexpr int kCellRows = 4;  // y rows per thread of a batch cell item

// One item: 64 x1 cells by 4 * kCellRows y rows of one x2 plane, threads 64 x 4 with rows
// j0, j0 + 4, ... each; items run x1 block fastest, then the y block, then the plane (ntx x1
// blocks, nty y blocks per plane). Every load of a thread is issued before its first store:
// rows past L load a clamped in-range row and store nothing. (One row per thread left the
// per-item set-up, the job search and argument loads, on every 256 cells: the copy stages
// ran 1.3-1.9x slower; a loop over all of a plane's rows waited on each row's loads in turn:
// CORRECT 1.26x slower.)
template <int OP>
__device__ __forceinline__ void cell_tile(const CellArgs& c, int t, int ntx, int nty) {
    constexpr int R = kCellRows;
    const Dims& d = c.d;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int xb = t % ntx, r = t / ntx;
    const int i = xb * 64 + tx;
    const int j0 = (r % nty) * (4 * R) + ty;
    const int k = r / nty;
    if (i >= d.M) return;
    const int I = i + 1, K = k + 1;
    size_t o[R];
#pragma unroll
    for (int q = 0; q < R; ++q) o[q] = sidx(d, I, min(j0 + 4 * q, d.L - 1) + 1, K);
    if (OP == BC_CORRECT) {
        const size_t pl = (size_t)d.nx * d.ny;
        const double alfxdt = c.dscale[0];
        const double au = c.mt.au[I], cw = c.mt.cw[K], rp = c.mt.cart ? 1.0 : c.mt.rp[I];
        double p0[R], pe[R], pn[R], pt[R], u[R], v[R], w[R], bv[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            p0[q] = c.dp[o[q]];
            pe[q] = c.dp[o[q] + 1];
            pn[q] = c.dp[o[q] + d.nx];
            pt[q] = c.dp[o[q] + pl];
            u[q] = c.us[o[q]];
            v[q] = c.vs[o[q]];
            w[q] = c.ws[o[q]];
            bv[q] = c.mt.bv[min(j0 + 4 * q, d.L - 1) + 1];
        }
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int j = j0 + 4 * q;
            if (j >= d.L) continue;
            const long gofs = (j == d.L - 1) ? -(long)d.nx * (d.ny - 2) : ((j == 0) ? (long)d.nx * (d.ny - 2) : 0);
            if (I <= d.nx - 3) {
                const double uu = u[q] - au * (pe[q] - p0[q]) * alfxdt;
                c.uo[o[q]] = uu;
                if (gofs) c.uo[o[q] + gofs] = uu;
            }
            const double gv = bv[q] * (pn[q] - p0[q]) * alfxdt;
            const double vv = v[q] - (c.mt.cart ? gv : gv / rp);
            c.vo[o[q]] = vv;
            if (gofs) c.vo[o[q] + gofs] = vv;
            if (K <= d.nz - 3) {
                const double ww = w[q] - cw * (pt[q] - p0[q]) * alfxdt;
                c.wo[o[q]] = ww;
                if (gofs) c.wo[o[q] + gofs] = ww;
            }
        }
    } else if (OP == BC_COPY_RHS) {
        double x[R];
#pragma unroll
        for (int q = 0; q < R; ++q) x[q] = c.dp[o[q]];
#pragma unroll
        for (int q = 0; q < R; ++q)
            if (j0 + 4 * q < d.L) c.rhs[o[q]] = x[q];
    } else {
        const double mean = c.red[RED_PHISUM] / ((double)d.M * (double)d.L * (double)d.N);
        double x[R];
#pragma unroll
        for (int q = 0; q < R; ++q) x[q] = c.dp[o[q]];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int j = j0 + 4 * q;
            if (j >= d.L) continue;
            const int J = j + 1;
            const double v = x[q] - mean;
            int Is[3], Js[3];
            int ni = 0, nj = 0;
            Is[ni++] = I;
            if (i == 0) Is[ni++] = 0;
            if (i == d.M - 1) Is[ni++] = d.nx - 1;
            Js[nj++] = J;
            if (j == d.L - 1) Js[nj++] = 0;
            if (j == 0) Js[nj++] = d.ny - 1;
            for (int p = 0; p < ni; ++p)
                for (int qq = 0; qq < nj; ++qq) c.dp[sidx(d, Is[p], Js[qq], K)] = v;
            if (k == 0 || k == d.N - 1) {
                const int Kg = (k == 0) ? 0 : d.nz - 1;
                for (int p = 0; p < ni; ++p)
                    for (int qq = 0; qq < nj; ++qq) c.dp[sidx(d, Is[p], Js[qq], Kg)] = 0.0;
            }
        }
    }
}

template <int OP>
__global__ void __launch_bounds__(256) k_batch_cells(const Job<CellArgs>* jobs, int njobs, int total) {
    run_item(jobs, njobs, total, [](const Job<CellArgs>& J, int t) { cell_tile<OP>(J.a, t, J.ntx, J.aux); });
}

template <typename T>
static int prepare_cells_t(const std::vector<BatchProblem<T>>& pr, int op, StageLaunch& out,
                           std::vector<void*>& owned) {
    std::vector<int> order(pr.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return pr[x].d.L > pr[y].d.L; });
    std::vector<Job<CellArgs>> jobs;
    int total = 0;
    for (int idx : order) {
        const BatchProblem<T>& p = pr[idx];
        Job<CellArgs> j{};
        CellArgs& c = j.a;
        c.d = p.d;
        c.mt = p.mt;
        c.us = p.us;
        c.vs = p.vs;
        c.ws = p.ws;
        c.uo = p.uo;
        c.vo = p.vo;
        c.wo = p.wo;
        c.dp = p.dp;
        c.rhs = const_cast<double*>(p.rhs);
        c.red = p.red;
        c.dscale = p.dscale;
        j.ntx = ceil_div(p.d.M, 64);
        j.aux = ceil_div(p.d.L, 4 * kCellRows);  // y blocks per plane
        j.ntiles = j.ntx * j.aux * p.d.N;
        j.first = total;
        total += j.ntiles;
        jobs.push_back(j);
    }
    void* buf = nullptr;
    if (hipMalloc(&buf, jobs.size() * sizeof(jobs[0])) != hipSuccess) return -1;
    owned.push_back(buf);
    if (hipMemcpy(buf, jobs.data(), jobs.size() * sizeof(jobs[0]), hipMemcpyHostToDevice) != hipSuccess) return -1;
    out.jobs = buf;
    out.njobs = (int)jobs.size();
    out.total = total;
    out.lds = 0;
    return 0;
}

int prepare_cells(const std::vector<BatchProblem<double>>& pr, int op, StageLaunch& out, std::vector<void*>& owned) {
    return prepare_cells_t(pr, op, out, owned);
}
int prepare_cells_f(const std::vector<BatchProblem<float>>& pr, int op, StageLaunch& out,
                    std::vector<void*>& owned) {
    return prepare_cells_t(pr, op, out, owned);
}

void batch_cells(hipStream_t st, const StageLaunch& sl, int op) {
    auto go = [&](auto fn) {
        hipLaunchKernelGGL(fn, dim3(std::max(sl.total, 1)), dim3(256), 0, st, (const Job<CellArgs>*)sl.jobs, sl.njobs,
                           sl.total);
    };
    if (op == BC_CORRECT) go(k_batch_cells<BC_CORRECT>);
    else if (op == BC_COPY_RHS) go(k_batch_cells<BC_COPY_RHS>);
    else go(k_batch_cells<BC_SHIFT_GHOSTS>);
}

__global__ void __launch_bounds__(1024) k_batch_fold(const FoldJob* jobs) {
    const FoldJob J = jobs[blockIdx.x];
    double v = (J.op == RED_OP_SUM) ? 0.0 : ((J.op == RED_OP_MIN) ? 1e300 : -1e300);
    for (int t = threadIdx.x; t < J.n; t += blockDim.x) {
        const double x = J.part[t];
        v = (J.op == RED_OP_SUM) ? v + x : ((J.op == RED_OP_MIN) ? fmin(v, x) : fmax(v, x));
    }
    __shared__ double w[16];
    if (J.op == RED_OP_SUM) v = wave_reduce<RED_OP_SUM>(v);
    else if (J.op == RED_OP_MIN) v = wave_reduce<RED_OP_MIN>(v);
    else v = wave_reduce<RED_OP_MAX>(v);
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = w[0];
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q)
            s = (J.op == RED_OP_SUM) ? s + w[q] : ((J.op == RED_OP_MIN) ? fmin(s, w[q]) : fmax(s, w[q]));
        *J.out = s;
    }
}

void batch_fold(hipStream_t st, const FoldJob* jobs, int njobs) {
    if (njobs > 0) hipLaunchKernelGGL(k_batch_fold, dim3(njobs), dim3(1024), 0, st, jobs);
}

void launch_fold(hipStream_t st, const double* part, int n, double* out, int op, int add) {
    switch (op) {
        case RED_OP_MIN: hipLaunchKernelGGL(k_fold<RED_OP_MIN>, dim3(1), dim3(1024), 0, st, part, n, out, add); break;
        case RED_OP_MAX: hipLaunchKernelGGL(k_fold<RED_OP_MAX>, dim3(1), dim3(1024), 0, st, part, n, out, add); break;
        default: hipLaunchKernelGGL(k_fold<RED_OP_SUM>, dim3(1), dim3(1024), 0, st, part, n, out, add); break;
    }
}

void launch_divergence(hipStream_t st, const Dims& d, const Metrics& mt, const double* us, const double* vs,
                       const double* ws, double* div, double scale, const Reduce& rd, int accumulate_compat,
                       const double* dscale) {
    hipLaunchKernelGGL(k_divergence, red_grid(d), kCellBlock, 0, st, d, mt, us, vs, ws, div, scale, rd.part,
                       accumulate_compat, dscale);
    if (accumulate_compat) {
        const int nb = reduce_blocks(d);
        launch_fold(st, rd.part, nb, rd.red + RED_VOLB, RED_OP_SUM, 0);
        launch_fill(st, rd.red + RED_VOL, 1, mt.vol_total);
    }
}

void launch_div_minmax(hipStream_t st, const Dims& d, const Metrics& mt, const double* us, const double* vs,
                       const double* ws, const Reduce& rd) {
    hipLaunchKernelGGL(k_div_minmax, red_grid(d), kCellBlock, 0, st, d, mt, us, vs, ws, rd.part);
    const int nb = reduce_blocks(d);
    launch_fold(st, rd.part, nb, rd.red + RED_DMIN, RED_OP_MIN, 0);
    launch_fold(st, rd.part + nb, nb, rd.red + RED_DMAX, RED_OP_MAX, 0);
}

void launch_correct(hipStream_t st, const Dims& d, const Metrics& mt, const double* us, const double* vs,
                    const double* ws, double* uo, double* vo, double* wo, const double* dp, double alfxdt,
                    int refresh_periodic, const double* dscale) {
    const dim3 grid(ceil_div(d.M, 64), ceil_div(d.L, 4 * kCorrectRows), d.N);
    hipLaunchKernelGGL(k_correct<kCorrectRows>, grid, kCellBlock, 0, st, d, mt, us, vs, ws, uo, vo, wo, dp, alfxdt,
                       refresh_periodic, dscale);
}

void launch_laplacian(hipStream_t st, const Dims& d, const Op64& op, const double* phi, double* out) {
    hipLaunchKernelGGL(k_laplacian, cell_grid(d), kCellBlock, 0, st, d, op, phi, out);
}

void launch_finalize(hipStream_t st, const Dims& d, const Op64& op, double* phi, const double* rhs,
                     const Reduce& rd) {
    // the residual reads phi's neighbours, so it runs before the shift rewrites phi
    if (rhs) {
        hipLaunchKernelGGL(k_resid_sumsq, red_grid(d), kCellBlock, 0, st, d, op, phi, rhs, rd.part);
        launch_fold(st, rd.part, reduce_blocks(d), rd.red + RED_RNORM2, RED_OP_SUM, 0);
    }
    hipLaunchKernelGGL(k_shift_ghosts, cell_grid(d), kCellBlock, 0, st, d, phi, (const double*)rd.red);
}

void launch_rhs_check(hipStream_t st, const Dims& d, const Metrics& mt, const double* b, const Reduce& rd) {
    const int nb = reduce_blocks(d);
    hipLaunchKernelGGL(k_rhs_check, red_grid(d), kCellBlock, 0, st, d, mt, b, rd.part, nb);
    launch_fold(st, rd.part, nb, rd.red + RED_CHK_VB, RED_OP_SUM, 0);
    launch_fold(st, rd.part + nb, nb, rd.red + RED_CHK_VABS, RED_OP_SUM, 0);
    launch_fold(st, rd.part + 2 * nb, nb, rd.red + RED_CHK_NF, RED_OP_SUM, 0);
}

void launch_sumsq(hipStream_t st, const Dims& d, const double* a, const Reduce& rd, int slot) {
    hipLaunchKernelGGL(k_sumsq, red_grid(d), kCellBlock, 0, st, d, a, rd.part);
    launch_fold(st, rd.part, reduce_blocks(d), rd.red + slot, RED_OP_SUM, 0);
}

void launch_copy_shift(hipStream_t st, const Dims& d, const double* src, double* dst, const double* red) {
    hipLaunchKernelGGL(k_copy_shift, cell_grid(d), kCellBlock, 0, st, d, src, dst, red);
}

void launch_refresh_z(hipStream_t st, const Dims& d, double* dp) {
    if (!d.cyc_z) return;
    const size_t pl = (size_t)d.nx * d.ny;
    hipLaunchKernelGGL(k_refresh_z, dim3((unsigned)((pl + 255) / 256)), dim3(256), 0, st, d, dp);
}

void launch_fill(hipStream_t st, double* p, size_t n, double v) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
}

}  // namespace eddy
