// xsolve.hip -- the x1 tridiagonal solves of every (y-mode j, x2-mode m) line.
//
// After the y-FFT and the x2 transform, DIRECT's per-mode 2-D problem (direct.f:788-797)
// separates into one tridiagonal system per line:
//   AM(i) x(i-1) + (BM(i) - AK_j / RP(i+1)^2 + lam_m) x(i) + CM(i) x(i+1) = d(i),
// with AM(1) = CM(M) = 0 (Neumann walls, direct.f:344-352). The singular line (j = 0, null
// x2 mode) is pinned with x(1) = 0, as the reference's IPRES=3 branch pins its singular mode
// (direct.f:849-852); the plain-mean removal of direct.f:895-898 follows in the inverse x2
// transform.
//
// Algorithm (per line, P threads, Q = M/P rows each, Q >= 3): every thread eliminates its
// own chunk with a modified Thomas sweep that expresses each row through the chunk's first
// and last unknowns; the 2P chunk-end unknowns form a tridiagonal system solved by parallel
// cyclic reduction across the P lanes of the line (wave shuffles); each thread then
// back-substitutes its chunk. The line block is staged through LDS so global reads/writes
// are contiguous (lines are x1-contiguous in S).
#include <algorithm>
#include <numeric>

#include "batch_device.h"
#include "kernels.h"
#include "stencil_device.h"

namespace eddy {

// Reciprocal from the hardware estimate and Newton steps (within about an ulp): the pivots
// of the eliminations are far from zero and denormals, so the scaling steps of a correctly
// rounded division buy nothing here.
__device__ __forceinline__ double fast_rcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}
__device__ __forceinline__ float fast_rcp(float x) {
    float r = __builtin_amdgcn_rcpf(x);
    return __builtin_fmaf(r, __builtin_fmaf(-x, r, 1.0f), r);
}

// Lane permute of a 32- or 64-bit value by a precomputed byte address (4 * source lane).
__device__ __forceinline__ float lane_get(float v, int addr) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(v)));
}
__device__ __forceinline__ double lane_get(double v, int addr) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_ds_bpermute(addr, (int)b);
    const int hi = __builtin_amdgcn_ds_bpermute(addr, (int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <typename T>
struct XArgs {
    T* S;
    T* linesum;
    Dims d;
    Tables<T> tb;
    int P, Q, ldw;
    long lbeg, lend;  // line range (j-major lines j*N + m) of this launch
    // periodic lines (the CYC kernels): row 0 couples to x(M-1) through beta (AM(1)) and row
    // M-1 to x(0) through alpha (CM(M)); the tables hold 0 in those two slots
    T alpha, beta;
};

// One block of 256/P lines (logical line block `lb`). CYC: periodic lines, solved as the
// tridiagonal part T' plus a rank-one correction (Sherman-Morrison, the method of the
// reference's CYCLIC, numrec.f:754-781): T = T' + u v^T with u = (g, 0, ..., 0, alpha),
// v = (1, 0, ..., 0, beta/g), g = -b(0); the elimination carries the second right-hand side
// u beside d, and x = y - z (y_0 + beta y_{M-1}/g) / (1 + z_0 + beta z_{M-1}/g). The singular
// line is pinned at x(0) = 0 and solved without the correction (its corner couplings act on
// x(0) = 0).
template <typename T, int QMAX, bool CYC = false>
__device__ __forceinline__ void xsolve_tile(const XArgs<T>& a, int lb) {
    extern __shared__ unsigned char smem_raw[];
    const Dims d = a.d;
    const int P = a.P, Q = a.Q, M = d.M, N = d.N;
    const int LPW = blockDim.x / P;
    const int ldw = a.ldw;  // P*(Q+1): padded line stride in LDS
    T* tile = reinterpret_cast<T*>(smem_raw);
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int lw = tid / P, p = tid - (tid / P) * P;
    // line numbers and element offsets in 32 bits: S holds fewer than 2^29 elements (the plan
    // rejects fields of 4 GiB and more)
    const int nlines = (int)a.lend;
    const int line0 = (int)a.lbeg + lb * LPW;

    // stage the block's lines (contiguous in S) into LDS, several loads in flight per thread;
    // small-integer quotients through a float reciprocal (exact: the dividends stay below 2^14)
    // and 24-bit products; global addresses as 32-bit byte offsets from the block's first line.
    // Elements past the last line read a clamped in-range element and are zeroed afterwards.
    T* const blk = a.S + (size_t)line0 * (size_t)M;
    const int nel = LPW * M;
    const int nvalid = min(nel, (nlines - line0) * M);
    const float rM = 1.0f / (float)M, rQ = 1.0f / (float)Q;
    constexpr int U = 8;
    for (int e0 = tid; e0 < nel; e0 += nthr * U) {
        T v[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
            v[u] = ld_off(blk, (unsigned)sizeof(T) * (unsigned)min(e0 + u * nthr, nvalid - 1));
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = e0 + u * nthr;
            if (e >= nel) continue;
            const int l2 = (int)(((float)e + 0.5f) * rM);
            const int i = e - (int)__umul24(l2, M);
            const int c = (int)(((float)i + 0.5f) * rQ);
            tile[__umul24(l2, ldw) + i + c] = (e < nvalid) ? v[u] : T(0);  // row c*(Q+1) + (i - c*Q)
        }
    }
    __syncthreads();

    const int l = line0 + lw;
    const bool valid = l < nlines;
    const int j = valid ? l / N : 0;
    const int m = valid ? l - j * N : 0;
    const T akj = a.tb.akj[j];
    const T lam = a.tb.lamj ? a.tb.lamj[valid ? l : 0] : a.tb.lam[m];
    const bool pin = valid && j == 0 && m == d.null_m;
    const T* row = tile + lw * ldw + p * (Q + 1);
    // periodic line: the corrected diagonals of rows 0 and M-1 and the second right-hand side
    const bool cyc = CYC && valid && !pin;
    const T g = CYC ? -(a.tb.bm[0] - akj * a.tb.r2inv[0] + lam) : T(0);

    T aa[QMAX], cc[QMAX], dd[QMAX];
    T ee[CYC ? QMAX : 1];
    const int ib = p * Q;
#pragma unroll
    for (int r = 0; r < QMAX; ++r) {
        if (r < Q) {
            const int i = ib + r;
            T av = T(0), cv = T(0), bv = T(1), dv = T(0);  // rows past M pad the line: x = 0
            if (i < M) {
                const int ti = r * P + p;  // tables in thread order (row p*Q + r)
                av = a.tb.am[ti];
                cv = a.tb.cm[ti];
                bv = a.tb.bm[ti] - akj * a.tb.r2inv[ti] + lam;
                dv = row[r];
            }
            if (!valid) {
                av = T(0);
                cv = T(0);
                bv = T(1);
                dv = T(0);
            }
            if (pin && i == 0) {
                av = T(0);
                bv = T(1);
                cv = T(0);
                dv = T(0);
            }
            T ev = T(0);
            if (CYC && cyc) {
                if (i == 0) {
                    bv = bv - g;
                    ev = g;
                }
                if (i == M - 1) {
                    bv = bv - a.alpha * a.beta / g;
                    ev = a.alpha;
                }
            }
            if (r < 2) {
                const T inv = fast_rcp(bv);
                aa[r] = av * inv;
                cc[r] = cv * inv;
                dd[r] = dv * inv;
                if constexpr (CYC) ee[r] = ev * inv;
            } else {
                const T inv = fast_rcp(bv - av * cc[r - 1]);
                dd[r] = (dv - av * dd[r - 1]) * inv;
                if constexpr (CYC) ee[r] = (ev - av * ee[r - 1]) * inv;
                aa[r] = -av * aa[r - 1] * inv;
                cc[r] = cv * inv;
            }
        }
    }
#pragma unroll
    for (int r = QMAX - 3; r >= 1; --r) {
        if (r <= Q - 3) {
            dd[r] -= cc[r] * dd[r + 1];
            if constexpr (CYC) ee[r] -= cc[r] * ee[r + 1];
            aa[r] -= cc[r] * aa[r + 1];
            cc[r] = -cc[r] * cc[r + 1];
        }
    }
    {
        const T inv = fast_rcp(T(1) - cc[0] * aa[1]);
        dd[0] = (dd[0] - cc[0] * dd[1]) * inv;
        if constexpr (CYC) ee[0] = (ee[0] - cc[0] * ee[1]) * inv;
        aa[0] = aa[0] * inv;
        cc[0] = -cc[0] * cc[1] * inv;
    }
    // reduced system: rows 2p (chunk first) and 2p+1 (chunk last)
    T lastA = T(0), lastC = T(0), lastD = T(0), lastE = T(0);
#pragma unroll
    for (int r = 0; r < QMAX; ++r)
        if (r == Q - 1) {
            lastA = aa[r];
            lastC = cc[r];
            lastD = dd[r];
            if constexpr (CYC) lastE = ee[r];
        }
    // The 2P chunk-end unknowns y (y_2p = chunk p's first row, y_2p+1 its last) form a
    // tridiagonal system with unit diagonal: lane p holds rows 2p (a0, c0, d0) and 2p+1
    // (a1, c1, d1). One cyclic-reduction step removes the odd unknowns from the even rows;
    // the P even unknowns are then solved by parallel cyclic reduction across the lanes
    // (wave shuffles), and each odd unknown follows from its own row:
    // y_2p+1 = d1 - a1 y_2p - c1 y_2p+2. Rows at a line end have the coupling to the missing
    // side at exactly 0 (AM(1) = CM(M) = 0 at the walls, identity padding rows), so the
    // neighbours fetched across a line end (wrapping inside the wave) are finite values
    // multiplied by 0.
    T a0 = aa[0], c0 = cc[0], d0 = dd[0], e0 = CYC ? ee[0] : T(0);
    const T a1 = lastA, c1 = lastC, d1 = lastD, e1 = lastE;
    const int lane = tid & 63;
    {  // even row 2p minus a0 * row 2p-1 (lane p-1's odd row) minus c0 * row 2p+1 (own odd row)
        const int up = ((lane - 1) & 63) << 2;
        const T ua = lane_get(a1, up), uc = lane_get(c1, up), ud = lane_get(d1, up);
        const T i0 = fast_rcp(T(1) - a0 * uc - c0 * a1);
        const T na0 = -a0 * ua * i0, nc0 = -c0 * c1 * i0, nd0 = (d0 - a0 * ud - c0 * d1) * i0;
        if constexpr (CYC) e0 = (e0 - a0 * lane_get(e1, up) - c0 * e1) * i0;
        a0 = na0;
        c0 = nc0;
        d0 = nd0;
    }
    for (int dl = 1; dl < P; dl <<= 1) {  // even rows at distance dl lanes
        const int up = ((lane - dl) & 63) << 2, dn = ((lane + dl) & 63) << 2;
        const T ua = lane_get(a0, up), uc = lane_get(c0, up), ud = lane_get(d0, up);
        const T va = lane_get(a0, dn), vc = lane_get(c0, dn), vd = lane_get(d0, dn);
        const T i0 = fast_rcp(T(1) - a0 * uc - c0 * va);
        const T na0 = -a0 * ua * i0, nc0 = -c0 * vc * i0, nd0 = (d0 - a0 * ud - c0 * vd) * i0;
        if constexpr (CYC) e0 = (e0 - a0 * lane_get(e0, up) - c0 * lane_get(e0, dn)) * i0;
        a0 = na0;
        c0 = nc0;
        d0 = nd0;
    }
    const T x0 = d0;
    const T xl = d1 - a1 * x0 - c1 * lane_get(x0, ((lane + 1) & 63) << 2);
    // periodic line: the correction factor from y_0, z_0 (lane p = 0, row 0) and y_{M-1},
    // z_{M-1} (lane (M-1)/Q, row (M-1) mod Q) of the line's lanes
    T fact = T(0), z0 = T(0), zl = T(0);
    if constexpr (CYC) {
        z0 = e0;
        zl = e1 - a1 * z0 - c1 * lane_get(z0, ((lane + 1) & 63) << 2);
        const int pl1 = (M - 1) / Q, rl1 = (M - 1) - pl1 * Q;
        T yM = T(0), zM = T(0);
#pragma unroll
        for (int r = 0; r < QMAX; ++r)
            if (r == rl1) {
                yM = (r == 0) ? x0 : ((r == Q - 1) ? xl : dd[r] - aa[r] * x0 - cc[r] * xl);
                zM = (r == 0) ? z0 : ((r == Q - 1) ? zl : ee[r] - aa[r] * z0 - cc[r] * zl);
            }
        const int first = lane - p;  // the line's lanes are consecutive within the wave
        const T y0l = lane_get(x0, first << 2), z0l = lane_get(z0, first << 2);
        yM = lane_get(yM, (first + pl1) << 2);
        zM = lane_get(zM, (first + pl1) << 2);
        if (cyc) fact = (y0l + a.beta * yM / g) / (T(1) + z0l + a.beta * zM / g);
    }
    // back-substitute into LDS (own rows only), accumulate the line sum
    T* wrow = tile + lw * ldw + p * (Q + 1);
    T part = T(0);
#pragma unroll
    for (int r = 0; r < QMAX; ++r) {
        if (r < Q) {
            T x;
            if (r == 0) x = x0;
            else if (r == Q - 1) x = xl;
            else x = dd[r] - aa[r] * x0 - cc[r] * xl;
            if constexpr (CYC) {
                T z;
                if (r == 0) z = z0;
                else if (r == Q - 1) z = zl;
                else z = ee[r] - aa[r] * z0 - cc[r] * zl;
                x -= fact * z;
            }
            wrow[r] = x;
            part += x;
        }
    }
    if (j == 0) {
        // reduce over the P lanes of this line (consecutive lanes of one wave, P <= 64)
        for (int off = P >> 1; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
        if (valid && p == 0) a.linesum[m] = part;
    }
    __syncthreads();
    for (int e = tid; e < nvalid; e += nthr) {
        const int l2 = (int)(((float)e + 0.5f) * rM);
        const int i = e - (int)__umul24(l2, M);
        const int c = (int)(((float)i + 0.5f) * rQ);
        st_off(blk, (unsigned)sizeof(T) * (unsigned)e, tile[__umul24(l2, ldw) + i + c]);
    }
}

template <typename T, int QMAX, bool CYC>
__global__ void __launch_bounds__(256) k_xsolve(XArgs<T> a) {
    xsolve_tile<T, QMAX, CYC>(a, xcd_tile(blockIdx.x, gridDim.x));
}

// Pick P (threads per line, a power of two <= 64) and Q (rows per thread, 3 <= Q <= qmax):
// an exact split M = P*Q when one exists, else Q = ceil(M/P) with the line padded by
// P*Q - M identity rows (CM(M) = 0 at the wall decouples them, so they solve to 0). Lines
// up to 64 x 32 = 2048 rows.
static bool pick_pq(int M, int qmax, int& P, int& Q) {
    for (P = 1; P <= 64; P <<= 1) {
        if (M % P) continue;
        Q = M / P;
        if (Q >= 3 && Q <= qmax) return true;
    }
    for (P = 1; P <= 64; P <<= 1) {
        Q = std::max(3, (M + P - 1) / P);
        if (Q <= qmax) return true;
    }
    return false;
}

static bool pick_pq_any(int M, int chunk, int& P, int& Q) {
    // the largest kernel template holds kXsolveMaxChunk rows per thread
    const int qmax = chunk > 0 ? std::min(chunk, kXsolveMaxChunk) : 16;
    return pick_pq(M, qmax, P, Q) || pick_pq(M, 16, P, Q) || pick_pq(M, 32, P, Q);
}

bool xsolve_partition(int M, int chunk, int& P, int& Q) { return pick_pq_any(M, chunk, P, Q); }

template <typename T, bool CYC>
static void launch_xsolve_q(hipStream_t st, dim3 grid, size_t sh, const XArgs<T>& a) {
    auto go = [&](auto fn) {
        (void)hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
        hipLaunchKernelGGL(fn, grid, dim3(256), sh, st, a);
    };
    if (a.Q <= 4) go(k_xsolve<T, 4, CYC>);
    else if (a.Q <= 8) go(k_xsolve<T, 8, CYC>);
    else if (a.Q <= 16) go(k_xsolve<T, 16, CYC>);
    else go(k_xsolve<T, 32, CYC>);
}

template <typename T>
int launch_xsolve(hipStream_t st, const Dims& d, const Tables<T>& tb, T* S, T* linesum, int chunk, int j0, int nj,
                  int cyc, double alpha, double beta) {
    XArgs<T> a{};
    a.S = S;
    a.linesum = linesum;
    a.d = d;
    a.tb = tb;
    a.alpha = T(alpha);
    a.beta = T(beta);
    int P = 0, Q = 0;
    if (!pick_pq_any(d.M, chunk, P, Q)) return -1;
    a.P = P;
    a.Q = Q;
    a.ldw = P * (a.Q + 1);
    const int LPW = 256 / P;
    a.lbeg = (long)j0 * d.N;
    a.lend = (long)(j0 + nj) * d.N;
    const long nlines = a.lend - a.lbeg;
    dim3 grid((unsigned)((nlines + LPW - 1) / LPW));
    size_t sh = (size_t)LPW * a.ldw * sizeof(T);
    if (cyc) launch_xsolve_q<T, true>(st, grid, sh, a);
    else launch_xsolve_q<T, false>(st, grid, sh, a);
    return 0;
}

// ------------------------------------------------------------------ batched (one item per workgroup)
// One kernel per rows-per-thread template: a kernel holding every template would take the
// registers of the largest (QMAX = 32) for all problems.
template <typename T, int QMAX>
// the 16-row class at no more than 128 VGPRs in FP64 (4 workgroups per CU; 130 unbounded) and
// 80 in FP32 (6 workgroups per CU; 83 unbounded)
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(QMAX == 16 ? (sizeof(T) == 4 ? 6 : 4) : 1)))
k_batch_xsolve(const Job<XArgs<T>>* jobs, int njobs, int total) {
    run_item(jobs, njobs, total, [](const Job<XArgs<T>>& J, int t) { xsolve_tile<T, QMAX>(J.a, t); });
}


static int qmax_class(int Q) { return Q <= 4 ? 4 : (Q <= 8 ? 8 : (Q <= 16 ? 16 : 32)); }

template <typename T>
int prepare_xsolve(const std::vector<BatchProblem<T>>& pr, StageLaunch& out, std::vector<void*>& owned) {
    struct Item {
        int idx, P, Q, cls;
    };
    std::vector<Item> items;
    for (int idx = 0; idx < (int)pr.size(); ++idx) {
        int P = 0, Q = 0;
        if (!pick_pq_any(pr[idx].d.M, pr[idx].chunk, P, Q)) return -2;
        items.push_back({idx, P, Q, qmax_class(Q)});
    }
    auto lds_of = [&](const Item& it) { return (size_t)(256 / it.P) * it.P * (it.Q + 1) * sizeof(T); };
    // grouped by template class (largest first); the tiles of most LDS first, then the
    // longest x1 lines
    std::stable_sort(items.begin(), items.end(), [&](const Item& x, const Item& y) {
        if (x.cls != y.cls) return x.cls > y.cls;
        const int ox = lds_occupancy(lds_of(x)), oy = lds_occupancy(lds_of(y));
        return ox != oy ? ox < oy : pr[x.idx].d.M > pr[y.idx].d.M;
    });
    std::vector<Job<XArgs<T>>> jobs;
    out = StageLaunch{};
    for (const Item& it : items) {
        const BatchProblem<T>& p = pr[it.idx];
        Job<XArgs<T>> j{};
        XArgs<T>& a = j.a;
        a.S = p.S;
        a.linesum = p.linesum;
        a.d = p.d;
        a.tb = p.tb;
        a.P = it.P;
        a.Q = it.Q;
        a.ldw = it.P * (a.Q + 1);
        a.lbeg = 0;
        a.lend = (long)p.d.L * p.d.N;
        const int LPW = 256 / it.P;
        j.ntx = 1;
        j.ntiles = (int)(((long)p.d.L * p.d.N + LPW - 1) / LPW);
        j.aux = it.cls;
        if (!add_class_job(out, jobs, j, it.cls, lds_of(it))) return -1;
    }
    return upload_jobs(jobs, out, owned);
}

template <typename T, int QMAX>
static void batch_xsolve_class(hipStream_t st, const StageLaunch& sl, int c) {
    auto fn = k_batch_xsolve<T, QMAX>;
    const size_t lds = sl.cls_lds[c];
    (void)hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    const int total = sl.cls_total[c];
    const Job<XArgs<T>>* jobs = (const Job<XArgs<T>>*)sl.jobs + sl.cls_job[c];
    hipLaunchKernelGGL(fn, dim3(std::max(total, 1)), dim3(256), lds, st, jobs, sl.cls_job[c + 1] - sl.cls_job[c], total);
}

template <typename T>
void batch_xsolve(hipStream_t st, const StageLaunch& sl) {
    for (int c = 0; c < sl.ncls; ++c) {
        switch (sl.cls_q[c]) {
            case 4: batch_xsolve_class<T, 4>(st, sl, c); break;
            case 8: batch_xsolve_class<T, 8>(st, sl, c); break;
            case 16: batch_xsolve_class<T, 16>(st, sl, c); break;
            default: batch_xsolve_class<T, 32>(st, sl, c); break;
        }
    }
}

template int prepare_xsolve<double>(const std::vector<BatchProblem<double>>&, StageLaunch&, std::vector<void*>&);
template int prepare_xsolve<float>(const std::vector<BatchProblem<float>>&, StageLaunch&, std::vector<void*>&);
template void batch_xsolve<double>(hipStream_t, const StageLaunch&);
template void batch_xsolve<float>(hipStream_t, const StageLaunch&);

template int launch_xsolve<double>(hipStream_t, const Dims&, const Tables<double>&, double*, double*, int, int, int,
                                   int, double, double);
template int launch_xsolve<float>(hipStream_t, const Dims&, const Tables<float>&, float*, float*, int, int, int, int,
                                  double, double);

}  // namespace eddy
