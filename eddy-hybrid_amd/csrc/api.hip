// api.hip -- plan construction, the solve pipelines and the C-ABI of libeddyproj.so.
//
// Pipeline of one FP64 solve (DIRECT, direct.f:590-1106), five HBM sweeps over the field:
//   1. y-FFT forward of every interior pencil            DP      -> S   (direct.f:620-629)
//   2. x2 transform (DCT-II or dense eigenvectors)       S       -> S
//   3. x1 tridiagonal solve of every (j, m) line          S       -> S   (direct.f:788-797)
//   4. x2 inverse transform (+ zero-mode mean removal)   S       -> S   (direct.f:894-898)
//   5. y-FFT inverse x 1/L, ghost fill                   S       -> DP  (direct.f:1035-1106)
// The MPI transposes of direct.f:651/989 have no counterpart on one GPU: the spectral array
// is laid out so that every sweep reads and writes rows of contiguous x1 values.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "batch_device.h"
#include "fft_device.h"
#include "host_eigen.h"
#include "host_stage.h"
#include "plan.h"
#include "reduce_device.h"

namespace eddy {

static thread_local std::string g_err = "no error";

void set_error(const std::string& s) { g_err = s; }

int device_error(hipError_t e, const char* what) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return EDDY_ERR_DEVICE;
}

#define HIP_TRY(call)                                             \
    do {                                                          \
        hipError_t e_ = (call);                                   \
        if (e_ != hipSuccess) return device_error(e_, #call);     \
    } while (0)

// ------------------------------------------------------------------ host numerics
// Radix plan of a length-n complex FFT: the power-of-two part in radices up to `maxr` (8, 4
// or 2), avoiding a lone radix 2 (2^(3a+1) = 8^(a-1) * 4 * 4), then 3 and 5. The smallest
// radix bounds the butterflies per pass of the in-place FFT, so it is kept large.
static bool radix_plan(int n, FftPlan& fp, int maxr = 8) {
    fp.n = n;
    fp.count = 0;
    int r = n, m = 0;
    while (r % 2 == 0) {
        r /= 2;
        ++m;
    }
    std::vector<int> pw;
    if (maxr >= 8) {
        int a = m / 3, b = m % 3;
        if (b == 1 && a >= 1) {
            --a;
            b = 4;  // 8 * 2 -> 4 * 4
        }
        for (int i = 0; i < a; ++i) pw.push_back(8);
        if (b == 2) pw.push_back(4);
        else if (b == 4) pw.insert(pw.end(), {4, 4});
        else if (b == 1) pw.push_back(2);
    } else if (maxr == 4) {
        for (int i = 0; i < m / 2; ++i) pw.push_back(4);
        if (m % 2) pw.push_back(2);
    } else {
        for (int i = 0; i < m; ++i) pw.push_back(2);
    }
    for (int f : pw) {
        if (fp.count >= kMaxRadices) return false;
        fp.rad[fp.count++] = f;
    }
    for (int f : {3, 5}) {
        while (r % f == 0 && r > 1) {
            if (fp.count >= kMaxRadices) return false;
            fp.rad[fp.count++] = f;
            r /= f;
        }
    }
    return r == 1;
}

template <typename T>
static int upload(DevTables<T>& dt, const std::vector<double>& h, const T** out) {
    if (h.empty()) {
        *out = nullptr;
        return 0;
    }
    std::vector<T> c(h.begin(), h.end());
    void* p = nullptr;
    HIP_TRY(hipMalloc(&p, c.size() * sizeof(T)));
    dt.owned.push_back(p);
    HIP_TRY(hipMemcpy(p, c.data(), c.size() * sizeof(T), hipMemcpyHostToDevice));
    *out = static_cast<const T*>(p);
    return 0;
}

// A table of L slots of `per` values each, slot j holding by_kw[(j + 1) / 2] (the y-mode
// slots 2k-1 and 2k share wavenumber k), converted to T one wavenumber at a time.
template <typename T>
static int upload_slots(DevTables<T>& dt, const std::vector<std::vector<double>>& by_kw, int L, size_t per,
                        const T** out) {
    void* p = nullptr;
    HIP_TRY(hipMalloc(&p, (size_t)L * per * sizeof(T)));
    dt.owned.push_back(p);
    T* base = static_cast<T*>(p);
    std::vector<T> c(per);
    for (size_t kw = 0; kw < by_kw.size(); ++kw) {
        std::copy(by_kw[kw].begin(), by_kw[kw].end(), c.begin());
        for (int j = (kw == 0) ? 0 : 2 * (int)kw - 1; j <= 2 * (int)kw && j < L; ++j)
            HIP_TRY(hipMemcpy(base + (size_t)j * per, c.data(), per * sizeof(T), hipMemcpyHostToDevice));
    }
    *out = base;
    return 0;
}

static int upload64(eddy_plan_s* P, const std::vector<double>& h, const double** out) {
    void* p = nullptr;
    HIP_TRY(hipMalloc(&p, std::max<size_t>(1, h.size()) * sizeof(double)));
    P->owned.push_back(p);
    if (!h.empty()) HIP_TRY(hipMemcpy(p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    *out = static_cast<const double*>(p);
    return 0;
}

// Host copies of every table, in FP64; uploaded once per precision.
struct HostTables {
    std::vector<double> am, bm, cm, r2inv, akj, lam, wsum, qf, qiT;
    // the x1 coefficients in the x1 solve's thread order: row p*Q + r of the line at r*P + p
    // (P threads per line, Q rows each; rows past M are padding), so a wave's loads of one
    // row index are contiguous
    std::vector<double> am_x, bm_x, cm_x, r2inv_x;
    std::vector<double> ytwc_re, ytwc_im, ytwr_re, ytwr_im;
    std::vector<double> ztwc_re, ztwc_im, ztwr_re, ztwr_im, zdct_re, zdct_im;
};

template <typename T>
static int upload_tables(DevTables<T>& dt, const HostTables& h) {
    int rc = 0;
    Tables<T>& t = dt.t;
    if ((rc = upload(dt, h.am_x, &t.am)) || (rc = upload(dt, h.bm_x, &t.bm)) || (rc = upload(dt, h.cm_x, &t.cm)) ||
        (rc = upload(dt, h.r2inv_x, &t.r2inv)) || (rc = upload(dt, h.akj, &t.akj)) || (rc = upload(dt, h.lam, &t.lam)) ||
        (rc = upload(dt, h.wsum, &t.wsum)) || (rc = upload(dt, h.qf, &t.qf)) || (rc = upload(dt, h.qiT, &t.qiT)) ||
        (rc = upload(dt, h.ytwc_re, &t.ytwc_re)) || (rc = upload(dt, h.ytwc_im, &t.ytwc_im)) ||
        (rc = upload(dt, h.ytwr_re, &t.ytwr_re)) || (rc = upload(dt, h.ytwr_im, &t.ytwr_im)) ||
        (rc = upload(dt, h.ztwc_re, &t.ztwc_re)) || (rc = upload(dt, h.ztwc_im, &t.ztwc_im)) ||
        (rc = upload(dt, h.ztwr_re, &t.ztwr_re)) || (rc = upload(dt, h.ztwr_im, &t.ztwr_im)) ||
        (rc = upload(dt, h.zdct_re, &t.zdct_re)) || (rc = upload(dt, h.zdct_im, &t.zdct_im)))
        return rc;
    return 0;
}

// exp(-2 pi i t / n) evaluated per entry from the reduced angle
static void twiddles(int n, int count, std::vector<double>& re, std::vector<double>& im) {
    re.resize(count);
    im.resize(count);
    const double tpi = 2.0 * M_PI;
    for (int t = 0; t < count; ++t) {
        const double a = tpi * (double)(t % n) / (double)n;
        re[t] = std::cos(a);
        im[t] = -std::sin(a);
    }
}

static void free_plan(eddy_plan_s* P) {
    if (!P) return;
    if (P->blas) {
        (void)hipStreamSynchronize(P->stream);
        zgemm_destroy(P->blas);
    }
    for (void* p : P->owned) (void)hipFree(p);
    for (void* p : P->t64.owned) (void)hipFree(p);
    for (void* p : P->t32.owned) (void)hipFree(p);
    for (void* p : P->tz64.owned) (void)hipFree(p);
    for (void* p : P->tz32.owned) (void)hipFree(p);
    for (auto& pt : P->pending) {
        (void)hipEventDestroy(pt.start);
        (void)hipEventDestroy(pt.stop);
    }
    for (auto e : P->event_pool) (void)hipEventDestroy(e);
    delete P;
}

// ------------------------------------------------------------------ timing
static hipEvent_t take_event(eddy_plan_s* P) {
    if (!P->event_pool.empty()) {
        hipEvent_t e = P->event_pool.back();
        P->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

struct KTimer {
    eddy_plan_s* P;
    int id;
    hipEvent_t a = nullptr;
    KTimer(eddy_plan_s* p, int k) : P(p), id(k) {
        if (P->timing) {
            a = take_event(P);
            (void)hipEventRecord(a, P->stream);
        }
    }
    ~KTimer() {
        if (P->timing) {
            hipEvent_t b = take_event(P);
            (void)hipEventRecord(b, P->stream);
            P->pending.push_back({id, a, b});
        }
    }
};

static void drain_timing(eddy_plan_s* P) {
    if (P->pending.empty()) return;
    (void)hipStreamSynchronize(P->stream);
    for (auto& pt : P->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, pt.start, pt.stop) == hipSuccess) {
            P->slots[pt.kernel].launches += 1;
            P->slots[pt.kernel].total_ms += ms;
        }
        P->event_pool.push_back(pt.start);
        P->event_pool.push_back(pt.stop);
    }
    P->pending.clear();
}

// ------------------------------------------------------------------ pipelines
template <typename T>
static const Tables<T>& tables(eddy_plan_s* P);
template <>
const Tables<double>& tables<double>(eddy_plan_s* P) { return P->t64.t; }
template <>
const Tables<float>& tables<float>(eddy_plan_s* P) { return P->t32.t; }
// the x2 line tables of the x1 GEMM form
template <typename T>
static const Tables<T>& xtables(eddy_plan_s* P);
template <>
const Tables<double>& xtables<double>(eddy_plan_s* P) { return P->tz64.t; }
template <>
const Tables<float>& xtables<float>(eddy_plan_s* P) { return P->tz32.t; }

static Reduce rd(eddy_plan_s* P) { return Reduce{P->part, P->red}; }

static int check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return device_error(e, what);
    return 0;
}

static FwdInput direct_input(const double* src, int shift) {
    FwdInput in{};
    in.mode = YF_DIRECT;
    in.src = src;
    in.shift = shift;
    return in;
}

// forward y-FFT of the input `in` into S, x2 transform (zero y-mode shifted by
// L*red[VOLB]/red[VOL] when zshift != 0), x1 solve, inverse x2 transform.
template <typename T>
static int solve_core(eddy_plan_s* P, const FwdInput& in, int zshift) {
    const Tables<T>& tb = tables<T>(P);
    T* S = reinterpret_cast<T*>(P->S);
    T* S2 = reinterpret_cast<T*>(P->S2);
    T* ls = reinterpret_cast<T*>(P->linesum);
    {
        KTimer kt(P, EDDY_K_YFFT_FWD);
        launch_yfft_fwd<T>(P->stream, P->d, P->fpy, tb, in, P->op, rd(P), S, P->tune.yfft_pencils,
                           P->tune.yfft_iblock);
    }
    // The x2 transform, x1 solve and inverse x2 transform of a y-mode touch only that mode's
    // slab of S; running the three over one chunk of modes at a time keeps the chunk's slabs
    // in the memory-side cache between the sweeps (tuning slab_chunks).
    const int L = P->d.L;
    const int nch = std::max(1, std::min(P->tune.slab_chunks, L));
    for (int c = 0; c < nch; ++c) {
        const int j0 = (int)((long)L * c / nch), j1 = (int)((long)L * (c + 1) / nch);
        if (j1 <= j0) continue;
        if (P->xgemm) {
            // x1 form: x1 eigen-transform (products), x2 line solves on S2, inverse products
            const Tables<T>& tz = xtables<T>(P);
            {
                KTimer kt(P, EDDY_K_ZTRANS_FWD);
                if (launch_xgemm_fwd<T>(P->stream, P->blas, P->d, tz, S, S2, zshift ? P->red : nullptr, j0,
                                        j1 - j0)) {
                    set_error("the forward x1 transform products failed to launch");
                    return EDDY_ERR_DEVICE;
                }
            }
            {
                KTimer kt(P, EDDY_K_XSOLVE);
                if (launch_xsolve<T>(P->stream, P->dz, tz, S2, ls, P->tune.xsolve_chunk, j0, j1 - j0, P->d.cyc_z,
                                     P->zalpha, P->zbeta)) {
                    set_error("x2 line solve: no thread partition for N");
                    return EDDY_ERR_UNSUPPORTED;
                }
            }
            {
                KTimer kt(P, EDDY_K_ZTRANS_INV);
                if (launch_xgemm_inv<T>(P->stream, P->blas, P->d, tz, S2, S, ls, P->dz.null_m, j0, j1 - j0)) {
                    set_error("the inverse x1 transform products failed to launch");
                    return EDDY_ERR_DEVICE;
                }
            }
            continue;
        }
        // the GEMM x2 transform works out of place: the x1 solve then runs on S2
        T* Sx = P->zgemm ? S2 : S;
        {
            KTimer kt(P, EDDY_K_ZTRANS_FWD);
            if (P->zgemm) {
                if (launch_zgemm_fwd<T>(P->stream, P->blas, P->d, tb, S, S2, zshift ? P->red : nullptr, j0, j1 - j0)) {
                    set_error("the forward x2 transform products failed to launch");
                    return EDDY_ERR_DEVICE;
                }
            } else {
                launch_ztrans_fwd<T>(P->stream, P->d, P->fpz, tb, S, P->tune.ztrans_pencils, P->dense,
                                     zshift ? P->red : nullptr, j0, j1 - j0);
            }
        }
        {
            KTimer kt(P, EDDY_K_XSOLVE);
            if (launch_xsolve<T>(P->stream, P->d, tb, Sx, ls, P->tune.xsolve_chunk, j0, j1 - j0, P->d.cyc_x,
                                 P->xalpha, P->xbeta)) {
                set_error("x1 solve: no thread partition for M");
                return EDDY_ERR_UNSUPPORTED;
            }
        }
        {
            KTimer kt(P, EDDY_K_ZTRANS_INV);
            if (P->zgemm) {
                if (launch_zgemm_inv<T>(P->stream, P->blas, P->d, tb, S2, S, ls, j0, j1 - j0)) {
                    set_error("the inverse x2 transform products failed to launch");
                    return EDDY_ERR_DEVICE;
                }
            } else {
                launch_ztrans_inv<T>(P->stream, P->d, P->fpz, tb, S, ls, P->tune.ztrans_pencils, P->dense, j0,
                                     j1 - j0);
            }
        }
    }
    return check_launch("solve kernels");
}

static int finish_direct64(eddy_plan_s* P, double* dp) {
    {
        KTimer kt(P, EDDY_K_YFFT_INV);
        launch_yfft_inv<double>(P->stream, P->d, P->fpy, P->t64.t, P->S, dp, rd(P), P->tune.yfft_pencils, YI_DIRECT);
    }
    return check_launch("y-FFT inverse");
}

static int run_direct64(eddy_plan_s* P, double* dp, int shift) {
    int rc = solve_core<double>(P, direct_input(dp, shift), 0);
    if (rc) return rc;
    return finish_direct64(P, dp);
}

static size_t field_elems(const eddy_plan_s* P) { return (size_t)P->d.nx * P->d.ny * P->d.nz; }

// FP64 DIRECT of the plan with caller-provided work buffers in place of the plan's own (the
// spectral array, the line sums, the reduction partials and scalars); nothing is allocated.
// The GEMM forms are excluded (their second spectral array and library workspace).
int direct64_with_buffers(eddy_plan_s* P, double* dp, const WorkBuffers& wb) {
    if (P->zgemm || P->xgemm) {
        set_error("the x2 / x1 transform products (GEMM forms) need the plan's own buffers");
        return EDDY_ERR_UNSUPPORTED;
    }
    double* save[4] = {P->S, P->linesum, P->part, P->red};
    const int timing = P->timing;
    P->S = wb.S;
    P->linesum = wb.linesum;
    P->part = wb.part;
    P->red = wb.red;
    P->timing = 0;  // the timing events are created on demand
    const int rc = run_direct64(P, dp, 0);
    P->S = save[0];
    P->linesum = save[1];
    P->part = save[2];
    P->red = save[3];
    P->timing = timing;
    return rc;
}

size_t direct64_work_doubles(const eddy_plan_s* P) {
    const size_t M = P->d.M, L = P->d.L, N = P->d.N;
    return M * L * N + std::max<size_t>(N, 1) + 2 * M * N + RED_COUNT;
}

// An RHS is compatible with the Neumann walls when its volume-weighted mean is zero; a
// divergence formed from velocities with zero normal wall flux meets that to roundoff.
constexpr double kCompatTol = 1e-10;

// |sum vol*b| / sum vol*|b| over the finite interior values of b, and the count of non-finite
// ones (synchronises the plan's stream).
static int rhs_check(eddy_plan_s* P, const double* b, double* rel, long* nonfinite) {
    launch_rhs_check(P->stream, P->d, P->mt, b, rd(P));
    double h[3] = {0.0, 0.0, 0.0};
    HIP_TRY(hipMemcpyAsync(h, P->red + RED_CHK_VB, sizeof h, hipMemcpyDeviceToHost, P->stream));
    HIP_TRY(hipStreamSynchronize(P->stream));
    *rel = h[1] > 0.0 ? std::fabs(h[0]) / h[1] : 0.0;
    *nonfinite = (long)h[2];
    return check_launch("RHS check");
}

static int ensure_rhs(eddy_plan_s* P) {
    if (P->rhs) return 0;
    void* p = nullptr;
    HIP_TRY(hipMalloc(&p, field_elems(P) * sizeof(double)));
    P->owned.push_back(p);
    P->rhs = static_cast<double*>(p);
    P->bytes += field_elems(P) * sizeof(double);
    return 0;
}

// FP32 solves + FP64 defect correction. dp: pressure out; the RHS is dp on entry (minus the
// solvability shift red[VOLB]/red[VOL] when shift != 0), or, when div_in is given, the
// divergence that input describes, formed inside the first FP32 forward transform, which
// also stores it in FP64 as the correction's RHS (unshifted: the shift is then subtracted
// in the zero y-mode of the first solve and on the fly in every residual). Without a requested
// residual and with at least one sweep, the closing mean removal and ghost fill are fused into
// the last update (the mean of the previous iterate: the last correction solve removes its
// own zero-mode mean, so its contribution is far below the FP64 rounding of phi).
static int run_refined(eddy_plan_s* P, double* dp, int sweeps, int want_resid, int shift,
                       const FwdInput* div_in = nullptr) {
    int rc = ensure_rhs(P);
    if (rc) return rc;
    const bool fused_final = !want_resid && sweeps >= 1 && !P->tune.yfft_gemm;
    int res_shift = 0;  // whether the residuals subtract the shift (RHS stored unshifted)
    if (div_in) {
        {
            KTimer kt(P, EDDY_K_OTHER);
            launch_fill(P->stream, P->red, RED_COUNT, 0.0);
        }
        FwdInput in = *div_in;
        in.rhs_out = P->rhs;
        rc = solve_core<float>(P, in, shift);
        res_shift = shift;
    } else {
        if (shift) launch_copy_shift(P->stream, P->d, dp, P->rhs, P->red);
        else HIP_TRY(hipMemcpyAsync(P->rhs, dp, field_elems(P) * sizeof(double), hipMemcpyDeviceToDevice, P->stream));
        {
            KTimer kt(P, EDDY_K_OTHER);
            launch_fill(P->stream, P->red, RED_COUNT, 0.0);
            if (want_resid) launch_sumsq(P->stream, P->d, P->rhs, rd(P), RED_BNORM2);
        }
        rc = solve_core<float>(P, direct_input(P->rhs, 0), 0);
    }
    if (rc) return rc;
    {
        KTimer kt(P, EDDY_K_YFFT_INV);
        launch_yfft_inv<float>(P->stream, P->d, P->fpy, P->t32.t, reinterpret_cast<float*>(P->S), dp, rd(P),
                               P->tune.yfft_pencils, YI_STORE);
    }
    for (int s = 0; s < sweeps; ++s) {
        const bool last_fused = fused_final && s == sweeps - 1;
        {
            KTimer kt(P, EDDY_K_OTHER);
            launch_fill(P->stream, P->red + RED_RNORM2, 1, 0.0);
            // the fused last update reads the previous pass's sum of phi
            if (!last_fused) launch_fill(P->stream, P->red + RED_PHISUM, 1, 0.0);
        }
        FwdInput in{};
        in.mode = YF_RESIDUAL;
        in.src = P->rhs;
        in.phi = dp;
        in.shift = res_shift;
        rc = solve_core<float>(P, in, 0);
        if (rc) return rc;
        {
            KTimer kt(P, EDDY_K_YFFT_INV);
            launch_yfft_inv<float>(P->stream, P->d, P->fpy, P->t32.t, reinterpret_cast<float*>(P->S), dp, rd(P),
                                   P->tune.yfft_pencils, last_fused ? YI_UPDATE_FINAL : YI_UPDATE);
        }
    }
    if (!fused_final) {
        KTimer kt(P, EDDY_K_OTHER);
        if (want_resid) launch_fill(P->stream, P->red + RED_RNORM2, 1, 0.0);
        launch_finalize(P->stream, P->d, P->op, dp, want_resid ? P->rhs : nullptr, rd(P));
    }
    return check_launch("refinement");
}

static int read_resid(eddy_plan_s* P, double* out) {
    double h[RED_COUNT];
    HIP_TRY(hipMemcpyAsync(h, P->red, sizeof(h), hipMemcpyDeviceToHost, P->stream));
    HIP_TRY(hipStreamSynchronize(P->stream));
    *out = (h[RED_BNORM2] > 0) ? std::sqrt(h[RED_RNORM2] / h[RED_BNORM2]) : std::sqrt(h[RED_RNORM2]);
    return 0;
}

// ------------------------------------------------------------------ plan creation
static int validate_dims(int nx, int ny, int nz, int mp, int lp, int np) {
    if (nx < 4 || ny < 4 || nz < 5) {
        set_error("dimensions too small (need NX>=4, NY>=4, NZ>=5)");
        return EDDY_ERR_ARG;
    }
    if ((unsigned long long)nx * ny * nz * sizeof(double) >= (1ull << 32)) {
        // the kernels address the FP64 fields with 32-bit byte offsets
        set_error("field too large: NX*NY*NZ*8 bytes must stay below 4 GiB");
        return EDDY_ERR_UNSUPPORTED;
    }
    if ((mp != 0 && mp != 1) || (lp != 0 && lp != 1) || (np != 0 && np != 1)) {
        set_error("boundary flags MP/LP/NP must be 0 or 1");
        return EDDY_ERR_ARG;
    }
    return 0;
}

// The x1 form of a stretched-x2 plan. For every distinct modified wavenumber AK (slots 2k-1
// and 2k share one, direct.f:321-331) the x1 operator of that y-mode,
//   T(i,i-1) = AM(i), T(i,i) = BM(i) - AK/RP(i+1)^2, T(i,i+1) = CM(i),
// is symmetric under the weights w(i) = RP(i+1)/AP(i+1) (w(i) CM(i) = w(i+1) AM(i+1) =
// AU(i+1) RU(i+1)), so H = W^{1/2} T W^{-1/2} is symmetric tridiagonal with off-diagonal
// CM(i) sqrt(w(i)/w(i+1)); H = U diag(mu) U^T gives QF(i,m) = U(i,m) sqrt(w(i)) (forward,
// X = V^{-1} x) and QI(i,m) = U(i,m)/sqrt(w(i)) (inverse, x = V X). The zero wavenumber's null
// pair is set exactly: mu = 0, U(:,null) ~ sqrt(w) (T has the constant null vector). The x2
// lines are then solved as tridiagonal lines (AN, BN + mu, CN) by the line-solve kernel:
// tables in its thread order for line length N, the shift per line (slot j, x1 mode m).
static int build_x1_basis(eddy_plan_s* P, int M, int L, int N, const double* ap, const double* rp,
                          const HostTables& h, const std::vector<double>& an, const std::vector<double>& bn,
                          const std::vector<double>& cn) {
    std::vector<double> w(M), sw(M);
    for (int i = 0; i < M; ++i) {
        w[i] = rp[i + 1] / ap[i + 1];  // > 0: the form is chosen only for RP > 0 on interior cells
        sw[i] = std::sqrt(w[i]);
    }
    const int nk = L / 2 + 1;  // distinct wavenumbers kw = 0..L/2
    std::vector<std::vector<double>> qf(nk), qi(nk), mu(nk);
    std::vector<int> fails(nk, 0);
    auto one = [&](int kw) {
        const double ak = h.akj[kw == 0 ? 0 : 2 * kw - 1];  // slot 2kw-1 holds wavenumber kw
        std::vector<double> d(M), e(M > 1 ? M - 1 : 0), ut, ev;
        for (int i = 0; i < M; ++i) {
            d[i] = h.bm[i] - ak * h.r2inv[i];
            if (i + 1 < M) e[i] = h.cm[i] * sw[i] / sw[i + 1];
        }
        // periodic x1 (MP=0): corners AM(1) sqrt(w(0)/w(M-1)) and CM(M) sqrt(w(M-1)/w(0))
        const double corner = P->mp == 0 ? 0.5 * (h.am[0] * sw[0] / sw[M - 1] + h.cm[M - 1] * sw[M - 1] / sw[0]) : 0.0;
        if (P->mp == 0 ? cyclic_eigen(M, d, e, corner, ut, ev) : tridiag_eigen(M, d, e, ut, ev)) {
            fails[kw] = 1;
            return;
        }
        if (kw == 0) {  // exact null pair
            int m0 = 0;
            for (int m = 1; m < M; ++m)
                if (std::fabs(ev[m]) < std::fabs(ev[m0])) m0 = m;
            double nrm = 0.0, sgn = 0.0;
            for (int i = 0; i < M; ++i) {
                nrm += w[i];
                sgn += ut[(size_t)m0 * M + i];
            }
            nrm = std::sqrt(nrm);
            sgn = sgn >= 0 ? 1.0 : -1.0;
            for (int i = 0; i < M; ++i) ut[(size_t)m0 * M + i] = sgn * sw[i] / nrm;
            ev[m0] = 0.0;
        }
        qf[kw].resize((size_t)M * M);
        qi[kw].resize((size_t)M * M);
        for (int m = 0; m < M; ++m)
            for (int i = 0; i < M; ++i) {
                const double u = ut[(size_t)m * M + i];
                qf[kw][i + (size_t)M * m] = u * sw[i];
                qi[kw][i + (size_t)M * m] = u / sw[i];
            }
        mu[kw] = ev;
    };
    {
        const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> th;
        for (int t = 0; t < nt; ++t)
            th.emplace_back([&, t] {
                for (int kw = t; kw < nk; kw += nt) one(kw);
            });
        for (auto& x : th) x.join();
    }
    for (int kw = 0; kw < nk; ++kw)
        if (fails[kw]) {
            set_error("x1 eigen-decomposition did not converge");
            return EDDY_ERR_NUMERIC;
        }
    int null_x = 0;  // the exact null pair set above
    for (int m = 1; m < M; ++m)
        if (std::fabs(mu[0][m]) < std::fabs(mu[0][null_x])) null_x = m;
    HostTables z;
    z.akj.assign(L, 0.0);
    z.wsum.assign(M, 0.0);
    for (int m = 0; m < M; ++m)
        for (int i = 0; i < M; ++i) z.wsum[m] += qi[0][i + (size_t)M * m];
    int zp = 0, zq = 0;
    xsolve_partition(N, P->tune.xsolve_chunk, zp, zq);
    const size_t n = (size_t)zp * zq;
    z.am_x.assign(n, 0.0);
    z.cm_x.assign(n, 0.0);
    z.bm_x.assign(n, 1.0);
    z.r2inv_x.assign(n, 0.0);
    for (int pp = 0; pp < zp; ++pp)
        for (int r = 0; r < zq; ++r) {
            const int k = pp * zq + r;
            if (k >= N) continue;
            const size_t t = (size_t)r * zp + pp;
            // periodic x2 lines: the corner couplings go to the line solve as scalars
            z.am_x[t] = k == 0 ? 0.0 : an[k];
            z.cm_x[t] = k == N - 1 ? 0.0 : cn[k];
            z.bm_x[t] = bn[k];
        }
    int rc;
    const size_t mm = (size_t)M * M;
    if ((rc = upload_tables(P->tz64, z)) || (rc = upload_tables(P->tz32, z)) ||
        (rc = upload_slots(P->tz64, mu, L, M, &P->tz64.t.lamj)) ||
        (rc = upload_slots(P->tz32, mu, L, M, &P->tz32.t.lamj)) ||
        (rc = upload_slots(P->tz64, qf, L, mm, &P->tz64.t.qf)) || (rc = upload_slots(P->tz32, qf, L, mm, &P->tz32.t.qf)) ||
        (rc = upload_slots(P->tz64, qi, L, mm, &P->tz64.t.qiT)) || (rc = upload_slots(P->tz32, qi, L, mm, &P->tz32.t.qiT)))
        return rc;
    P->dz = P->d;
    P->dz.M = N;
    P->dz.N = M;
    P->dz.null_m = null_x;
    P->dz.cyc_x = P->d.cyc_z;
    P->dz.cyc_z = P->d.cyc_x;
    return 0;
}

static int create_plan(eddy_plan_s** out, int nx, int ny, int nz, double delysq, const double* ap, const double* au,
                       const double* bp, const double* bv, const double* cp, const double* cw, const double* rp,
                       const double* ru, int mp, int lp, int np, const eddy_tuning* tuning, void* stream,
                       bool batch = false) {
    *out = nullptr;
    int rc = validate_dims(nx, ny, nz, mp, lp, np);
    if (rc) return rc;
    if (!ap || !au || !cp || !cw || !rp || !ru || !(delysq > 0.0)) {
        set_error("null metric array or non-positive DELYSQ");
        return EDDY_ERR_ARG;
    }
    std::unique_ptr<eddy_plan_s, void (*)(eddy_plan_s*)> P(new eddy_plan_s(), free_plan);
    const int M = nx - 2, L = ny - 2, N = nz - 2;
    P->d = Dims{nx, ny, nz, M, L, N, 0, mp == 0 ? 1 : 0, np == 0 ? 1 : 0, lp == 1 ? 1 : 0};
    P->mp = mp;
    P->lp = lp;
    P->np = np;
    P->delysq = delysq;
    P->stream = static_cast<hipStream_t>(stream);
    // y-transform tiles of at least 4096 elements where M allows: L/2-point FFTs over TI
    // pencils then give every one of the 256 threads a radix-8 butterfly in each pass
    P->tune.yfft_pencils = 16;
    while (P->tune.yfft_pencils < 64 && P->tune.yfft_pencils < M && (long)P->tune.yfft_pencils * L < 4096)
        P->tune.yfft_pencils *= 2;
    P->tune.ztrans_pencils = 16;
    P->tune.xsolve_chunk = 12;
    P->tune.use_dense_z = 0;
    P->tune.fft_radix = 8;
    P->tune.yfft_gemm = 0;
    P->tune.slab_chunks = 1;
    if (tuning) {
        if (tuning->yfft_pencils > 0) P->tune.yfft_pencils = tuning->yfft_pencils;
        if (tuning->ztrans_pencils > 0) P->tune.ztrans_pencils = tuning->ztrans_pencils;
        if (tuning->xsolve_chunk > 0) P->tune.xsolve_chunk = tuning->xsolve_chunk;
        P->tune.use_dense_z = tuning->use_dense_z;
        if (tuning->fft_radix > 0) P->tune.fft_radix = tuning->fft_radix;
        P->tune.yfft_gemm = tuning->yfft_gemm ? 1 : 0;
        if (tuning->slab_chunks > 0) P->tune.slab_chunks = tuning->slab_chunks;
        P->tune.zgemm = tuning->zgemm;
        if (tuning->yfft_iblock > 0) P->tune.yfft_iblock = tuning->yfft_iblock;
        P->tune.zgemm_lib = tuning->zgemm_lib ? 1 : 0;
    }
    if (P->tune.fft_radix != 8 && P->tune.fft_radix != 4 && P->tune.fft_radix != 2) {
        set_error("fft_radix must be 8, 4 or 2");
        return EDDY_ERR_ARG;
    }
    auto pow2 = [](int v) { return v > 0 && (v & (v - 1)) == 0; };
    if (!pow2(P->tune.yfft_pencils) || !pow2(P->tune.ztrans_pencils) || P->tune.ztrans_pencils < 4 ||
        P->tune.yfft_pencils > 64 || P->tune.ztrans_pencils > 64 || P->tune.xsolve_chunk < 3 ||
        P->tune.xsolve_chunk > kXsolveMaxChunk) {
        set_error("yfft_pencils and ztrans_pencils must be powers of two (ztrans >= 4, both <= 64), "
                  "3 <= xsolve_chunk <= 32");
        return EDDY_ERR_ARG;
    }
    HIP_TRY(hipGetDevice(&P->device));
    // Keep every kernel's LDS within the 160 KiB of a gfx950 CU (FP64 sizes bound FP32), and
    // the in-place FFTs at one butterfly per thread and pass (tile width x half-length /
    // smallest radix <= kFftThreads = 512).
    const size_t lds_cap = 160 * 1024;
    // the y transform: the radix FFT of L/2 complex points for an even L whose half factors
    // into 2, 3 and 5; any other periodic length (fftpack's general odd factors, fftpack.f:2096-
    // 2123) takes the dense real DFT, marked by a radix plan of length 0
    const bool ydense = (L % 2) || !radix_plan(L / 2, P->fpy, P->tune.fft_radix);
    if (ydense) P->fpy = FftPlan{0, 0, {}};
    {
        auto ylds = [&](int ti) {
            return ydense ? ((size_t)ti * fft_ld(L) + 2 * (size_t)L) * sizeof(double)
                          : (2 * (size_t)ti * fft_ld(L / 2) + 4 * (size_t)fft_ld(L / 2)) * sizeof(double);
        };
        auto yok = [&](int ti) {
            return ylds(ti) <= lds_cap && (ydense || (long)ti * (L / 2) <= (long)min_radix(P->fpy) * kFftThreads);
        };
        while (P->tune.yfft_pencils > 1 && !yok(P->tune.yfft_pencils)) P->tune.yfft_pencils /= 2;
        if (!yok(P->tune.yfft_pencils)) {
            set_error("periodic length too large for the in-place FFT");
            return EDDY_ERR_UNSUPPORTED;
        }
    }
    // RP is read on interior cells only (the ghost radius is negative when x1 starts on the
    // axis of a cylindrical grid, grid.f:462-466); RU(1) = 0 on the axis is allowed
    for (int i = 0; i < nx; ++i)
        if (!(ap[i] > 0) || !std::isfinite(au[i]) || !std::isfinite(rp[i]) || !std::isfinite(ru[i]) ||
            (i >= 1 && i <= nx - 2 && !(rp[i] > 0))) {
            set_error("non-finite or non-positive x1 metric");
            return EDDY_ERR_ARG;
        }
    for (int k = 0; k < nz; ++k)
        if (!(cp[k] > 0) || !std::isfinite(cw[k])) {
            set_error("non-finite or non-positive x2 metric");
            return EDDY_ERR_ARG;
        }

    HostTables h;
    // x1 operator, IPRES=1 form (direct.f:344-352); r2inv multiplies AK (direct.f:536)
    h.am.resize(M);
    h.bm.resize(M);
    h.cm.resize(M);
    h.r2inv.resize(M);
    for (int i = 1; i <= M; ++i) {
        h.am[i - 1] = ap[i] / rp[i] * au[i - 1] * ru[i - 1];
        h.cm[i - 1] = ap[i] / rp[i] * au[i] * ru[i];
        h.r2inv[i - 1] = 1.0 / (rp[i] * rp[i]);
    }
    if (mp == 1) {  // x1 walls (direct.f:347-350); MP=0 keeps the periodic couplings
        h.am[0] = 0.0;
        h.cm[M - 1] = 0.0;
    }
    for (int i = 0; i < M; ++i) h.bm[i] = -h.am[i] - h.cm[i];
    // periodic x1 lines: the line solve takes the two corner couplings as scalars and the
    // tridiagonal part without them
    P->xalpha = mp == 0 ? h.cm[M - 1] : 0.0;
    P->xbeta = mp == 0 ? h.am[0] : 0.0;
    {
        int xp = 0, xq = 0;
        if (!xsolve_partition(M, P->tune.xsolve_chunk, xp, xq)) {
            set_error("x1 solve: no thread partition for M");
            return EDDY_ERR_UNSUPPORTED;
        }
        const size_t n = (size_t)xp * xq;
        h.am_x.assign(n, 0.0);
        h.cm_x.assign(n, 0.0);
        h.bm_x.assign(n, 1.0);
        h.r2inv_x.assign(n, 0.0);
        for (int pp = 0; pp < xp; ++pp)
            for (int r = 0; r < xq; ++r) {
                const int i = pp * xq + r;
                if (i >= M) continue;
                const size_t t = (size_t)r * xp + pp;
                h.am_x[t] = i == 0 ? 0.0 : h.am[i];
                h.cm_x[t] = i == M - 1 ? 0.0 : h.cm[i];
                h.bm_x[t] = h.bm[i];
                h.r2inv_x[t] = h.r2inv[i];
            }
    }
    // x2 operator (direct.f:370-378)
    std::vector<double> an(N), bn(N), cn(N);
    for (int k = 1; k <= N; ++k) {
        an[k - 1] = cp[k] * cw[k - 1];
        cn[k - 1] = cp[k] * cw[k];
    }
    if (np == 1) {  // x2 walls (direct.f:372-375); NP=0 keeps the periodic couplings
        an[0] = 0.0;
        cn[N - 1] = 0.0;
    }
    for (int k = 0; k < N; ++k) bn[k] = -an[k] - cn[k];
    P->zalpha = np == 0 ? cn[N - 1] : 0.0;
    P->zbeta = np == 0 ? an[0] : 0.0;
    // A periodic x2 operator T = diag(CP) S is symmetric under the weights 1/CP when its two
    // corner couplings match, CW(1) = CW(N+1) (a periodic grid); then an eigenbasis of x2
    // exists as for the walls. Otherwise x1 is diagonalised instead (the x1 form below).
    const bool zsym = np == 1 || std::fabs(cw[0] - cw[N]) <= 1e-12 * std::max(std::fabs(cw[0]), std::fabs(cw[N]));
    // likewise a periodic x1 operator under the weights RP/AP: AU(1) RU(1) = AU(M+1) RU(M+1)
    const bool xsym = mp == 1 || std::fabs(au[0] * ru[0] - au[M] * ru[M]) <=
                                     1e-12 * std::max(std::fabs(au[0] * ru[0]), std::fabs(au[M] * ru[M]));
    // modified wavenumbers AK(J/2+1) per fftpack slot (direct.f:321-331): slot j (0-based)
    // holds wavenumber (j+1)/2, eigenvalue 2(1-cos(2 pi kw/L))/DELYSQ = 4 sin^2(pi kw/L)/DELYSQ
    h.akj.resize(L);
    for (int j = 0; j < L; ++j) {
        const int kw = (j + 1) / 2;
        const double s = std::sin(M_PI * (double)kw / (double)L);
        h.akj[j] = 4.0 * s * s / delysq;
    }
    // x2 transform: the DCT diagonalises the uniform walled x2 only
    bool uniform = (N % 2 == 0) && np == 1;
    for (int k = 2; k <= N + 1 && uniform; ++k)
        if (std::fabs(cp[k] - cp[1]) > 1e-15 * cp[1]) uniform = false;
    for (int k = 1; k <= N && uniform; ++k)
        if (std::fabs(cw[k] - cw[1]) > 1e-15 * cw[1]) uniform = false;
    if (N > 1 && std::fabs(cp[1] * cw[1] - cn[0]) > 1e-14 * cn[0]) uniform = false;
    P->dense = (!uniform || P->tune.use_dense_z) ? 1 : 0;
    h.lam.resize(N);
    h.wsum.assign(N, 0.0);
    if (!P->dense) {
        const double c = cn[0];
        for (int m = 0; m < N; ++m) {
            const double s = std::sin(M_PI * (double)m / (2.0 * N));
            h.lam[m] = -4.0 * c * s * s;
        }
        h.wsum[0] = 1.0;
        P->d.null_m = 0;
        if (!radix_plan(N / 2, P->fpz, P->tune.fft_radix)) {
            P->dense = 1;
        } else {
            twiddles(N / 2, N / 2, h.ztwc_re, h.ztwc_im);
            twiddles(N, N / 2 + 1, h.ztwr_re, h.ztwr_im);
            h.zdct_re.resize(N);
            h.zdct_im.resize(N);
            for (int m = 0; m < N; ++m) {
                const double a = M_PI * (double)m / (2.0 * N);
                h.zdct_re[m] = std::cos(a);
                h.zdct_im[m] = -std::sin(a);
            }
        }
    }
    // x2 tile width: the DCT's in-place FFT as above; the dense transform works on groups of
    // 4 lines staged whole in LDS.
    {
        auto zok = [&](int ti) {
            if (P->dense) return (size_t)N * ti * sizeof(double) <= lds_cap;
            return (2 * (size_t)ti * fft_ld(N / 2) + 4 * (size_t)N + 2) * sizeof(double) <= lds_cap &&
                   (long)ti * (N / 2) <= (long)min_radix(P->fpz) * kFftThreads;
        };
        const int zmin = P->dense ? 4 : 1;
        // the DCT of a short x2 takes more pencils per tile (at least 4096 elements where M
        // allows), as the y transform does
        if (!P->dense && !(tuning && tuning->ztrans_pencils > 0))
            while (P->tune.ztrans_pencils < 64 && P->tune.ztrans_pencils < M &&
                   (long)P->tune.ztrans_pencils * N < 4096 && zok(2 * P->tune.ztrans_pencils))
                P->tune.ztrans_pencils *= 2;
        while (P->tune.ztrans_pencils > zmin && !zok(P->tune.ztrans_pencils)) P->tune.ztrans_pencils /= 2;
        if (!zok(P->tune.ztrans_pencils)) {
            set_error("x2 length too large for the x2 transform");
            return EDDY_ERR_UNSUPPORTED;
        }
    }
    // The GEMM forms of a stretched x2 (tuning zgemm: 0 auto, 1 x2 products, 2 x1 products,
    // -1 the fused tile kernel): a separable transform costs 2 * (its axis length) flops per
    // element, so auto diagonalises the shorter of x1 and x2 once either is GEMM-sized.
    // Batch plans that take the x1 form also hold the x2 eigenvectors: the one-launch batch
    // mode applies them with its tile kernel.
    if (P->dense && P->tune.zgemm != -1) {
        const bool big = std::max(M, N) >= 128;
        int form = P->tune.zgemm;
        if (form == 0) form = big ? (M < N ? 2 : 1) : 0;
        int zp = 0, zq = 0;
        bool rp_pos = true;  // the x1 form's symmetrising weights RP/AP need RP > 0
        for (int i = 1; i <= M; ++i) rp_pos = rp_pos && rp[i] > 0.0;
        if (form == 2 && !rp_pos && P->tune.zgemm == 2) {
            set_error("zgemm = 2 (x1 form) needs RP > 0 on every interior cell");
            return EDDY_ERR_UNSUPPORTED;
        }
        if (form == 2 && (!rp_pos || !xsym || !xsolve_partition(N, P->tune.xsolve_chunk, zp, zq))) form = 1;
        if (!zsym) form = 2;  // no x2 eigenbasis: the x1 form is the only separable one
        P->zgemm = form == 1;
        P->xgemm = form == 2;
    }
    if (!zsym) {
        bool rp_pos = true;
        for (int i = 1; i <= M; ++i) rp_pos = rp_pos && rp[i] > 0.0;
        int zp = 0, zq = 0;
        if (!P->xgemm || !rp_pos || !xsym || !xsolve_partition(N, P->tune.xsolve_chunk, zp, zq)) {
            set_error("periodic x2 with CW(1) != CW(N+1) needs the x1 eigenbasis (zgemm 0 or 2, RP > 0, x1 walls "
                      "or a symmetric periodic x1)");
            return EDDY_ERR_UNSUPPORTED;
        }
    }
    if (P->xgemm) {
        rc = build_x1_basis(P.get(), M, L, N, ap, rp, h, an, bn, cn);
        if (rc) return rc;
    }
    if (P->dense && (!P->xgemm || batch)) {
        // H = CP^{1/2} S CP^{1/2}, T = diag(CP) S  (generalized symmetric form of the x2 operator):
        // symmetric tridiagonal, diagonal BN, off-diagonal CN(k)/CP(k+1) sqrt(CP(k) CP(k+1))
        std::vector<double> sq(N), hd(N), he(N > 1 ? N - 1 : 0), Ut, w;
        for (int k = 0; k < N; ++k) sq[k] = std::sqrt(cp[k + 1]);
        for (int k = 0; k < N; ++k) {
            hd[k] = bn[k];
            if (k + 1 < N) he[k] = cn[k] / cp[k + 1] * sq[k] * sq[k + 1];
        }
        // periodic x2: the corner couplings H(0, N-1) = AN(1) sqrt(CP(N)/CP(1)) and
        // H(N-1, 0) = CN(N) sqrt(CP(1)/CP(N)) (equal when zsym; their mean)
        const double corner = np == 0 ? 0.5 * (an[0] * sq[N - 1] / sq[0] + cn[N - 1] * sq[0] / sq[N - 1]) : 0.0;
        if (np == 0 ? cyclic_eigen(N, hd, he, corner, Ut, w) : tridiag_eigen(N, hd, he, Ut, w)) {
            set_error("x2 eigen-decomposition did not converge");
            return EDDY_ERR_NUMERIC;
        }
        std::vector<double> U((size_t)N * N);  // eigenvectors as columns: U[k*N + m]
        for (int m = 0; m < N; ++m)
            for (int k = 0; k < N; ++k) U[(size_t)k * N + m] = Ut[(size_t)m * N + k];
        int m0 = 0;
        for (int m = 1; m < N; ++m)
            if (std::fabs(w[m]) < std::fabs(w[m0])) m0 = m;
        // exact null pair: lam = 0, u ~ CP^{-1/2}
        double nrm = 0.0, sgn = 0.0;
        for (int k = 0; k < N; ++k) {
            nrm += 1.0 / cp[k + 1];
            sgn += U[(size_t)k * N + m0];
        }
        nrm = std::sqrt(nrm);
        sgn = sgn >= 0 ? 1.0 : -1.0;
        for (int k = 0; k < N; ++k) U[(size_t)k * N + m0] = sgn / (sq[k] * nrm);
        w[m0] = 0.0;
        P->d.null_m = m0;
        h.lam = w;
        h.qf.resize((size_t)N * N);
        h.qiT.resize((size_t)N * N);
        for (int k = 0; k < N; ++k)
            for (int m = 0; m < N; ++m) {
                const double u = U[(size_t)k * N + m];
                h.qf[(size_t)k * N + m] = u / sq[k];   // A[s=k][o=m] = Qfwd(m,k)
                h.qiT[(size_t)m * N + k] = u * sq[k];  // A[s=m][o=k] = Qinv(k,m)
                h.wsum[m] += u * sq[k];
            }
        P->fpz = FftPlan{1, 0, {}};
    }
    // y real FFT tables (L = 2H; radix plan made above); the dense DFT reads all L roots
    if (ydense) {
        twiddles(L, L, h.ytwr_re, h.ytwr_im);
    } else {
        twiddles(L / 2, L / 2, h.ytwc_re, h.ytwc_im);
        twiddles(L, L / 2 + 1, h.ytwr_re, h.ytwr_im);
    }

    if ((rc = upload_tables(P->t64, h)) || (rc = upload_tables(P->t32, h))) return rc;
    // dense real-DFT matrices of the FP32 matrix-core y transform (tuning yfft_gemm)
    if (P->tune.yfft_gemm && ygemm_supported(P->d)) {
        std::vector<double> ft((size_t)L * L), gt((size_t)L * L);
        for (int j = 0; j < L; ++j) {
            for (int p = 0; p < L; ++p) {
                double f, g;
                if (p == 0) {
                    f = 1.0;
                    g = 1.0;
                } else if (p == L - 1) {
                    f = (j & 1) ? -1.0 : 1.0;
                    g = f;
                } else {
                    const int m = (p + 1) / 2;
                    const long t = ((long)m * j) % L;  // reduced angle 2 pi t / L
                    const double ang = 2.0 * M_PI * (double)t / (double)L;
                    f = (p & 1) ? std::cos(ang) : -std::sin(ang);
                    g = 2.0 * f;
                }
                ft[(size_t)j * L + p] = f;            // FT[j][p] = F(p, j)
                gt[(size_t)p * L + j] = g / (double)L;  // GT[p][j] = G(j, p)
            }
        }
        if ((rc = upload(P->t32, ft, &P->t32.t.ydft_f)) || (rc = upload(P->t32, gt, &P->t32.t.ydft_b))) return rc;
    } else {
        P->tune.yfft_gemm = 0;
    }

    // FP64 operator and metrics
    std::vector<double> vd;
    if ((rc = upload64(P.get(), h.am, &P->op.am)) || (rc = upload64(P.get(), h.bm, &P->op.bm)) ||
        (rc = upload64(P.get(), h.cm, &P->op.cm)) || (rc = upload64(P.get(), h.r2inv, &P->op.r2inv)) ||
        (rc = upload64(P.get(), an, &P->op.an)) || (rc = upload64(P.get(), bn, &P->op.bn)) ||
        (rc = upload64(P.get(), cn, &P->op.cn)))
        return rc;
    P->op.inv_delysq = 1.0 / delysq;
    std::vector<double> hap(ap, ap + nx), hau(au, au + nx), hrp(rp, rp + nx), hru(ru, ru + nx), hcp(cp, cp + nz),
        hcw(cw, cw + nz), hbp(ny), hbv(ny);
    for (int j = 0; j < ny; ++j) {
        hbp[j] = bp ? bp[j] : 1.0 / std::sqrt(delysq);
        hbv[j] = bv ? bv[j] : 1.0 / std::sqrt(delysq);
    }
    if ((rc = upload64(P.get(), hap, &P->mt.ap)) || (rc = upload64(P.get(), hau, &P->mt.au)) ||
        (rc = upload64(P.get(), hrp, &P->mt.rp)) || (rc = upload64(P.get(), hru, &P->mt.ru)) ||
        (rc = upload64(P.get(), hbp, &P->mt.bp)) || (rc = upload64(P.get(), hbv, &P->mt.bv)) ||
        (rc = upload64(P.get(), hcp, &P->mt.cp)) || (rc = upload64(P.get(), hcw, &P->mt.cw)))
        return rc;
    // separable cell volumes RP/(AP*BP*CP) (the solvability weights) and their interior sum
    std::vector<double> hvx(nx, 0.0), hvy(ny, 0.0), hvz(nz, 0.0);
    long double sx = 0.0L, sy = 0.0L, sz = 0.0L;
    for (int i = 0; i < nx; ++i) hvx[i] = hrp[i] / hap[i];
    for (int j = 0; j < ny; ++j) hvy[j] = 1.0 / hbp[j];
    for (int k = 0; k < nz; ++k) hvz[k] = 1.0 / hcp[k];
    for (int i = 1; i <= M; ++i) sx += hvx[i];
    for (int j = 1; j <= L; ++j) sy += hvy[j];
    for (int k = 1; k <= N; ++k) sz += hvz[k];
    P->mt.vol_total = (double)(sx * sy * sz);
    P->mt.cart = 1;
    for (int i = 0; i < nx; ++i)
        if (hrp[i] != 1.0 || hru[i] != 1.0) P->mt.cart = 0;
    P->mt.bp0 = hbp[0];
    for (int j = 1; j < ny; ++j)
        if (hbp[j] != hbp[0]) P->mt.bp0 = 0.0;
    P->mt.vy0 = P->mt.bp0 != 0.0 ? hvy[0] : 0.0;
    if ((rc = upload64(P.get(), hvx, &P->mt.volx)) || (rc = upload64(P.get(), hvy, &P->mt.voly)) ||
        (rc = upload64(P.get(), hvz, &P->mt.volz)))
        return rc;
    {
        std::vector<double> vt(1, P->mt.vol_total);
        if ((rc = upload64(P.get(), vt, &P->vol_total_dev))) return rc;
    }

    // work buffers
    const size_t cells = (size_t)M * L * N;
    void* p = nullptr;
    HIP_TRY(hipMalloc(&p, cells * sizeof(double)));
    P->owned.push_back(p);
    P->S = static_cast<double*>(p);
    HIP_TRY(hipMalloc(&p, std::max(N, 1) * sizeof(double)));
    P->owned.push_back(p);
    P->linesum = static_cast<double*>(p);
    HIP_TRY(hipMalloc(&p, RED_COUNT * sizeof(double)));
    P->owned.push_back(p);
    P->red = static_cast<double*>(p);
    // per-workgroup partials: two quantities of up to M*N workgroups each
    HIP_TRY(hipMalloc(&p, 2 * (size_t)M * N * sizeof(double)));
    P->owned.push_back(p);
    P->part = static_cast<double*>(p);
    HIP_TRY(hipMemset(P->red, 0, RED_COUNT * sizeof(double)));
    P->bytes = cells * sizeof(double) + (size_t)N * sizeof(double) + RED_COUNT * sizeof(double) +
               2 * (size_t)M * N * sizeof(double);
    // the GEMM forms: a second spectral array; with rocBLAS (tuning zgemm_lib) also its handle,
    // with one product of each kind and precision run now, at the shapes the solves use, so
    // that rocBLAS has loaded its kernels and sized its workspace before any graph capture of
    // the plan's work (the library's MFMA kernel needs neither)
    if (P->zgemm || P->xgemm) {
        HIP_TRY(hipMalloc(&p, cells * sizeof(double)));
        P->owned.push_back(p);
        P->S2 = static_cast<double*>(p);
        P->bytes += cells * sizeof(double);
    }
    if ((P->zgemm || P->xgemm) && P->tune.zgemm_lib) {
        if (zgemm_create(&P->blas, P->stream)) {
            set_error("rocBLAS handle creation failed");
            return EDDY_ERR_DEVICE;
        }
        HIP_TRY(hipMemsetAsync(P->S, 0, cells * sizeof(double), P->stream));
        HIP_TRY(hipMemsetAsync(P->S2, 0, cells * sizeof(double), P->stream));
        HIP_TRY(hipMemsetAsync(P->linesum, 0, std::max(N, 1) * sizeof(double), P->stream));
        float* S32 = reinterpret_cast<float*>(P->S);
        float* S232 = reinterpret_cast<float*>(P->S2);
        const bool fail =
            P->zgemm ? (launch_zgemm_fwd<double>(P->stream, P->blas, P->d, P->t64.t, P->S, P->S2, nullptr, 0, L) ||
                        launch_zgemm_fwd<float>(P->stream, P->blas, P->d, P->t32.t, S32, S232, nullptr, 0, L))
                     : (launch_xgemm_fwd<double>(P->stream, P->blas, P->d, P->tz64.t, P->S, P->S2, nullptr, 0, L) ||
                        launch_xgemm_fwd<float>(P->stream, P->blas, P->d, P->tz32.t, S32, S232, nullptr, 0, L) ||
                        launch_xgemm_inv<double>(P->stream, P->blas, P->d, P->tz64.t, P->S2, P->S,
                                                 reinterpret_cast<double*>(P->linesum), P->dz.null_m, 0, L) ||
                        launch_xgemm_inv<float>(P->stream, P->blas, P->d, P->tz32.t, S232, S32,
                                                reinterpret_cast<float*>(P->linesum), P->dz.null_m, 0, L));
        if (fail) {
            set_error("rocBLAS refused the transform products");
            return EDDY_ERR_DEVICE;
        }
        HIP_TRY(hipStreamSynchronize(P->stream));
    }
    *out = P.release();
    return 0;
}

template <typename T>
static BatchProblem<T> problem_of(eddy_plan_s* P, const DevTables<T>& t) {
    BatchProblem<T> b{};
    b.d = P->d;
    b.fpy = P->fpy;
    b.fpz = P->fpz;
    b.tb = t.t;
    b.op = P->op;
    b.mt = P->mt;
    b.S = reinterpret_cast<T*>(P->S);
    b.linesum = reinterpret_cast<T*>(P->linesum);
    b.part = P->part;
    b.red = P->red;
    b.tiy = P->tune.yfft_pencils;
    b.tiz = P->tune.ztrans_pencils;
    b.dense = P->dense;
    b.chunk = P->tune.xsolve_chunk;
    return b;
}

// ------------------------------------------------------------------ DIRECT drop-in cache
struct DirectCache {
    std::mutex mu;
    eddy_plan_s* plan = nullptr;
    uint64_t key = 0;
    HostStage* stage = nullptr;  // pinned staging of DP, kept across calls (pinning is slow)
};
static DirectCache g_cache;

static uint64_t fnv(uint64_t h, const void* p, size_t n) {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) {
        h ^= c[i];
        h *= 1099511628211ull;
    }
    return h;
}

}  // namespace eddy

using namespace eddy;

extern "C" {

const char* eddy_last_error(void) { return g_err.c_str(); }
const char* eddy_version(void) { return "eddyproj 0.1 (gfx950)"; }

int eddy_plan_create(eddy_plan_t* plan, int nx, int ny, int nz, double delysq, const double* ap, const double* au,
                     const double* bp, const double* bv, const double* cp, const double* cw, const double* rp,
                     const double* ru, int mp, int lp, int np, const eddy_tuning* tuning, void* stream) {
    if (!plan) {
        set_error("null plan pointer");
        return EDDY_ERR_ARG;
    }
    return create_plan(plan, nx, ny, nz, delysq, ap, au, bp, bv, cp, cw, rp, ru, mp, lp, np, tuning, stream);
}

int eddy_plan_destroy(eddy_plan_t plan) {
    if (!plan) return 0;
    if (plan->stream) (void)hipStreamSynchronize(plan->stream);
    else (void)hipDeviceSynchronize();
    free_plan(plan);
    return 0;
}

int eddy_plan_set_stream(eddy_plan_t plan, void* stream) {
    if (!plan) return EDDY_ERR_ARG;
    plan->stream = static_cast<hipStream_t>(stream);
    if (plan->blas && zgemm_set_stream(plan->blas, plan->stream)) {
        set_error("rocBLAS refused the plan's new stream");
        return EDDY_ERR_DEVICE;
    }
    return 0;
}

int eddy_plan_get_tuning(eddy_plan_t plan, eddy_tuning* out) {
    if (!plan || !out) return EDDY_ERR_ARG;
    *out = plan->tune;
    out->use_dense_z = plan->dense;
    out->zgemm = plan->xgemm ? 2 : (plan->zgemm ? 1 : -1);
    return 0;
}

size_t eddy_plan_device_bytes(eddy_plan_t plan) { return plan ? plan->bytes : 0; }

int eddy_direct_dev(eddy_plan_t plan, double* d_dp) {
    if (!plan || !d_dp) {
        set_error("null plan or DP");
        return EDDY_ERR_ARG;
    }
    return run_direct64(plan, d_dp, 0);
}

int eddy_direct_refined_dev(eddy_plan_t plan, double* d_dp, int sweeps, double* resid_out) {
    if (!plan || !d_dp || sweeps < 0 || sweeps > 16) {
        set_error("null plan/DP or sweeps outside 0..16");
        return EDDY_ERR_ARG;
    }
    int rc = run_refined(plan, d_dp, sweeps, resid_out != nullptr, 0);
    if (rc) return rc;
    if (resid_out) return read_resid(plan, resid_out);
    return 0;
}

int eddy_direct_refined_adaptive_dev(eddy_plan_t plan, double* d_dp, double tol, int max_sweeps, int* sweeps_out,
                                     double* resid_out) {
    if (!plan || !d_dp || max_sweeps < 0 || max_sweeps > 16 || !(tol > 0)) {
        set_error("bad arguments to the adaptive refinement");
        return EDDY_ERR_ARG;
    }
    // keep the RHS: each trial restarts from it, growing the sweep count until tol is met
    int rc = ensure_rhs(plan);
    if (rc) return rc;
    double* keep = nullptr;
    HIP_TRY(hipMalloc(&keep, field_elems(plan) * sizeof(double)));
    // freed on every return below, after the stream has finished with it
    struct Keep {
        double* p;
        hipStream_t st;
        ~Keep() {
            (void)hipStreamSynchronize(st);
            (void)hipFree(p);
        }
    } guard{keep, plan->stream};
    HIP_TRY(hipMemcpyAsync(keep, d_dp, field_elems(plan) * sizeof(double), hipMemcpyDeviceToDevice, plan->stream));
    double r = 0.0;
    int s = 0;
    for (s = 0; s <= max_sweeps; ++s) {
        if (s > 0)
            HIP_TRY(hipMemcpyAsync(d_dp, keep, field_elems(plan) * sizeof(double), hipMemcpyDeviceToDevice,
                                   plan->stream));
        rc = run_refined(plan, d_dp, s, 1, 0);
        if (!rc) rc = read_resid(plan, &r);
        if (rc || r <= tol) break;
    }
    if (rc) return rc;
    if (sweeps_out) *sweeps_out = std::min(s, max_sweeps);
    if (resid_out) *resid_out = r;
    if (r > tol) {
        set_error("defect correction did not reach the tolerance");
        return EDDY_ERR_NUMERIC;
    }
    return 0;
}

int eddy_divergence_dev(eddy_plan_t plan, const double* us, const double* vs, const double* ws, double* div,
                        double scale) {
    if (!plan || !us || !vs || !ws || !div) return EDDY_ERR_ARG;
    KTimer kt(plan, EDDY_K_DIVERGENCE);
    launch_divergence(plan->stream, plan->d, plan->mt, us, vs, ws, div, scale, rd(plan), 0);
    return check_launch("divergence");
}

int eddy_correct_dev(eddy_plan_t plan, const double* us, const double* vs, const double* ws, double* uo, double* vo,
                     double* wo, const double* dp, double alfxdt) {
    if (!plan || !us || !vs || !ws || !uo || !vo || !wo || !dp) return EDDY_ERR_ARG;
    KTimer kt(plan, EDDY_K_CORRECT);
    launch_correct(plan->stream, plan->d, plan->mt, us, vs, ws, uo, vo, wo, dp, alfxdt, 0);
    return check_launch("correct");
}

}  // extern "C"

namespace eddy {
// The projection of eddy_project_dev. dscale: null (ALFXDT from the argument), or device
// {ALFXDT, 1/ALFXDT} read by the kernels at run time, so that a captured graph of the step
// serves every time step (the reference recomputes ALFXDT = ALF(IS)*DTM1 each step,
// Eddy6.f:2122, 2149).
static int project_core(eddy_plan_s* plan, const double* us, const double* vs, const double* ws, double* uo,
                        double* vo, double* wo, double* dp, double alfxdt, const double* dscale, int remove_mean,
                        int precision, int sweeps) {
    {
        KTimer kt(plan, EDDY_K_OTHER);
        launch_fill(plan->stream, plan->red, RED_COUNT, 0.0);
    }
    int rc = 0;
    if (precision == EDDY_FP64) {
        // divergence fused into the forward y-FFT: the RHS is never written to DP, and its
        // volume-weighted mean is removed from the zero y-mode in the x2 transform
        FwdInput in{};
        in.mode = YF_DIVERGENCE;
        in.us = us;
        in.vs = vs;
        in.ws = ws;
        in.mt = plan->mt;
        in.scale = 1.0 / alfxdt;
        in.dscale = dscale;
        rc = solve_core<double>(plan, in, remove_mean);
        if (!rc) rc = finish_direct64(plan, dp);
    } else {
        // the divergence is formed inside the first FP32 forward transform, which also stores
        // it in FP64 as the defect correction's RHS
        FwdInput in{};
        in.mode = YF_DIVERGENCE;
        in.us = us;
        in.vs = vs;
        in.ws = ws;
        in.mt = plan->mt;
        in.scale = 1.0 / alfxdt;
        in.dscale = dscale;
        rc = run_refined(plan, dp, sweeps, 0, remove_mean, &in);
    }
    if (rc) return rc;
    // REFRESHBC(DP) (Eddy6.f:2730): the periodic images of a periodic x2 into the z ghosts
    launch_refresh_z(plan->stream, plan->d, dp);
    {
        KTimer kt(plan, EDDY_K_CORRECT);
        launch_correct(plan->stream, plan->d, plan->mt, us, vs, ws, uo, vo, wo, dp, alfxdt, 1, dscale);
    }
    return check_launch("projection");
}
}  // namespace eddy

extern "C" {

int eddy_project_dev(eddy_plan_t plan, const double* us, const double* vs, const double* ws, double* uo, double* vo,
                     double* wo, double* dp, double alfxdt, int remove_mean, int precision, int sweeps) {
    if (!plan || !us || !vs || !ws || !uo || !vo || !wo || !dp || alfxdt == 0.0) {
        set_error("null pointer or zero ALFXDT");
        return EDDY_ERR_ARG;
    }
    if (precision != EDDY_FP64 && precision != EDDY_FP32_REFINE) return EDDY_ERR_ARG;
    return project_core(plan, us, vs, ws, uo, vo, wo, dp, alfxdt, nullptr, remove_mean, precision, sweeps);
}

int eddy_divergence_minmax_dev(eddy_plan_t plan, const double* us, const double* vs, const double* ws, double* dmin,
                               double* dmax) {
    if (!plan || !us || !vs || !ws || !dmin || !dmax) return EDDY_ERR_ARG;
    launch_div_minmax(plan->stream, plan->d, plan->mt, us, vs, ws, rd(plan));
    double h[2];
    HIP_TRY(hipMemcpyAsync(h, plan->red + RED_DMIN, sizeof(h), hipMemcpyDeviceToHost, plan->stream));
    HIP_TRY(hipStreamSynchronize(plan->stream));
    *dmin = h[0];
    *dmax = h[1];
    return 0;
}

int eddy_laplacian_dev(eddy_plan_t plan, const double* phi, double* out) {
    if (!plan || !phi || !out) return EDDY_ERR_ARG;
    launch_laplacian(plan->stream, plan->d, plan->op, phi, out);
    return check_launch("laplacian");
}

// ---------------------------------------------------------------- graphs
struct eddy_graph_s {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t stream = nullptr;
};

int eddy_graph_create(eddy_graph_t* graph, eddy_plan_t plan, const double* us, const double* vs, const double* ws,
                      double* uo, double* vo, double* wo, double* dp, double alfxdt, int remove_mean, int precision,
                      int sweeps) {
    if (!graph || !plan) return EDDY_ERR_ARG;
    if (!plan->stream) {
        set_error("graph capture needs a non-default stream on the plan");
        return EDDY_ERR_ARG;
    }
    if (plan->timing) {
        set_error("disable kernel timing before capturing a graph");
        return EDDY_ERR_ARG;
    }
    int rc = ensure_rhs(plan);
    if (rc) return rc;
    std::unique_ptr<eddy_graph_s> G(new eddy_graph_s());
    G->stream = plan->stream;
    HIP_TRY(hipStreamBeginCapture(plan->stream, hipStreamCaptureModeThreadLocal));
    rc = eddy_project_dev(plan, us, vs, ws, uo, vo, wo, dp, alfxdt, remove_mean, precision, sweeps);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(plan->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return device_error(e, "hipStreamEndCapture");
    G->graph = g;
    HIP_TRY(hipGraphInstantiate(&G->exec, g, nullptr, nullptr, 0));
    *graph = G.release();
    return 0;
}

int eddy_graph_launch(eddy_graph_t graph) {
    if (!graph) return EDDY_ERR_ARG;
    HIP_TRY(hipGraphLaunch(graph->exec, graph->stream));
    return 0;
}

int eddy_graph_destroy(eddy_graph_t graph) {
    if (!graph) return 0;
    if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
    if (graph->graph) (void)hipGraphDestroy(graph->graph);
    delete graph;
    return 0;
}

// ---------------------------------------------------------------- pressure update
int eddy_plan_set_pressure_bc(eddy_plan_t plan, int implx, int imply, double ru1, const double* rvimplx,
                              const double* rvimply, int itype1, int itype5, int itype6) {
    if (!plan) return EDDY_ERR_ARG;
    if ((implx && !rvimplx) || (imply && !rvimply)) {
        set_error("implicit corrections need their RVIMPLX / RVIMPLY masks");
        return EDDY_ERR_ARG;
    }
    if (itype5 == 0 || itype6 == 0) {
        set_error("one GPU holds both x2 boundaries: ITYPE(5) and ITYPE(6) must be boundary types");
        return EDDY_ERR_UNSUPPORTED;
    }
    const int nx = plan->d.nx;
    std::vector<double> hx(nx, 0.0), hy(nx, 0.0);
    if (rvimplx) hx.assign(rvimplx, rvimplx + nx);
    if (rvimply) hy.assign(rvimply, rvimply + nx);
    int rc = 0;
    if (!plan->rvimplx && (rc = upload64(plan, hx, &plan->rvimplx))) return rc;
    if (!plan->rvimply && (rc = upload64(plan, hy, &plan->rvimply))) return rc;
    HIP_TRY(hipMemcpy(const_cast<double*>(plan->rvimplx), hx.data(), nx * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(const_cast<double*>(plan->rvimply), hy.data(), nx * sizeof(double), hipMemcpyHostToDevice));
    plan->implx = implx ? 1 : 0;
    plan->imply = imply ? 1 : 0;
    plan->ru1 = ru1;
    plan->itype1 = itype1;
    plan->itype5 = itype5;
    plan->itype6 = itype6;
    return 0;
}

int eddy_pressure_update_dev(eddy_plan_t plan, double* d_p, const double* d_dp, double alfxdt) {
    if (!plan || !d_p || !d_dp || d_p == d_dp) return EDDY_ERR_ARG;
    if (!plan->rvimplx) {
        int rc = eddy_plan_set_pressure_bc(plan, 0, 0, 0.0, nullptr, nullptr, 0, 1, 1);
        if (rc) return rc;
    }
    PressArgs a{};
    a.d = plan->d;
    a.mt = plan->mt;
    a.p = d_p;
    a.dp = d_dp;
    a.rvx = plan->rvimplx;
    a.rvy = plan->rvimply;
    a.red = plan->red;
    a.alfxdt = alfxdt;
    a.ru1 = plan->ru1;
    a.dely = std::sqrt(plan->delysq);
    a.implx = plan->implx;
    a.imply = plan->imply;
    a.mp = plan->mp;
    a.lp = plan->lp;
    a.itype1 = plan->itype1;
    a.itype5 = plan->itype5;
    a.itype6 = plan->itype6;
    launch_pressure_update(plan->stream, a);
    return check_launch("pressure update");
}

// ---------------------------------------------------------------- resident loop
struct eddy_resident_s {
    eddy_plan_s* plan = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double* star = nullptr;  // predicted velocities us, vs, ws (3 fields)
    double* log = nullptr;   // 2 * cap doubles
    int* step = nullptr;     // device step counter
    int cap = 0;
    int enqueued = 0;        // steps enqueued so far (host side bound)
};

static int resident_free(eddy_resident_s* R) {
    if (R->exec) (void)hipGraphExecDestroy(R->exec);
    if (R->graph) (void)hipGraphDestroy(R->graph);
    if (R->star) (void)hipFree(R->star);
    if (R->log) (void)hipFree(R->log);
    if (R->step) (void)hipFree(R->step);
    delete R;
    return 0;
}

int eddy_perturb_dev(eddy_plan_t plan, const double* u, const double* v, const double* w, double* us, double* vs,
                     double* ws, double sigma, unsigned seed, const int* d_step) {
    if (!plan || !u || !v || !w || !us || !vs || !ws || !d_step) return EDDY_ERR_ARG;
    if (us == u || vs == v || ws == w) {
        set_error("eddy_perturb_dev: the predicted velocities must not alias the inputs");
        return EDDY_ERR_ARG;
    }
    launch_perturb(plan->stream, plan->d, u, v, w, us, vs, ws, sigma, (uint32_t)seed, d_step);
    return check_launch("perturb");
}

int eddy_resident_create(eddy_resident_t* res, eddy_plan_t plan, double* u, double* v, double* w, double* dp,
                         double alfxdt, double sigma, unsigned seed, int precision, int sweeps, int max_steps) {
    if (!res || !plan || !u || !v || !w || !dp || max_steps < 1 || !(alfxdt > 0.0)) return EDDY_ERR_ARG;
    if (precision != EDDY_FP64 && precision != EDDY_FP32_REFINE) return EDDY_ERR_ARG;
    if (!plan->stream) {
        set_error("the resident loop is a captured graph: give the plan a non-default stream");
        return EDDY_ERR_ARG;
    }
    if (plan->timing) {
        set_error("disable kernel timing before capturing the resident loop");
        return EDDY_ERR_ARG;
    }
    std::unique_ptr<eddy_resident_s, int (*)(eddy_resident_s*)> R(new eddy_resident_s(), resident_free);
    R->plan = plan;
    R->cap = max_steps;
    const size_t n = field_elems(plan);
    HIP_TRY(hipMalloc(&R->star, 3 * n * sizeof(double)));
    HIP_TRY(hipMalloc(&R->log, 2 * (size_t)max_steps * sizeof(double)));
    HIP_TRY(hipMalloc(&R->step, sizeof(int)));
    HIP_TRY(hipMemset(R->step, 0, sizeof(int)));
    HIP_TRY(hipMemset(R->log, 0, 2 * (size_t)max_steps * sizeof(double)));
    int rc = ensure_rhs(plan);
    if (rc) return rc;
    HIP_TRY(hipDeviceSynchronize());
    double* us = R->star;
    double* vs = us + n;
    double* ws = vs + n;
    HIP_TRY(hipStreamBeginCapture(plan->stream, hipStreamCaptureModeThreadLocal));
    launch_perturb(plan->stream, plan->d, u, v, w, us, vs, ws, sigma, (uint32_t)seed, R->step);
    rc = eddy_project_dev(plan, us, vs, ws, u, v, w, dp, alfxdt, 1, precision, sweeps);
    if (!rc) {
        launch_div_minmax(plan->stream, plan->d, plan->mt, u, v, w, rd(plan));
        launch_log_step(plan->stream, plan->red, R->log, R->step, R->cap);
    }
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(plan->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return device_error(e, "hipStreamEndCapture");
    R->graph = g;
    HIP_TRY(hipGraphInstantiate(&R->exec, g, nullptr, nullptr, 0));
    *res = R.release();
    return 0;
}

int eddy_resident_run(eddy_resident_t res, int nsteps) {
    if (!res || nsteps < 0) return EDDY_ERR_ARG;
    if (nsteps > res->cap - res->enqueued) {
        set_error("eddy_resident_run: " + std::to_string(res->enqueued) + " + " + std::to_string(nsteps) +
                  " steps exceed max_steps " + std::to_string(res->cap));
        return EDDY_ERR_ARG;
    }
    for (int s = 0; s < nsteps; ++s) HIP_TRY(hipGraphLaunch(res->exec, res->plan->stream));
    res->enqueued += nsteps;
    return 0;
}

int eddy_resident_log(eddy_resident_t res, double* minmax, int capacity, int* steps) {
    if (!res || (capacity > 0 && !minmax)) return EDDY_ERR_ARG;
    HIP_TRY(hipStreamSynchronize(res->plan->stream));
    int s = 0;
    HIP_TRY(hipMemcpy(&s, res->step, sizeof(int), hipMemcpyDeviceToHost));
    const int k = std::min(std::min(s, capacity), res->cap);
    if (k > 0) HIP_TRY(hipMemcpy(minmax, res->log, 2 * (size_t)k * sizeof(double), hipMemcpyDeviceToHost));
    if (steps) *steps = s;
    return 0;
}

int eddy_resident_destroy(eddy_resident_t res) {
    if (!res) return 0;
    if (res->plan && res->plan->stream) (void)hipStreamSynchronize(res->plan->stream);
    return resident_free(res);
}

// ---------------------------------------------------------------- batches
struct eddy_batch_s {
    std::vector<eddy_plan_s*> plans;
    hipStream_t stream = nullptr;
    std::vector<void*> owned;
    bool bound = false, has_vel = false, has_f32 = false;
    std::vector<BatchProblem<double>> p64;
    std::vector<BatchProblem<float>> p32;
    StageLaunch d_yf, d_zf, d_xs, d_zi, d_yi;              // FP64 DIRECT
    StageLaunch d_yf_div, d_zf_shift, d_cells_corr;        // FP64 projection
    StageLaunch f_copy, f_yf, f_zf, f_xs, f_zi, f_yi_store, f_yf_res, f_yi_upd, f_shift;  // FP32 + refine
    FoldJob* fold_div = nullptr;   // per problem: VOLB, VOL
    FoldJob* fold_phi = nullptr;   // per problem: PHISUM
    // {ALFXDT, 1/ALFXDT} of the projection, written by a fill kernel before each launch and
    // read by the kernels: one captured graph (branches) or one set of job arrays (one launch per stage)
    // serves every time step
    double* dscale = nullptr;
    long cells = 0, items = 0;
    // branch mode: each problem's own plan pipeline, problems spread over parallel branches of
    // one captured graph (instantiated per operation and cached until the next bind)
    int mode = EDDY_BATCH_BRANCHES;
    hipStream_t cap = nullptr;  // capture origin (the caller's stream may be the null stream)
    std::vector<hipStream_t> side;
    std::vector<int> order;  // problems by decreasing cell count
    std::vector<double*> dp, uo, vo, wo;
    std::vector<const double*> us, vs, ws;
    struct CachedGraph {
        int op, precision, sweeps, remove_mean;
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    std::vector<CachedGraph> graphs;
};

static void drop_graphs(eddy_batch_s* B) {
    for (auto& g : B->graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
    }
    B->graphs.clear();
}

static void free_batch(eddy_batch_s* B) {
    if (!B) return;
    drop_graphs(B);
    for (hipStream_t st : B->side) (void)hipStreamDestroy(st);
    if (B->cap) (void)hipStreamDestroy(B->cap);
    for (auto* P : B->plans) free_plan(P);
    for (void* p : B->owned) (void)hipFree(p);
    delete B;
}

static constexpr int kBatchBranches = 8;

int eddy_batch_create(eddy_batch_t* batch, int nprob, const int* nx, const int* ny, const int* nz, const double* delysq,
                      const double* const* ap, const double* const* au, const double* const* bp,
                      const double* const* bv, const double* const* cp, const double* const* cw,
                      const double* const* rp, const double* const* ru, const eddy_tuning* tuning, void* stream) {
    if (!batch || nprob < 1 || !nx || !ny || !nz || !delysq || !ap || !au || !cp || !cw || !rp || !ru) {
        set_error("bad batch arguments");
        return EDDY_ERR_ARG;
    }
    std::unique_ptr<eddy_batch_s, void (*)(eddy_batch_s*)> B(new eddy_batch_s(), free_batch);
    B->stream = static_cast<hipStream_t>(stream);
    for (int p = 0; p < nprob; ++p) {
        eddy_plan_s* P = nullptr;
        int rc = create_plan(&P, nx[p], ny[p], nz[p], delysq[p], ap[p], au[p], bp ? bp[p] : nullptr,
                             bv ? bv[p] : nullptr, cp[p], cw[p], rp[p], ru[p], 1, 0, 1, tuning ? &tuning[p] : nullptr,
                             stream, true);
        if (rc) {
            set_error("problem " + std::to_string(p) + ": " + g_err);
            return rc;
        }
        if (P->fpy.n == 0) {  // the batched kernels carry the radix FFT only
            free_plan(P);
            set_error("problem " + std::to_string(p) + ": batches need an even L whose half factors into 2, 3, 5");
            return EDDY_ERR_UNSUPPORTED;
        }
        B->plans.push_back(P);
        B->cells += (long)P->d.M * P->d.L * P->d.N;
    }
    void* c = nullptr;
    HIP_TRY(hipMalloc(&c, 2 * sizeof(double)));
    B->owned.push_back(c);
    B->dscale = static_cast<double*>(c);
    HIP_TRY(hipStreamCreateWithFlags(&B->cap, hipStreamNonBlocking));
    const int nside = std::min(nprob, kBatchBranches);
    for (int k = 0; k < nside; ++k) {
        hipStream_t st = nullptr;
        HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));  // default priority
        B->side.push_back(st);
    }
    B->order.resize(nprob);
    for (int p = 0; p < nprob; ++p) B->order[p] = p;
    std::stable_sort(B->order.begin(), B->order.end(), [&](int x, int y) {
        const Dims &a = B->plans[x]->d, &b = B->plans[y]->d;
        return (long)a.M * a.L * a.N > (long)b.M * b.L * b.N;
    });
    *batch = B.release();
    return 0;
}

int eddy_batch_set_mode(eddy_batch_t B, int mode) {
    if (!B || (mode != EDDY_BATCH_BRANCHES && mode != EDDY_BATCH_QUEUES)) return EDDY_ERR_ARG;
    B->mode = mode;
    return 0;
}

// Point a plan's launches (its kernels and its rocBLAS products) at stream st.
static int plan_use_stream(eddy_plan_s* P, hipStream_t st) {
    P->stream = st;
    if (P->blas && zgemm_set_stream(P->blas, st)) {
        set_error("rocBLAS refused a stream of the batch");
        return EDDY_ERR_DEVICE;
    }
    return 0;
}

// Capture op (0: DIRECT, 1: projection) of every problem on parallel branches: problem r of the
// size order goes to branch r mod K, each branch runs its problems' own plan pipelines in turn.
// The projection reads ALFXDT from B->dscale, so one graph per (op, precision, sweeps,
// remove_mean) serves every time step.
static int batch_branch_graph(eddy_batch_s* B, int op, int precision, int sweeps, int remove_mean,
                              hipGraphExec_t* out) {
    for (auto& g : B->graphs)
        if (g.op == op && g.precision == precision && g.sweeps == sweeps && g.remove_mean == remove_mean) {
            *out = g.exec;
            return 0;
        }
    const int K = (int)B->side.size();
    std::vector<hipEvent_t> ev(K + 1, nullptr);
    for (auto& e : ev) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    auto destroy_events = [&]() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
    };
    hipError_t e = hipStreamBeginCapture(B->cap, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        destroy_events();
        return device_error(e, "hipStreamBeginCapture (batch)");
    }
    int rc = 0;
    (void)hipEventRecord(ev[K], B->cap);
    for (int k = 0; k < K; ++k) (void)hipStreamWaitEvent(B->side[k], ev[K], 0);
    for (size_t r = 0; r < B->order.size() && !rc; ++r) {
        const int p = B->order[r];
        eddy_plan_s* P = B->plans[p];
        hipStream_t keep = P->stream;
        rc = plan_use_stream(P, B->side[r % K]);
        if (!rc) {
            if (op == 0) {
                rc = (precision == EDDY_FP64) ? run_direct64(P, B->dp[p], 0)
                                              : run_refined(P, B->dp[p], sweeps, 0, 0);
            } else {
                rc = project_core(P, B->us[p], B->vs[p], B->ws[p], B->uo[p], B->vo[p], B->wo[p], B->dp[p], 1.0,
                                  B->dscale, remove_mean, precision, sweeps);
            }
        }
        const int rc2 = plan_use_stream(P, keep);
        if (!rc) rc = rc2;
    }
    for (int k = 0; k < K; ++k) {
        (void)hipEventRecord(ev[k], B->side[k]);
        (void)hipStreamWaitEvent(B->cap, ev[k], 0);
    }
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(B->cap, &g);
    destroy_events();
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return device_error(e, "hipStreamEndCapture (batch)");
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return device_error(e, "hipGraphInstantiate (batch)");
    }
    B->graphs.push_back({op, precision, sweeps, remove_mean, g, x});
    *out = x;
    return 0;
}

int eddy_batch_bind(eddy_batch_t B, double* const* dp, const double* const* us, const double* const* vs,
                    const double* const* ws, double* const* uo, double* const* vo, double* const* wo) {
    if (!B || !dp) return EDDY_ERR_ARG;
    const int n = (int)B->plans.size();
    B->has_vel = us && vs && ws && uo && vo && wo;
    drop_graphs(B);
    B->dp.assign(dp, dp + n);
    B->us.assign(n, nullptr);
    B->vs.assign(n, nullptr);
    B->ws.assign(n, nullptr);
    B->uo.assign(n, nullptr);
    B->vo.assign(n, nullptr);
    B->wo.assign(n, nullptr);
    if (B->has_vel) {
        B->us.assign(us, us + n);
        B->vs.assign(vs, vs + n);
        B->ws.assign(ws, ws + n);
        B->uo.assign(uo, uo + n);
        B->vo.assign(vo, vo + n);
        B->wo.assign(wo, wo + n);
    }
    B->p64.clear();
    B->p32.clear();
    for (int p = 0; p < n; ++p) {
        eddy_plan_s* P = B->plans[p];
        int rc = ensure_rhs(P);
        if (rc) return rc;
        BatchProblem<double> b = problem_of<double>(P, P->t64);
        b.dp = dp[p];
        b.dscale = B->dscale;
        if (B->has_vel) {
            b.us = us[p];
            b.vs = vs[p];
            b.ws = ws[p];
            b.uo = uo[p];
            b.vo = vo[p];
            b.wo = wo[p];
        }
        B->p64.push_back(b);
        BatchProblem<float> f = problem_of<float>(P, P->t32);
        f.dp = dp[p];
        f.rhs = P->rhs;
        B->p32.push_back(f);
    }
    std::vector<void*>& own = B->owned;
    int rc = 0;
    if ((rc = prepare_yfft<double>(B->p64, 1, YF_DIRECT, B->d_yf, own)) ||
        (rc = prepare_ztrans<double>(B->p64, 0, 0, B->d_zf, own)) || (rc = prepare_xsolve<double>(B->p64, B->d_xs, own)) ||
        (rc = prepare_ztrans<double>(B->p64, 1, 0, B->d_zi, own)) ||
        (rc = prepare_yfft<double>(B->p64, 0, YI_DIRECT, B->d_yi, own)))
        return rc == -2 ? (set_error("x1 solve: no thread partition for some M"), EDDY_ERR_UNSUPPORTED)
                        : (set_error("batch preparation failed"), EDDY_ERR_DEVICE);
    if (B->has_vel) {
        if ((rc = prepare_yfft<double>(B->p64, 1, YF_DIVERGENCE, B->d_yf_div, own)) ||
            (rc = prepare_ztrans<double>(B->p64, 0, 1, B->d_zf_shift, own)) ||
            (rc = prepare_cells(B->p64, BC_CORRECT, B->d_cells_corr, own)))
            return (set_error("batch preparation failed"), EDDY_ERR_DEVICE);
    }
    if ((rc = prepare_cells_f(B->p32, BC_COPY_RHS, B->f_copy, own)) ||
        (rc = prepare_yfft<float>(B->p32, 1, YF_DIRECT, B->f_yf, own)) ||
        (rc = prepare_ztrans<float>(B->p32, 0, 0, B->f_zf, own)) || (rc = prepare_xsolve<float>(B->p32, B->f_xs, own)) ||
        (rc = prepare_ztrans<float>(B->p32, 1, 0, B->f_zi, own)) ||
        (rc = prepare_yfft<float>(B->p32, 0, YI_STORE, B->f_yi_store, own)) ||
        (rc = prepare_yfft<float>(B->p32, 1, YF_RESIDUAL, B->f_yf_res, own)) ||
        (rc = prepare_yfft<float>(B->p32, 0, YI_UPDATE, B->f_yi_upd, own)) ||
        (rc = prepare_cells_f(B->p32, BC_SHIFT_GHOSTS, B->f_shift, own)))
        return (set_error("batch preparation failed"), EDDY_ERR_DEVICE);
    // fold jobs: the y-stage partials of each problem (one slot per y tile)
    std::vector<FoldJob> fd, fp;
    for (int p = 0; p < n; ++p) {
        eddy_plan_s* P = B->plans[p];
        const int ntiles = ceil_div(P->d.M, P->tune.yfft_pencils) * P->d.N;
        fd.push_back(FoldJob{P->part, ntiles, RED_OP_SUM, P->red + RED_VOLB});
        fd.push_back(FoldJob{P->vol_total_dev, 1, RED_OP_SUM, P->red + RED_VOL});
        fp.push_back(FoldJob{P->part, ntiles, RED_OP_SUM, P->red + RED_PHISUM});
    }
    void* q = nullptr;
    HIP_TRY(hipMalloc(&q, (fd.size() + fp.size()) * sizeof(FoldJob)));
    own.push_back(q);
    B->fold_div = static_cast<FoldJob*>(q);
    B->fold_phi = B->fold_div + fd.size();
    HIP_TRY(hipMemcpy(B->fold_div, fd.data(), fd.size() * sizeof(FoldJob), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(B->fold_phi, fp.data(), fp.size() * sizeof(FoldJob), hipMemcpyHostToDevice));
    B->items = std::max({(long)B->d_yf.total, (long)B->d_zf.total, (long)B->d_xs.total});
    B->bound = true;
    return 0;
}

int eddy_batch_direct_dev(eddy_batch_t B, int precision, int sweeps) {
    if (!B || !B->bound || sweeps < 0 || sweeps > 16) {
        set_error("batch not bound or sweeps outside 0..16");
        return EDDY_ERR_ARG;
    }
    if (precision != EDDY_FP64 && precision != EDDY_FP32_REFINE) return EDDY_ERR_ARG;
    if (B->mode == EDDY_BATCH_BRANCHES) {
        hipGraphExec_t x = nullptr;
        int rc = batch_branch_graph(B, 0, precision, precision == EDDY_FP64 ? 0 : sweeps, 0, &x);
        if (rc) return rc;
        HIP_TRY(hipGraphLaunch(x, B->stream));
        return 0;
    }
    hipStream_t st = B->stream;
    if (precision == EDDY_FP64) {
        batch_yfft<double>(st, B->d_yf, 1, YF_DIRECT);
        batch_ztrans<double>(st, B->d_zf, 0);
        batch_xsolve<double>(st, B->d_xs);
        batch_ztrans<double>(st, B->d_zi, 1);
        batch_yfft<double>(st, B->d_yi, 0, YI_DIRECT);
        return check_launch("batched FP64 solve");
    }
    const int n = (int)B->plans.size();
    batch_cells(st, B->f_copy, BC_COPY_RHS);
    batch_yfft<float>(st, B->f_yf, 1, YF_DIRECT);
    batch_ztrans<float>(st, B->f_zf, 0);
    batch_xsolve<float>(st, B->f_xs);
    batch_ztrans<float>(st, B->f_zi, 1);
    batch_yfft<float>(st, B->f_yi_store, 0, YI_STORE);
    batch_fold(st, B->fold_phi, n);
    for (int k = 0; k < sweeps; ++k) {
        batch_yfft<float>(st, B->f_yf_res, 1, YF_RESIDUAL);
        batch_ztrans<float>(st, B->f_zf, 0);
        batch_xsolve<float>(st, B->f_xs);
        batch_ztrans<float>(st, B->f_zi, 1);
        batch_yfft<float>(st, B->f_yi_upd, 0, YI_UPDATE);
        batch_fold(st, B->fold_phi, n);
    }
    batch_cells(st, B->f_shift, BC_SHIFT_GHOSTS);
    return check_launch("batched FP32 + refine solve");
}

int eddy_batch_project_dev(eddy_batch_t B, double alfxdt, int remove_mean) {
    if (!B || !B->bound || !B->has_vel || alfxdt == 0.0) {
        set_error("batch not bound with velocities, or zero ALFXDT");
        return EDDY_ERR_ARG;
    }
    // this call's ALFXDT, read by the divergence and correction kernels of either mode
    launch_fill(B->stream, B->dscale, 1, alfxdt);
    launch_fill(B->stream, B->dscale + 1, 1, 1.0 / alfxdt);
    if (B->mode == EDDY_BATCH_BRANCHES) {
        hipGraphExec_t x = nullptr;
        int rc = batch_branch_graph(B, 1, EDDY_FP64, 0, remove_mean ? 1 : 0, &x);
        if (rc) return rc;
        HIP_TRY(hipGraphLaunch(x, B->stream));
        return 0;
    }
    hipStream_t st = B->stream;
    batch_yfft<double>(st, B->d_yf_div, 1, YF_DIVERGENCE);
    batch_fold(st, B->fold_div, 2 * (int)B->plans.size());
    batch_ztrans<double>(st, remove_mean ? B->d_zf_shift : B->d_zf, 0);
    batch_xsolve<double>(st, B->d_xs);
    batch_ztrans<double>(st, B->d_zi, 1);
    batch_yfft<double>(st, B->d_yi, 0, YI_DIRECT);
    batch_cells(st, B->d_cells_corr, BC_CORRECT);
    return check_launch("batched projection");
}

int eddy_batch_info(eddy_batch_t B, long* cells, long* items) {
    if (!B) return EDDY_ERR_ARG;
    if (cells) *cells = B->cells;
    if (items) *items = B->items;
    return 0;
}

int eddy_batch_destroy(eddy_batch_t B) {
    if (!B) return 0;
    if (B->stream) (void)hipStreamSynchronize(B->stream);
    else (void)hipDeviceSynchronize();
    free_batch(B);
    return 0;
}

// ---------------------------------------------------------------- timing API
int eddy_plan_enable_timing(eddy_plan_t plan, int enable) {
    if (!plan) return EDDY_ERR_ARG;
    plan->timing = enable ? 1 : 0;
    return 0;
}

int eddy_plan_kernel_time(eddy_plan_t plan, int kernel_id, long* launches, double* total_ms) {
    if (!plan || kernel_id < 0 || kernel_id >= EDDY_K_COUNT) return EDDY_ERR_ARG;
    drain_timing(plan);
    if (launches) *launches = plan->slots[kernel_id].launches;
    if (total_ms) *total_ms = plan->slots[kernel_id].total_ms;
    return 0;
}

int eddy_plan_reset_timing(eddy_plan_t plan) {
    if (!plan) return EDDY_ERR_ARG;
    drain_timing(plan);
    for (auto& s : plan->slots) s = TimerSlot{};
    return 0;
}

// ---------------------------------------------------------------- DIRECT drop-in
int eddy_direct(double* dp, double* dummy, double* dummyt, const double* delysq, const double* delzsq,
                const double* ap, const double* au, const double* cp, const double* cw, const double* rp,
                const double* ru, const int* nx, const int* ny, const int* nz, const int* nzg, const int* mysize,
                const int* myrank, const int* mpi_comm_eddy, const int* mtype, const int* ipres, const int* mp,
                const int* lp, const int* np) {
    (void)dummy;
    (void)dummyt;
    (void)myrank;
    (void)mpi_comm_eddy;
    (void)mtype;
    if (!dp || !delysq || !ap || !au || !cp || !cw || !rp || !ru || !nx || !ny || !nz || !nzg || !mysize || !ipres ||
        !mp || !lp || !np) {
        set_error("null argument to DIRECT");
        return EDDY_ERR_ARG;
    }
    if (*mysize != 1) {
        set_error("MYSIZE must be 1: one GPU solves the whole field");
        return EDDY_ERR_MPI;
    }
    if (*nz != *nzg) {
        set_error("NZ must equal NZG on one rank");
        return EDDY_ERR_ARG;
    }
    if (*ipres < 1 || *ipres > 6 || *ipres == 2) {
        set_error("IPRES must be 1, 3, 4, 5 or 6 (the GENBUN branch, 2, is not wired up in direct.f: its RHS "
                  "array is never filled)");
        return EDDY_ERR_UNSUPPORTED;
    }
    if (*ipres >= 4 && (*mp == 0 || *np == 0)) {
        set_error("IPRES 4-6 (pdc2dn) with MP=0 or NP=0: the reference's branch drops the periodic couplings "
                  "(direct.f:344-367); use IPRES=1 (BLKTRI's periodic system) or 3");
        return EDDY_ERR_UNSUPPORTED;
    }
    const int NX = *nx, NY = *ny, NZ = *nz;
    // IPRES=3 (direct.f:834-869, 560-580): an FFT along a uniform periodic x2 whose spacing is
    // sqrt(DELZSQ) (CP, CW and NP are not read), TRIDAG or CYCLIC along x1
    int NP = *np;
    std::vector<double> cz;
    if (*ipres == 3) {
        if (!delzsq || !(*delzsq > 0.0)) {
            set_error("IPRES=3 needs DELZSQ > 0");
            return EDDY_ERR_ARG;
        }
        cz.assign(NZ, 1.0 / std::sqrt(*delzsq));
        cp = cz.data();
        cw = cz.data();
        NP = 0;
    }
    int rc = validate_dims(NX, NY, NZ, *mp, *lp, NP);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(g_cache.mu);
    uint64_t key = 1469598103934665603ull;
    const int dims[6] = {NX, NY, NZ, *mp, *lp, NP};
    key = fnv(key, dims, sizeof(dims));
    key = fnv(key, delysq, sizeof(double));
    key = fnv(key, ap, NX * sizeof(double));
    key = fnv(key, au, NX * sizeof(double));
    key = fnv(key, rp, NX * sizeof(double));
    key = fnv(key, ru, NX * sizeof(double));
    key = fnv(key, cp, NZ * sizeof(double));
    key = fnv(key, cw, NZ * sizeof(double));
    if (!g_cache.plan || g_cache.key != key) {
        if (g_cache.plan) {
            free_plan(g_cache.plan);
            g_cache.plan = nullptr;
        }
        eddy_plan_s* P = nullptr;
        rc = create_plan(&P, NX, NY, NZ, *delysq, ap, au, nullptr, nullptr, cp, cw, rp, ru, *mp, *lp, NP, nullptr,
                         nullptr);
        if (rc) return rc;
        g_cache.plan = P;
        g_cache.key = key;
    }
    eddy_plan_s* P = g_cache.plan;
    const size_t n = field_elems(P);
    if (!P->dp_dev) {
        void* p = nullptr;
        HIP_TRY(hipMalloc(&p, n * sizeof(double)));
        P->owned.push_back(p);
        P->dp_dev = static_cast<double*>(p);
    }
    if (!g_cache.stage) {
        HostStage* sg = nullptr;
        try {
            sg = new HostStage();  // starts the copy threads
        } catch (const std::exception& ex) {
            set_error(std::string("DIRECT: host staging threads: ") + ex.what());
            return EDDY_ERR_DEVICE;
        }
        const hipError_t e = sg->init();
        if (e != hipSuccess) {
            delete sg;
            return device_error(e, "pinned staging of DP");
        }
        g_cache.stage = sg;
    }
    HIP_TRY(g_cache.stage->to_device(P->dp_dev, dp, n, P->stream));
    // The reference solves whatever it is given; here the RHS is checked on the device first:
    // non-finite values stop the call with DP untouched, and an incompatible RHS (its
    // volume-weighted mean not at roundoff) is solved as any other but reported.
    double rel = 0.0;
    long nonfinite = 0;
    if ((rc = rhs_check(P, P->dp_dev, &rel, &nonfinite))) return rc;
    if (nonfinite) {
        set_error("DIRECT: the RHS holds " + std::to_string(nonfinite) + " non-finite interior values");
        return EDDY_ERR_NUMERIC;
    }
    rc = run_direct64(P, P->dp_dev, 0);
    if (rc) return rc;
    HIP_TRY(g_cache.stage->to_host(dp, P->dp_dev, n, P->stream));
    if (rel > kCompatTol) {
        char msg[200];
        std::snprintf(msg, sizeof msg,
                      "DIRECT: incompatible RHS, volume-weighted mean %.3e of its magnitude; solved with the "
                      "singular mode pinned, the residual carries the incompatible part", rel);
        set_error(msg);
        return EDDY_ERR_INCOMPATIBLE;
    }
    return 0;
}

int eddy_rhs_check_dev(eddy_plan_t plan, const double* d_dp, double* rel_mean, long* nonfinite) {
    if (!plan || !d_dp || !rel_mean || !nonfinite) {
        set_error("null plan, DP or output");
        return EDDY_ERR_ARG;
    }
    return rhs_check(plan, d_dp, rel_mean, nonfinite);
}

int eddy_direct_release(void) {
    std::lock_guard<std::mutex> lock(g_cache.mu);
    if (g_cache.plan) {
        (void)hipDeviceSynchronize();
        free_plan(g_cache.plan);
        g_cache.plan = nullptr;
    }
    delete g_cache.stage;
    g_cache.stage = nullptr;
    return 0;
}

}  // extern "C"
