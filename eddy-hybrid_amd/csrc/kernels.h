// kernels.h -- host-side launchers of the projection kernels (one per sweep).
#pragma once

#include <algorithm>
#include <vector>

#include "common.h"

namespace eddy {

// FP64 coefficients of the 7-point operator DIRECT inverts (direct.f:344-352, 370-378, 536),
// used by the residual of the defect correction and by eddy_laplacian_dev.
struct Op64 {
    const double* am;
    const double* bm;
    const double* cm;
    const double* r2inv;
    const double* an;
    const double* bn;
    const double* cn;
    double inv_delysq;
};

// Metric arrays of the stencils (timeindeg.f:944-946, 1002-1022), device copies.
struct Metrics {
    const double *ap, *au, *rp, *ru;  // length NX
    const double *bp, *bv;            // length NY
    const double *cp, *cw;            // length NZ
    // cell volumes RP/(AP*BP*CP) as separable factors RP/AP, 1/BP, 1/CP, and their interior sum
    const double *volx, *voly, *volz;
    double vol_total;
    double bp0, vy0;  // BP(J) and 1/BP(J) when BP is the same for every J (uniform y), else 0
    int cart;  // RP = RU = 1 everywhere: the /RP and RU* factors are exact identities and skipped
};

// Reduction workspace: per-workgroup partials (plan-owned, >= 2 * max grid size doubles) and
// the FP64 scalars they fold into (indices RedSlot).
struct Reduce {
    double* part;
    double* red;
};

enum YfftFwdMode { YF_DIRECT = 0, YF_RESIDUAL = 1, YF_DIVERGENCE = 2 };

// What the forward y-FFT reads: YF_DIRECT the RHS in src (minus the solvability shift
// red[RED_VOLB]/red[RED_VOL] when shift != 0); YF_RESIDUAL r = (src - that shift when
// shift != 0) - L(phi) in FP64 (src is the RHS; sum r^2 -> red[RED_RNORM2]); YF_DIVERGENCE
// b = DIV(us, vs, ws) * scale formed on the fly (sum vol*b -> red[RED_VOLB], sum vol ->
// red[RED_VOL]; the solvability shift is then applied to the zero y-mode by the x2 transform;
// the FP32 transform also stores b in FP64 at rhs_out, the RHS of the defect correction).
struct FwdInput {
    int mode;
    const double* src;
    const double* phi;
    const double *us, *vs, *ws;
    Metrics mt;
    double scale;
    // when non-null: device {ALFXDT, 1/ALFXDT}, read at run time in place of `scale` (one
    // captured graph then serves every time step)
    const double* dscale;
    int shift;
    double* rhs_out;  // YF_DIVERGENCE of the FP32 transform: the FP64 RHS, DP layout (interior)
};
// YI_UPDATE_FINAL: the last update of the defect correction with its closing pass fused in:
// phi = phi + delta - red[RED_PHISUM]/(M L N) (the plain mean the previous pass summed) on the
// interior, the ghosts as DIRECT fills them and zero z ghost planes; no sum.
enum YfftInvMode { YI_DIRECT = 0, YI_STORE = 1, YI_UPDATE = 2, YI_UPDATE_FINAL = 3 };

// Forward y-FFT of every interior (i,k) pencil of the input described by `in` (FwdInput).
// Output: S[i + M*(k + N*j)] in the fftpack packing along j.
// iblk: tiles per x1 block of the tile order (0: whole rows), see iblock_tile.
template <typename T>
void launch_yfft_fwd(hipStream_t st, const Dims& d, const FftPlan& fp, const Tables<T>& tb, const FwdInput& in,
                     const Op64& op, const Reduce& rd, T* S, int TI, int iblk = 0);

// Inverse y-FFT (x 1/L) of S into DP layout. YI_DIRECT writes interior + ghosts like
// direct.f:1035-1106; YI_STORE writes the interior only and adds sum(phi) to
// red[RED_PHISUM]; YI_UPDATE adds into phi's interior and adds sum(new phi).
template <typename T>
void launch_yfft_inv(hipStream_t st, const Dims& d, const FftPlan& fp, const Tables<T>& tb,
                     const T* S, double* dp, const Reduce& rd, int TI, int mode);

// The y real DFT as FP32 matrix-core products (ygemm.hip), for L % 16 == 0, L <= 512:
// forward from the YF_DIRECT / YF_RESIDUAL input into S, inverse (YI_STORE / YI_UPDATE)
// from S into DP, with the same reductions as the FFT launchers.
bool ygemm_supported(const Dims& d);
void launch_ygemm_fwd(hipStream_t st, const Dims& d, const float* FT, const FwdInput& in, const Op64& op,
                      const Reduce& rd, float* S);
void launch_ygemm_inv(hipStream_t st, const Dims& d, const float* GT, const float* S, double* dp, const Reduce& rd,
                      int mode);

// Forward x2 transform (DCT-II when uniform, dense eigenvectors otherwise), in place on S.
// When shift_red != null, the zero y-mode first loses L * red[RED_VOLB]/red[RED_VOL] (the
// volume-weighted solvability shift of an RHS whose y-FFT was taken before its mean was known).
template <typename T>
void launch_ztrans_fwd(hipStream_t st, const Dims& d, const FftPlan& fpz, const Tables<T>& tb, T* S,
                       int TI, int dense, const double* shift_red, int j0, int nj);
// Inverse x2 transform; for the y-mode j=0 it first applies the plain-mean correction from
// the line sums (direct.f:894-898).
template <typename T>
void launch_ztrans_inv(hipStream_t st, const Dims& d, const FftPlan& fpz, const Tables<T>& tb, T* S,
                       const T* linesum, int TI, int dense, int j0, int nj);
// The dense x2 transform as strided-batched GEMMs (rocBLAS handle `blas` on the plan's
// stream), out of place: forward S -> S2 (with the zero-mode shift), inverse S2 -> S (with
// the plain-mean correction). Return 0, or -1 if rocBLAS refused the call.
template <typename T>
int launch_zgemm_fwd(hipStream_t st, void* blas, const Dims& d, const Tables<T>& tb, T* S, T* S2,
                     const double* shift_red, int j0, int nj);
template <typename T>
int launch_zgemm_inv(hipStream_t st, void* blas, const Dims& d, const Tables<T>& tb, T* S2, T* S, const T* linesum,
                     int j0, int nj);
template <typename T>
int launch_xgemm_fwd(hipStream_t st, void* blas, const Dims& d, const Tables<T>& tz, T* S, T* S2,
                     const double* shift_red, int j0, int nj);
template <typename T>
int launch_xgemm_inv(hipStream_t st, void* blas, const Dims& d, const Tables<T>& tz, T* S2, T* S, const T* linesum,
                     int null_x, int j0, int nj);
int zgemm_create(void** blas, hipStream_t st);
int zgemm_set_stream(void* blas, hipStream_t st);
void zgemm_destroy(void* blas);

// Most rows per thread of the x1 solve (its largest kernel template).
constexpr int kXsolveMaxChunk = 32;

// Threads per line P and rows per thread Q of the x1 solve for M rows and the tuning chunk
// (the plan lays the x1 coefficient tables out in this thread order).
bool xsolve_partition(int M, int chunk, int& P, int& Q);

// x1 tridiagonal solves of every (j, m) line, in place on S; line sums of j=0 -> linesum.
// cyc != 0: periodic lines (MP=0), row 0 coupled to x(M-1) by beta = AM(1) and row M-1 to
// x(0) by alpha = CM(M) (the tables hold 0 in those slots).
template <typename T>
int launch_xsolve(hipStream_t st, const Dims& d, const Tables<T>& tb, T* S, T* linesum, int chunk, int j0, int nj,
                  int cyc = 0, double alpha = 0.0, double beta = 0.0);

// ------------------------------------------------------------------ batched launches
// One problem of a batch, with everything its stages read and write (device pointers).
template <typename T>
struct BatchProblem {
    Dims d;
    FftPlan fpy, fpz;
    Tables<T> tb;
    Op64 op;
    Metrics mt;
    T* S;
    T* linesum;
    double* part;       // reduction partials (>= 2 x the largest stage's tile count)
    double* red;        // reduction scalars (RedSlot)
    double* dp;         // DP: RHS in / pressure out
    const double* rhs;  // FP64 RHS kept for the defect correction
    const double *us, *vs, *ws;
    double *uo, *vo, *wo;
    const double* dscale;  // device {ALFXDT, 1/ALFXDT} of the projection
    int tiy, tiz, dense, chunk;
};

// A prepared batched stage: the job array (device), its size, the item count and the LDS
// bytes the stage kernel needs.
constexpr int kMaxClasses = 4;  // kernel templates of a stage
struct StageLaunch {
    void* jobs = nullptr;
    int njobs = 0;
    int total = 0;
    size_t lds = 0;
    // launch classes (one launch each, in turn): class c holds jobs [cls_job[c], cls_job[c + 1])
    // with items numbered from 0, cls_total[c] items, kernel template cls_q[c] (the x1 solve's
    // rows per thread, the x2 transform's kind)
    int ncls = 0;
    int cls_q[kMaxClasses] = {};
    int cls_job[kMaxClasses + 1] = {};
    int cls_total[kMaxClasses] = {};
    size_t cls_lds[kMaxClasses] = {};
    // the forward y transform with the divergence: 1 every problem Cartesian, 2 every one
    // cylindrical (a kernel with only that metric's registers), 0 mixed
    int kind = 0;
};

// Workgroups of 256 threads that one CU holds at once when each takes `lds` bytes of its
// 160 KiB of LDS (at most 8: 32 waves per CU).
inline int lds_occupancy(size_t lds) {
    return lds == 0 ? 8 : (int)std::max<size_t>(1, std::min<size_t>(8, ((size_t)160 << 10) / lds));
}

// Appends a job of kernel template `tmpl` whose tile takes `lds` bytes to a stage split into
// one launch class per template; jobs arrive grouped by template. (Classes split further by
// LDS occupancy, so that small tiles are not held to the occupancy of the largest, measured
// 1-3 % slower on uneven batches: the extra launches' tails cost more.) False when the stage
// has more templates than kMaxClasses.
template <typename J>
bool add_class_job(StageLaunch& out, std::vector<J>& jobs, J j, int tmpl, size_t lds) {
    int c = out.ncls - 1;
    if (c < 0 || out.cls_q[c] != tmpl) {
        if (out.ncls == kMaxClasses) return false;
        c = out.ncls++;
        out.cls_q[c] = tmpl;
        out.cls_job[c] = (int)jobs.size();
    }
    j.first = out.cls_total[c];
    out.cls_total[c] += j.ntiles;
    out.total += j.ntiles;
    out.cls_lds[c] = std::max(out.cls_lds[c], lds);
    out.lds = std::max(out.lds, lds);
    jobs.push_back(j);
    out.cls_job[out.ncls] = (int)jobs.size();
    return true;
}

// Uploads a stage's job array (synchronous copy; the buffer joins `owned`).
template <typename J>
int upload_jobs(const std::vector<J>& jobs, StageLaunch& out, std::vector<void*>& owned) {
    void* buf = nullptr;
    if (hipMalloc(&buf, std::max<size_t>(1, jobs.size()) * sizeof(J)) != hipSuccess) return -1;
    owned.push_back(buf);
    if (!jobs.empty() && hipMemcpy(buf, jobs.data(), jobs.size() * sizeof(J), hipMemcpyHostToDevice) != hipSuccess)
        return -1;
    out.jobs = buf;
    out.njobs = (int)jobs.size();
    return 0;
}

// Each prepare_* builds and uploads the stage's job array (owned by the caller through
// `owned`), problems longest first; each batch_* enqueues ONE launch with a workgroup per
// work item of every problem (the x1 solve and the x2 transforms: one per template class,
// in turn on the stream).
template <typename T>
int prepare_yfft(const std::vector<BatchProblem<T>>& pr, int fwd, int mode, StageLaunch& out,
                 std::vector<void*>& owned);
template <typename T>
void batch_yfft(hipStream_t st, const StageLaunch& sl, int fwd, int mode);
template <typename T>
int prepare_ztrans(const std::vector<BatchProblem<T>>& pr, int dir, int shift, StageLaunch& out,
                   std::vector<void*>& owned);
template <typename T>
void batch_ztrans(hipStream_t st, const StageLaunch& sl, int dir);
template <typename T>
int prepare_xsolve(const std::vector<BatchProblem<T>>& pr, StageLaunch& out, std::vector<void*>& owned);
template <typename T>
void batch_xsolve(hipStream_t st, const StageLaunch& sl);
// cell-wise FP64 stages over all problems: 0 = CORRECT (+ periodic ghost refresh),
// 1 = phi -= mean, ghosts (finalize of the defect correction), 2 = rhs copy of dp
enum BatchCellOp { BC_CORRECT = 0, BC_SHIFT_GHOSTS = 1, BC_COPY_RHS = 2 };
int prepare_cells(const std::vector<BatchProblem<double>>& pr, int op, StageLaunch& out,
                  std::vector<void*>& owned);
int prepare_cells_f(const std::vector<BatchProblem<float>>& pr, int op, StageLaunch& out,
                    std::vector<void*>& owned);
void batch_cells(hipStream_t st, const StageLaunch& sl, int op);
// fold per-problem partials: one workgroup per (problem, quantity)
struct FoldJob {
    const double* part;
    int n;
    int op;
    double* out;
};
void batch_fold(hipStream_t st, const FoldJob* jobs, int njobs);

// Stencils (FP64).
int reduce_blocks(const Dims& d);
void launch_fold(hipStream_t st, const double* part, int n, double* out, int op, int add);
// red[RED_CHK_VB] = sum vol*b, red[RED_CHK_VABS] = sum vol*|b| over the finite interior values
// of b, red[RED_CHK_NF] = the count of non-finite ones.
void launch_rhs_check(hipStream_t st, const Dims& d, const Metrics& mt, const double* b, const Reduce& rd);
// dscale (both): null, or device {ALFXDT, 1/ALFXDT} read at run time in place of the value
// argument (the divergence takes 1/ALFXDT, the correction ALFXDT).
void launch_divergence(hipStream_t st, const Dims& d, const Metrics& mt, const double* us,
                       const double* vs, const double* ws, double* div, double scale, const Reduce& rd,
                       int accumulate_compat, const double* dscale = nullptr);
void launch_correct(hipStream_t st, const Dims& d, const Metrics& mt, const double* us, const double* vs,
                    const double* ws, double* uo, double* vo, double* wo, const double* dp, double alfxdt,
                    int refresh_periodic, const double* dscale = nullptr);
void launch_laplacian(hipStream_t st, const Dims& d, const Op64& op, const double* phi, double* out);
// Final pass of the defect correction: when rhs != null, red[RED_RNORM2] = sum (rhs - L phi)^2;
// then phi -= red[RED_PHISUM]/(M*L*N) on the interior and the ghosts as DIRECT fills them.
void launch_finalize(hipStream_t st, const Dims& d, const Op64& op, double* phi, const double* rhs,
                     const Reduce& rd);
void launch_sumsq(hipStream_t st, const Dims& d, const double* a, const Reduce& rd, int slot);
void launch_fill(hipStream_t st, double* p, size_t n, double v);
// REFRESHBC on one GPU (refreshbc.f:1-28): nothing with x2 walls; with a periodic x2 the z
// ghost planes take the periodic images DP(:,:,1) = DP(:,:,NZ-1), DP(:,:,NZ) = DP(:,:,2).
void launch_refresh_z(hipStream_t st, const Dims& d, double* dp);
void launch_copy_shift(hipStream_t st, const Dims& d, const double* src, double* dst, const double* red);
void launch_div_minmax(hipStream_t st, const Dims& d, const Metrics& mt, const double* us,
                       const double* vs, const double* ws, const Reduce& rd);

// Pressure update after CORRECT (Eddy6.f:2772-2846): P += DP minus the implicit corrections,
// PREF subtracted (red[RED_PREF]), ghost values from the boundary copies.
struct PressArgs {
    Dims d;
    Metrics mt;
    double* p;
    const double* dp;
    const double* rvx;  // RVIMPLX per x1 index (NX)
    const double* rvy;  // RVIMPLY per x1 index (NX)
    double* red;
    double alfxdt, ru1, dely;
    int implx, imply, mp, lp, itype1, itype5, itype6;
};
void launch_pressure_update(hipStream_t st, const PressArgs& a);

// Resident time loop: us = u + sigma * N(0,1) on the faces CORRECT updates (Philox4x32-10,
// counter {cell, step}, key {seed, 0}; step read from device memory); the log epilogue
// stores red[RED_DMIN], red[RED_DMAX] at log[2 step] and increments *step.
void launch_perturb(hipStream_t st, const Dims& d, const double* u, const double* v, const double* w, double* us,
                    double* vs, double* ws, double sigma, uint32_t seed, const int* step);
void launch_log_step(hipStream_t st, const double* red, double* log, int* step, int cap);

}  // namespace eddy
