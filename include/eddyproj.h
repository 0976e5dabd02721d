/*
 * eddyproj.h -- C-ABI of the MI355X pressure-projection library (libeddyproj.so).
 *
 * The library replaces the reference's projection path on one GPU:
 *   DIRECT      src/direct.f:2-3 (SUBROUTINE DIRECT, 23 arguments, called at src/Eddy6.f:2722)
 *   DIVERGENCE  src/timeindeg.f:908 (SUBROUTINE DIVERGENCE(US,VS,WS,DIV,IM,JM,KM))
 *   CORRECT     src/timeindeg.f:981 (SUBROUTINE CORRECT(US,VS,WS,UO,VO,WO,DP,IM,JM,KM,ALFXDT))
 *
 * Conventions (all entry points):
 *   - Fields use the reference's Fortran layout A(NX,NY,NZ): element (I,J,K), 1-based, at
 *     offset (I-1) + NX*((J-1) + NY*(K-1)). NX = M+2 (x1 walls), NY = L+2 (periodic FFT axis),
 *     NZ = N+2 (x2 walls); one ghost layer on each side.
 *   - Metric arrays are the reference's: AP, AU, RP, RU of length >= NX; BP, BV of length
 *     >= NY; CP, CW (CPG, CWG) of length >= NZ (grid.f:35-173, 454-478).
 *   - Every function returns an eddy_status (0 = success). Nothing aborts the process.
 *   - Supported boundary flags are the walls+periodic set used by the projection path:
 *     MP=1 (x walls), LP=0 (periodic y), NP=1 (z walls). Others return EDDY_ERR_UNSUPPORTED.
 *   - Device-pointer entry points (suffix _dev) take device memory and a hipStream_t passed
 *     as void*; they enqueue work and return without synchronizing.
 *   - Not thread-safe per plan: calls on one plan must not overlap.
 */
#ifndef EDDYPROJ_H
#define EDDYPROJ_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    EDDY_OK = 0,
    EDDY_ERR_ARG = -1,          /* invalid dimension, null pointer, bad flag value */
    EDDY_ERR_UNSUPPORTED = -2,  /* valid for the reference, not supported here (e.g. MP=0) */
    EDDY_ERR_DEVICE = -3,       /* HIP runtime error (message in eddy_last_error()) */
    EDDY_ERR_NUMERIC = -4,      /* eigen-decomposition failure, NaN/Inf detected, no convergence */
    EDDY_ERR_MPI = -5,          /* MYSIZE != 1: one GPU per solve only */
    EDDY_ERR_INCOMPATIBLE = -6, /* RHS fails the volume-weighted solvability check (strict mode) */
    EDDY_ERR_IO = -7            /* field file missing, short, or not in the expected record layout */
} eddy_status;

typedef enum { EDDY_FP64 = 0, EDDY_FP32_REFINE = 1 } eddy_precision;

typedef struct eddy_plan_s* eddy_plan_t;

/* Text of the last error on this thread (never NULL). */
const char* eddy_last_error(void);
/* Library version string. */
const char* eddy_version(void);

/* ---------------------------------------------------------------------------------------
 * Drop-in DIRECT (src/direct.f:2-3). Same 23 arguments, all by reference, as the Fortran
 * shim eddy-hybrid_amd/fortran/direct_hip.f90 passes them. DP (host) holds the RHS on the
 * interior on entry and the pressure correction on exit; x and y ghosts are filled and z
 * ghosts zeroed as direct.f:1035-1106 does. DUMMY/DUMMYT (scratch in the reference,
 * direct.f:85-86) are not touched. DELZSQ, MPI_COMM_EDDY and MTYPE are accepted and
 * ignored (one GPU, FP64). IPRES selects nothing: every IPRES value solves the same discrete
 * system (IPRES 2 and 3 are rejected as the reference cannot run them with Neumann walls).
 * A device plan is cached, keyed on the dimensions, flags and a hash of the metric arrays
 * (the reference SAVEs its plan on the first call, direct.f:121-123, and never rebuilds it).
 * The RHS is checked on the device first (eddy_rhs_check_dev), where the reference solves
 * whatever it is given: non-finite interior values return EDDY_ERR_NUMERIC with DP untouched;
 * an RHS whose volume-weighted mean exceeds 1e-10 of sum vol*|b| is solved as any other (the
 * singular mode pinned) and the call returns EDDY_ERR_INCOMPATIBLE with DP written, which
 * the shim prints and continues on, as the reference prints FISHPACK's IERROR.
 * DP travels through two 32 MiB pinned buffers kept with the cached plan, host threads
 * copying one chunk while the DMA engine moves the other.
 * ------------------------------------------------------------------------------------- */
int eddy_direct(double* dp, double* dummy, double* dummyt, const double* delysq,
                const double* delzsq, const double* ap, const double* au, const double* cp,
                const double* cw, const double* rp, const double* ru, const int* nx,
                const int* ny, const int* nz, const int* nzg, const int* mysize,
                const int* myrank, const int* mpi_comm_eddy, const int* mtype,
                const int* ipres, const int* mp, const int* lp, const int* np);

/* Release the cached plan of eddy_direct and its pinned staging (e.g. before the process
 * exits). */
int eddy_direct_release(void);

/* ---------------------------------------------------------------------------------------
 * Plans (explicit form of the SAVEd state of direct.f:121-585).
 * ------------------------------------------------------------------------------------- */
typedef struct {
    /* Zero fields take the defaults. The plan lowers the two tile widths until the in-place
       FFTs hold one butterfly per thread and pass (width * transform/2 / smallest radix <= 512)
       and the LDS fits; eddy_plan_get_tuning reports what the plan uses. */
    int yfft_pencils;   /* pencils (consecutive x1 lines) per workgroup in the y-FFT: power of 2
                           <= 64 (default: the smallest of 16, 32, 64 with pencils x L >= 4096,
                           at most the next power of 2 of M) */
    int ztrans_pencils; /* lines per workgroup in the x2 transform: power of 2, >= 4 (default 16) */
    int xsolve_chunk;   /* max rows per thread in the x1 solve: the PCR-to-Thomas switch (3..32,
                           default 12; EDDY_ERR_ARG outside) */
    int use_dense_z;    /* 1: force the dense x2 eigen-transform even on a uniform x2 grid */
    int fft_radix;      /* largest radix of the FFT radix plan: 8 (default), 4 or 2 */
    int yfft_gemm;      /* 1: FP32 y transforms of EDDY_FP32_REFINE as dense FP32 matrix-core
                           products instead of the radix FFT (L % 16 == 0, L <= 512) */
    int slab_chunks;    /* run the x2 transform, x1 solve and inverse x2 transform over this many
                           chunks of y-modes in turn (default 1) */
    int zgemm;          /* stretched-x2 transform as batched matrix products on the matrix cores
                           (out of place): 1 diagonalise x2 (N x N products per y-mode,
                           then x1 lines), 2 diagonalise x1 (M x M products per y-mode, then x2
                           lines; single plans), 0 auto (when max(M, N) >= 128: form 2 if M < N,
                           else form 1), -1 never (the fused tile kernel) */
    int yfft_iblock;    /* forward y transform tile order: blocks of this many x1 tiles walked
                           down x2 (x2-neighbour planes stay in L2); 0 = whole x1 rows */
    int zgemm_lib;      /* the GEMM forms' products: 0 the library's MFMA kernel (default),
                           1 rocBLAS strided-batched GEMMs */
} eddy_tuning;

/* Create a plan for DP(NX,NY,NZ) with the given metrics (host arrays, copied). Allocates
 * the device work buffers on the current device. stream may be NULL (default stream).
 * tuning may be NULL (defaults, or the autotuned table when one was recorded). */
int eddy_plan_create(eddy_plan_t* plan, int nx, int ny, int nz, double delysq,
                     const double* ap, const double* au, const double* bp, const double* bv,
                     const double* cp, const double* cw, const double* rp, const double* ru,
                     int mp, int lp, int np, const eddy_tuning* tuning, void* stream);
int eddy_plan_destroy(eddy_plan_t plan);
int eddy_plan_set_stream(eddy_plan_t plan, void* stream);
int eddy_plan_get_tuning(eddy_plan_t plan, eddy_tuning* out);
/* Bytes of device memory the plan holds. */
size_t eddy_plan_device_bytes(eddy_plan_t plan);

/* In-place FP64 DIRECT on device memory d_dp (NX*NY*NZ doubles). */
int eddy_direct_dev(eddy_plan_t plan, double* d_dp);

/* eddy_direct_dev without any allocation by the call: no counterpart in the reference. For
 * callers that cannot allocate device memory while they run (hipGraph capture, serving and
 * time-stepping loops that forbid hipMalloc). The work buffers (spectral array, line sums,
 * reduction partials) live in a fixed 64 MiB device array of the library, one per GPU, held
 * from library load to process exit; grids needing more (M*L*N + N + 2*M*N + 11 doubles above
 * 8,388,608, about 8.3 million cells) return EDDY_ERR_UNSUPPORTED, as do plans that take the
 * GEMM transform forms (stretched x2 with max(M, N) >= 128; tuning zgemm = -1 avoids them).
 * Every call on one GPU shares that array: calls that overlap on one GPU (other streams,
 * other threads) race and must be serialised by the caller. Same result as eddy_direct_dev. */
int eddy_direct_noalloc_dev(eddy_plan_t plan, double* d_dp);

/* FP32 transforms/eliminations + FP64 defect correction (7-point FP64 residual, FP64
 * update, FP64 plain-mean removal). d_dp: RHS in, pressure out (same contract as
 * eddy_direct_dev). Runs exactly `sweeps` correction sweeps after the first FP32 solve
 * (no host synchronisation, capturable in a graph). If resid_out is non-NULL, the final
 * relative residual ||b - L phi||_2 / ||b||_2 is computed on the device and copied back
 * (this synchronises the stream). */
int eddy_direct_refined_dev(eddy_plan_t plan, double* d_dp, int sweeps, double* resid_out);

/* Adaptive form: sweeps until ||b - L phi||_2/||b||_2 <= tol or max_sweeps; reports the
 * sweeps run and the final residual. Synchronises once per sweep. */
int eddy_direct_refined_adaptive_dev(eddy_plan_t plan, double* d_dp, double tol, int max_sweeps,
                                     int* sweeps_out, double* resid_out);

/* DIVERGENCE (timeindeg.f:908-975) with the DP=DP/ALFXDT scaling of Eddy6.f:2715 fused:
 * d_div(I,J,K) = DIV(us,vs,ws)(I,J,K) * scale on the interior; ghosts are not written. */
int eddy_divergence_dev(eddy_plan_t plan, const double* d_us, const double* d_vs,
                        const double* d_ws, double* d_div, double scale);

/* CORRECT (timeindeg.f:981-1030): uo/vo/wo = us/vs/ws - grad(dp)*alfxdt on the corrected
 * faces (walls excluded, setup.f:17-72); other entries of uo/vo/wo are not written. */
int eddy_correct_dev(eddy_plan_t plan, const double* d_us, const double* d_vs,
                     const double* d_ws, double* d_uo, double* d_vo, double* d_wo,
                     const double* d_dp, double alfxdt);

/* One fractional-step projection: DIVERGENCE -> RHS/alfxdt -> (volume-weighted mean
 * removal if remove_mean) -> DIRECT (FP64, or FP32+refine with `sweeps` corrections when
 * precision == EDDY_FP32_REFINE) -> CORRECT. d_dp receives the pressure correction. */
int eddy_project_dev(eddy_plan_t plan, const double* d_us, const double* d_vs,
                     const double* d_ws, double* d_uo, double* d_vo, double* d_wo,
                     double* d_dp, double alfxdt, int remove_mean, int precision, int sweeps);

/* Post-projection divergence report (SCRINFO, info.f:5-165): min and max of DIV(u) over the
 * interior (synchronises). */
int eddy_divergence_minmax_dev(eddy_plan_t plan, const double* d_us, const double* d_vs,
                               const double* d_ws, double* dmin, double* dmax);

/* Apply the operator DIRECT inverts to d_phi (interior) into d_out (interior). */
int eddy_laplacian_dev(eddy_plan_t plan, const double* d_phi, double* d_out);

/* Checks of an RHS d_dp on the interior (no counterpart in the reference, which checks
 * nothing at run time): *rel_mean = |sum vol*b| / sum vol*|b| over the finite values (the
 * Neumann solvability condition, vol = RP/(AP*BP*CP)), *nonfinite = the count of NaN/Inf
 * values. Synchronises the plan's stream. */
int eddy_rhs_check_dev(eddy_plan_t plan, const double* d_dp, double* rel_mean, long* nonfinite);

/* ---------------------------------------------------------------------------------------
 * Graph-captured projection (hipGraph of divergence -> solve -> correction).
 * ------------------------------------------------------------------------------------- */
typedef struct eddy_graph_s* eddy_graph_t;
int eddy_graph_create(eddy_graph_t* graph, eddy_plan_t plan, const double* d_us,
                      const double* d_vs, const double* d_ws, double* d_uo, double* d_vo,
                      double* d_wo, double* d_dp, double alfxdt, int remove_mean,
                      int precision, int sweeps);
int eddy_graph_launch(eddy_graph_t graph);
int eddy_graph_destroy(eddy_graph_t graph);

/* ---------------------------------------------------------------------------------------
 * Batches: many independent problems of uneven sizes. Each solve stage of every problem runs
 * either in the problems' own kernels on parallel graph branches, or in ONE launch with a
 * workgroup per work item of all problems (longest problem first). The launch sequence has
 * no host synchronisation and can be captured in a hipGraph by the caller.
 * ------------------------------------------------------------------------------------- */
typedef struct eddy_batch_s* eddy_batch_t;
/* Per-problem arrays of length nprob (dimensions, DELYSQ, metric array pointers as in
 * eddy_plan_create). tuning: NULL, or one eddy_tuning per problem. */
int eddy_batch_create(eddy_batch_t* batch, int nprob, const int* nx, const int* ny, const int* nz,
                      const double* delysq, const double* const* ap, const double* const* au,
                      const double* const* bp, const double* const* bv, const double* const* cp,
                      const double* const* cw, const double* const* rp, const double* const* ru,
                      const eddy_tuning* tuning, void* stream);
/* Bind the device fields of every problem (arrays of nprob device pointers). dp is required;
 * the velocity arrays may be NULL when only eddy_batch_direct_dev is used. Builds the job
 * arrays of the one-launch mode (synchronous host-to-device copies). */
int eddy_batch_bind(eddy_batch_t batch, double* const* dp, const double* const* us,
                    const double* const* vs, const double* const* ws, double* const* uo,
                    double* const* vo, double* const* wo);
/* DIRECT on every bound DP (FP64, or FP32 + `sweeps` FP64 defect corrections). */
int eddy_batch_direct_dev(eddy_batch_t batch, int precision, int sweeps);
/* FP64 projection of every problem: divergence/alfxdt (fused into the y-FFT), optional
 * volume-weighted mean removal, solve, correction (+ periodic y ghosts of the velocities). */
int eddy_batch_project_dev(eddy_batch_t batch, double alfxdt, int remove_mean);
/* Execution of the batched operations:
 *   EDDY_BATCH_BRANCHES (default): each problem's own single-problem kernels, problems spread
 *     (largest first) over 8 parallel branches of one hipGraph, captured on the first call of
 *     each operation and replayed after;
 *   EDDY_BATCH_QUEUES: one launch per stage (per template class for the x1 solve and the x2
 *     transforms) with a workgroup per work item of all problems. */
enum { EDDY_BATCH_BRANCHES = 0, EDDY_BATCH_QUEUES = 1 };
int eddy_batch_set_mode(eddy_batch_t batch, int mode);
/* Total interior cells and work items of the largest stage. */
int eddy_batch_info(eddy_batch_t batch, long* cells, long* items);
int eddy_batch_destroy(eddy_batch_t batch);

/* ---------------------------------------------------------------------------------------
 * Pressure update after CORRECT (Eddy6.f:2772-2846): P = P + DP; when IMPLY / IMPLX == 1 the
 * implicit-viscous corrections with RU1 and the RVIMPLY / RVIMPLX masks (:2777-2796);
 * PREF = sum(P(2, 2:NY-1, 2)) / (NY-2) subtracted everywhere (:2801-2806); ghost values by
 * the boundary copies in x (MP; ITYPE(1) = 300 axis through JSYM), y (LP) and x2 (ITYPE(5),
 * ITYPE(6): 500 periodic, other types Neumann) (:2813-2847). REFRESHBC(P) is a no-op on one
 * GPU.
 * ------------------------------------------------------------------------------------- */
/* Store the update's parameters in the plan (host masks of NX doubles, copied; NULL = 0).
 * Until called: no implicit corrections, wall at x1 = 1 (ITYPE(1) = 0), Neumann x2 walls. */
int eddy_plan_set_pressure_bc(eddy_plan_t plan, int implx, int imply, double ru1,
                              const double* rvimplx, const double* rvimply, int itype1,
                              int itype5, int itype6);
/* In place on d_p (NX*NY*NZ doubles) from d_dp (DIRECT's output, ghosts filled). */
int eddy_pressure_update_dev(eddy_plan_t plan, double* d_p, const double* d_dp, double alfxdt);

/* ---------------------------------------------------------------------------------------
 * Resident time loop: the substep of Eddy6.f:2715-2730 repeated on the device with no host
 * round trip. Each step: u* = u + sigma*N(0,1) on the faces CORRECT updates (the stand-in
 * for the momentum predictor), the projection of u* back into u (DIVERGENCE, DIRECT,
 * CORRECT), then SCRINFO's divergence range (info.f:5-165) of the corrected u logged on the
 * device. The step is one captured hipGraph; the step counter lives in device memory.
 * ------------------------------------------------------------------------------------- */
/* Noise alone: us = u + sigma*g (and vs, ws likewise) where g are N(0,1) draws from
 * Philox4x32-10 with counter {cell lo, cell hi, *d_step, 0}, key {seed, 0}, Box-Muller on
 * words (0,1) -> u, v and (2,3) -> w; cell = 0-based offset (I + NX*(J + NY*K)) of the cell,
 * or of its periodic image for the ghost rows J = 0, NY-1. Faces CORRECT does not update are
 * copied. d_step is a device int. */
int eddy_perturb_dev(eddy_plan_t plan, const double* d_u, const double* d_v, const double* d_w,
                     double* d_us, double* d_vs, double* d_ws, double sigma, unsigned seed,
                     const int* d_step);
typedef struct eddy_resident_s* eddy_resident_t;
/* Capture the step for the fields d_u, d_v, d_w (updated in place every step) with d_dp as
 * the pressure correction; the plan must have a non-default stream. Holds the predicted
 * velocities (3 fields), the device step counter (starts at 0) and a log of max_steps
 * (min, max) pairs. */
int eddy_resident_create(eddy_resident_t* res, eddy_plan_t plan, double* d_u, double* d_v,
                         double* d_w, double* d_dp, double alfxdt, double sigma, unsigned seed,
                         int precision, int sweeps, int max_steps);
/* Enqueue nsteps graph replays (asynchronous). Fails with EDDY_ERR_ARG past max_steps. */
int eddy_resident_run(eddy_resident_t res, int nsteps);
/* Synchronise; copy min/max of DIV(u) after each completed step into minmax[2*s], [2*s+1]
 * for s < min(steps, capacity); *steps = steps completed. */
int eddy_resident_log(eddy_resident_t res, double* minmax, int capacity, int* steps);
int eddy_resident_destroy(eddy_resident_t res);

/* ---------------------------------------------------------------------------------------
 * Field files of the reference: IOSCALAR (src/io.f:114-200; the restart files u_, v_, w_,
 * p_<step>.res of src/Eddy6.f:780-796) and IOVECTOR (src/io.f:8-108). Fortran sequential
 * unformatted records (4-byte length markers around each record), INTEGER*4, REAL*8:
 *   DIR=1 IOSCALAR [NX NY NZG 1], NZG plane records, [NSTEP], [TIME], [DTM1 GRAV]
 *   DIR=1 IOVECTOR [NX NY NZG 3], NZG records of the U, V, W planes
 *   DIR=2 IOSCALAR [NX NY NZ 1 TIME DTM1 GRAV] + planes; IOVECTOR [NX NY NZ 3 TIME] + planes
 * NZG = MYSIZE*(NZ-2)+2. Reads take the slab of rank myrank of mysize (NZ planes from
 * myrank*(NZ-2)+1 on, ghost planes included) as the reference's reads do (io.f:26-28,
 * 125-129); mysize = 1 reads the whole file, whose NZ must then equal the caller's.
 * The header and the first and last plane records are checked against the caller's NX, NY,
 * NZ before anything is stored (the reference does not check): a mismatch returns
 * EDDY_ERR_IO and stores nothing; a damaged record met later returns EDDY_ERR_IO after the
 * planes before it were stored. IOVECTOR headers are read
 * with or without TIME (the reference's DIR=1 write omits the TIME its read expects, io.f:24
 * vs :47). Writes produce the DIR=1 layout of a single writer (MYSIZE=1: the GPU holds the
 * whole field) or, with per_rank = 1, the DIR=2 layout. The _dev forms move the planes
 * between the file and device memory through pinned staging buffers on `stream` and
 * synchronise it before returning.
 * ------------------------------------------------------------------------------------- */
typedef struct {
    int nx, ny, nz;           /* header dimensions (NZ is NZG for the DIR=1 layout) */
    int nvar;                 /* 1: IOSCALAR, 3: IOVECTOR */
    int per_rank;             /* 1: IOSCALAR DIR=2 header (TIME, DTM1, GRAV in the header) */
    int has_time;             /* TIME was present (header or IOSCALAR trailer) */
    int nstep;                /* IOSCALAR DIR=1 trailer, else 0 */
    double time, dtm1, grav;  /* as present in the file, else 0 */
} eddy_field_info;

/* Header (and IOSCALAR trailer) of a field file, with its record structure checked. */
int eddy_field_info_read(const char* path, eddy_field_info* info);
/* IOSCALAR(NAME, P, DP, NX, NY, NZ, DIR=-1, TIME, DTM1, NSTEP) into host P(NX,NY,NZ). */
int eddy_ioscalar_read(const char* path, double* p, int nx, int ny, int nz, int myrank, int mysize,
                       eddy_field_info* info);
int eddy_ioscalar_read_dev(const char* path, double* d_p, int nx, int ny, int nz, int myrank,
                           int mysize, eddy_field_info* info, void* stream);
/* IOSCALAR(..., DIR=1) (per_rank = 0) or DIR=2 (per_rank = 1; NSTEP is not stored there). */
int eddy_ioscalar_write(const char* path, const double* p, int nx, int ny, int nz, int per_rank,
                        int nstep, double time, double dtm1, double grav);
int eddy_ioscalar_write_dev(const char* path, const double* d_p, int nx, int ny, int nz,
                            int per_rank, int nstep, double time, double dtm1, double grav,
                            void* stream);
/* IOVECTOR(NAME, UO, VO, WO, UT, VT, WT, NX, NY, NZ, DIR=-1, TIME) into host U, V, W. */
int eddy_iovector_read(const char* path, double* u, double* v, double* w, int nx, int ny, int nz,
                       int myrank, int mysize, eddy_field_info* info);
int eddy_iovector_read_dev(const char* path, double* d_u, double* d_v, double* d_w, int nx, int ny,
                           int nz, int myrank, int mysize, eddy_field_info* info, void* stream);
/* IOVECTOR(..., DIR=1) (per_rank = 0: TIME is not stored, as io.f:47) or DIR=2 (per_rank = 1). */
int eddy_iovector_write(const char* path, const double* u, const double* v, const double* w,
                        int nx, int ny, int nz, int per_rank, double time);
int eddy_iovector_write_dev(const char* path, const double* d_u, const double* d_v,
                            const double* d_w, int nx, int ny, int nz, int per_rank, double time,
                            void* stream);

/* ---------------------------------------------------------------------------------------
 * Kernel timing (hipEvents around each launch on the plan's stream).
 * ------------------------------------------------------------------------------------- */
enum {
    EDDY_K_YFFT_FWD = 0, EDDY_K_ZTRANS_FWD, EDDY_K_XSOLVE, EDDY_K_ZTRANS_INV, EDDY_K_YFFT_INV,
    EDDY_K_DIVERGENCE, EDDY_K_CORRECT, EDDY_K_OTHER, EDDY_K_COUNT
};
int eddy_plan_enable_timing(eddy_plan_t plan, int enable);
/* Synchronises, then returns launches and total milliseconds recorded for kernel id. */
int eddy_plan_kernel_time(eddy_plan_t plan, int kernel_id, long* launches, double* total_ms);
int eddy_plan_reset_timing(eddy_plan_t plan);

#ifdef __cplusplus
}
#endif
#endif
