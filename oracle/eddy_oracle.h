/*
 * eddy_oracle.h -- CPU oracle for the pressure projection (TEST INFRASTRUCTURE ONLY).
 *
 * It is a reference to compare against, never the product: only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it. It restates,
 * in plain C (FP64, OpenMP over pencils and modes), what the reference computes on the
 * projection path:
 *
 *   DIVERGENCE      src/timeindeg.f:908-975   (stencil :944-946)
 *   DIRECT          src/direct.f:2-1118        (IPRES=1 operator, :321-378, :533-545)
 *     RFFTF/RFFTB   src/fftpack.f:2087-2980    (conventions documented at :140-240)
 *     per-mode 2-D  src/direct.f:710-981       (solve + zero-mode mean removal :894-898)
 *     ghost fill    src/direct.f:1035-1106
 *   CORRECT         src/timeindeg.f:981-1030
 *
 * Arrays use the reference's Fortran layout: A(NX,NY,NZ) column-major, element (I,J,K)
 * (1-based) at offset (I-1) + NX*((J-1) + NY*(K-1)). Metric arrays are passed as the
 * reference passes them (AP(1..NX), AU, RP, RU; CP(1..NZG), CW).
 *
 * The per-mode separable solve is done by a generalized eigen-decomposition of the x2 (Z)
 * operator (implicit QL on the symmetrized tridiagonal; cyclic Jacobi when x2 is periodic)
 * followed by Thomas (or, x1 periodic, Sherman-Morrison cyclic) elimination along x1 (X). BLKTRI, which the reference's production path calls, is not in the
 * reference tree; this solves the same discrete system exactly (the solution is unique
 * up to the Neumann constant, which is fixed by the plain-mean removal of direct.f:895-898).
 */
#ifndef EDDY_ORACLE_H
#define EDDY_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

void orc_set_threads(int n);
int orc_get_threads(void);

/* fftpack-convention real transforms of `count` contiguous vectors of length L (in place).
 * Forward: r[0]=sum x; r[2k-1]=sum x cos, r[2k]=-sum x sin; r[L-1]=sum (-1)^j x (L even).
 * Backward: the unnormalized synthesis, so backward(forward(x)) = L*x (fftpack.f:140-240). */
int orc_rfftf_many(int L, long count, double* data);
int orc_rfftb_many(int L, long count, double* data);

/* Eigen-decomposition of a symmetric tridiagonal matrix (diag d[n], off-diag e[n-1]).
 * On return d holds eigenvalues (ascending) and z (n*n, column m = eigenvector m,
 * z[k + n*m]) orthonormal eigenvectors. Returns 0 on success. */
int orc_tridiag_eig(int n, double* d, const double* e, double* z);

/* DIVERGENCE (timeindeg.f:944-946) on interior cells I=2..NX-1, J=2..NY-1, K=2..NZ-1.
 * div is written on the interior only; other entries are left untouched. */
void orc_divergence(int nx, int ny, int nz, const double* us, const double* vs, const double* ws,
                    const double* ap, const double* bp, const double* cp, const double* rp,
                    const double* ru, double* div);

/* CORRECT (timeindeg.f:981-1030). Faces corrected: U on I=2..NX-2 (MP=1 walls) else
 * 2..NX-1; V on J=2..NY-2 (LP=1) else 2..NY-1; W on K=2..NZ-2 (NP=1) else 2..NZ-1
 * (setup.f:17-72). uo/vo/wo are written on corrected faces only. */
void orc_correct(int nx, int ny, int nz, const double* us, const double* vs, const double* ws,
                 double* uo, double* vo, double* wo, const double* dp,
                 const double* au, const double* bv, const double* cw, const double* rp,
                 double alfxdt, int mp, int lp, int np);

/* Apply the operator DIRECT inverts (direct.f:344-352, 370-378, 536) to phi on interior
 * cells; phi's ghosts are ignored (the operator's own BCs are used). out: interior only. */
int orc_laplacian(int nx, int ny, int nz, double delysq, const double* ap, const double* au,
                  const double* cp, const double* cw, const double* rp, const double* ru,
                  int mp, int lp, int np, const double* phi, double* out);

/* DIRECT: dp holds the RHS on the interior on entry and the pressure correction on exit,
 * with ghosts filled as direct.f:1035-1106 does (x: MP=1 copies DP(2), DP(NX-1), MP=0 the
 * periodic images; y: LP=0 periodic images, LP=1 copies; z ghost planes zero). The y
 * direction is always the periodic FFT axis (any L); MP=0 / NP=0 make the x1 / x2 operator
 * cyclic (AM(1), CM(M) / AN(1), CN(N) kept, direct.f:344-352, 370-378). Per y-mode: x2
 * eigenbasis + x1 lines (Thomas, or Sherman-Morrison cyclic), or, when the periodic x2
 * operator has no symmetric form (CW(1) != CW(N+1)), the x1 eigenbasis + x2 lines.
 * Returns 0 on success, <0 on unsupported input. */
int orc_direct(int nx, int ny, int nz, double delysq, const double* ap, const double* au,
               const double* cp, const double* cw, const double* rp, const double* ru,
               int mp, int lp, int np, double* dp);
/* orc_direct with the per-mode method forced: path 1 = x2 eigenbasis, 2 = x1 eigenbasis,
 * 0 = automatic. Returns -2 if the forced axis operator has no symmetric form. */
int orc_direct_path(int nx, int ny, int nz, double delysq, const double* ap, const double* au,
                    const double* cp, const double* cw, const double* rp, const double* ru,
                    int mp, int lp, int np, double* dp, int path);

/* One per-mode 2-D solve (direct.f:753-941 for one J): rhs is (N,M) with z fastest as
 * RHS1D(k + N*(i-1)) (direct.f:755-762), overwritten with the solution. J is the 1-based
 * fftpack mode index; the J==1 mode gets the plain-mean removal of direct.f:895-898. */
int orc_mode_solve(int m, int n, int l, int jmode, double delysq, const double* ap,
                   const double* au, const double* cp, const double* cw, const double* rp,
                   const double* ru, double* rhs);

/* Metrics of a wall-bounded stretched direction from its face coordinates (grid.f:35-66):
 * faces xu[0..nf-1] are xu(1..nf); nf = NX-1. Writes ap[NX], au[NX] (1-based in Fortran). */
void orc_metrics_from_faces(int nx, const double* xu, double* ap, double* au);

/* Pressure update after CORRECT (Eddy6.f:2772-2846): P = P + DP; the implicit-viscous
 * corrections when IMPLY / IMPLX == 1 (RU1, DELY, RVIMPLY/RVIMPLX masks of length NX);
 * PREF = sum(P(2, 2:NY-1, 2))/(NY-2) subtracted everywhere; ghost BCs x (MP; ITYPE(1)==300
 * axis with JSYM), y (LP), z (ITYPE(5), ITYPE(6): 500 periodic, other nonzero Neumann). */
int orc_pressure_update(int nx, int ny, int nz, double* p, const double* dp, double alfxdt, double ru1,
                        double dely, const double* ap, const double* au, const double* rp, const double* ru,
                        const double* rvimplx, const double* rvimply, int implx, int imply, int mp, int lp,
                        int itype1, int itype5, int itype6);

/* Workload noise of the resident time loop (not a reference function):
 * Philox4x32-10 of counter ctr[4] under key[2] (Salmon et al., SC'11) -> out[4]. */
void orc_philox4x32_10(const unsigned* ctr, const unsigned* key, unsigned* out);
/* Three N(0,1) draws for cell `cell` (0-based linear offset in the DP(NX,NY,NZ) array) at
 * `step`: Philox counter {cell lo, cell hi, step, 0}, key {seed, 0}; Box-Muller on words
 * (0,1) -> g[0], g[1] and (2,3) -> g[2]. */
void orc_normals(unsigned long cell, unsigned step, unsigned seed, double* g);
/* us = u + sigma*g0, vs = v + sigma*g1, ws = w + sigma*g2 on the faces CORRECT updates
 * (U: I=2..NX-2, V: J=2..NY-1, W: K=2..NZ-2, over the interior of the other two indices);
 * the periodic ghost rows J=1, J=NY take their images' values; everything else is copied. */
void orc_perturb(int nx, int ny, int nz, const double* u, const double* v, const double* w, double* us,
                 double* vs, double* ws, double sigma, unsigned seed, unsigned step);

#ifdef __cplusplus
}
#endif
#endif
