/*
 * eddy_oracle.c -- CPU oracle for the pressure projection (TEST INFRASTRUCTURE ONLY).
 * See eddy_oracle.h for the contract and the reference lines each function restates.
 * Nothing in the product links this file; it is what the HIP path is compared to.
 */
#include "eddy_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX(I, J, K) ((size_t)((I) - 1) + (size_t)nx * ((size_t)((J) - 1) + (size_t)ny * (size_t)((K) - 1)))

/* Working precision of the transforms and eliminations: FP64 by default; the ORC_LONG build
 * (liboracle_ld.so) carries them in x87 extended precision (64-bit mantissa) so that its
 * answer is exact to well below FP64 roundoff and a comparison against it measures the
 * other side's error alone. Stencils stay FP64 in both builds (the reference's arithmetic). */
#ifdef ORC_LONG
typedef long double real;
#define RCOS cosl
#define RSIN sinl
#define RSQRT sqrtl
#define RFABS fabsl
#define RHYPOT hypotl
#define RATAN atanl
#define REPS 1.0842021724855044e-19L
#else
typedef double real;
#define RCOS cos
#define RSIN sin
#define RSQRT sqrt
#define RFABS fabs
#define RHYPOT hypot
#define RATAN atan
#define REPS 2.220446049250313e-16
#endif

static int g_threads = 1;

void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int orc_get_threads(void) { return g_threads; }

/* ------------------------------------------------------------------------------------ */
/* Complex mixed-radix FFT (recursive decimation in time), used to evaluate the fftpack   */
/* real-transform sums of fftpack.f:140-240. Twiddles come from one table of length n,    */
/* tw[t] = exp(-2*pi*i*t/n), each entry evaluated directly from its own angle.            */
/* ------------------------------------------------------------------------------------ */
typedef struct {
    int n;
    real* twr;
    real* twi;
} cplan;

static int smallest_factor(int n) {
    if (n % 4 == 0 && n != 4) return 4;  /* radix 4 first, as fftpack's factor order (fftpack.f:2096) */
    for (int p = 2; p * p <= n; ++p)
        if (n % p == 0) return p;
    return n;
}

static void cplan_init(cplan* P, int n) {
    P->n = n;
    P->twr = (real*)malloc(sizeof(real) * n);
    P->twi = (real*)malloc(sizeof(real) * n);
    const real tpi = 8 * RATAN((real)1);
    for (int t = 0; t < n; ++t) {
        real a = tpi * (real)t / (real)n;
        P->twr[t] = RCOS(a);
        P->twi[t] = -RSIN(a);
    }
}

static void cplan_free(cplan* P) {
    free(P->twr);
    free(P->twi);
}

/* y[0..n) = sum_j x[j*s] W^{jk}, W = exp(sign*2 pi i / n); sign -1 forward, +1 backward. */
static void cfft_rec(const cplan* P, int n, const real* xr, const real* xi, long s, real* yr,
                     real* yi, int sign, real* tr, real* ti) {
    if (n == 1) {
        yr[0] = xr[0];
        yi[0] = xi[0];
        return;
    }
    int p = smallest_factor(n);
    int m = n / p;
    for (int r = 0; r < p; ++r) cfft_rec(P, m, xr + r * s, xi + r * s, s * p, yr + r * m, yi + r * m, sign, tr, ti);
    int stride = P->n / n; /* table step for roots of unity of order n */
    for (int k = 0; k < m; ++k) {
        for (int q = 0; q < p; ++q) {
            int kk = k + q * m;
            real accr = 0, acci = 0;
            for (int r = 0; r < p; ++r) {
                long t = ((long)r * kk) % n;
                real wr = P->twr[t * stride];
                real wi = P->twi[t * stride] * (sign < 0 ? 1 : -1);
                real ar = yr[r * m + k], ai = yi[r * m + k];
                accr += ar * wr - ai * wi;
                acci += ar * wi + ai * wr;
            }
            tr[kk] = accr;
            ti[kk] = acci;
        }
    }
    memcpy(yr, tr, sizeof(real) * n);
    memcpy(yi, ti, sizeof(real) * n);
}

/* fftpack RFFTF packing of one real vector r[0..L). work: 6*L doubles. */
static void rfftf_one(const cplan* P, int L, real* r, real* w) {
    real *xr = w, *xi = w + L, *yr = w + 2 * L, *yi = w + 3 * L, *tr = w + 4 * L, *ti = w + 5 * L;
    for (int j = 0; j < L; ++j) {
        xr[j] = r[j];
        xi[j] = 0;
    }
    cfft_rec(P, L, xr, xi, 1, yr, yi, -1, tr, ti);
    r[0] = yr[0];
    for (int k = 1; k <= (L - 1) / 2; ++k) {
        r[2 * k - 1] = yr[k];
        r[2 * k] = yi[k];
    }
    if (L % 2 == 0) r[L - 1] = yr[L / 2];
}

/* fftpack RFFTB synthesis of one packed vector. */
static void rfftb_one(const cplan* P, int L, real* r, real* w) {
    real *xr = w, *xi = w + L, *yr = w + 2 * L, *yi = w + 3 * L, *tr = w + 4 * L, *ti = w + 5 * L;
    xr[0] = r[0];
    xi[0] = 0;
    for (int k = 1; k <= (L - 1) / 2; ++k) {
        xr[k] = r[2 * k - 1];
        xi[k] = r[2 * k];
        xr[L - k] = r[2 * k - 1];
        xi[L - k] = -r[2 * k];
    }
    if (L % 2 == 0) {
        xr[L / 2] = r[L - 1];
        xi[L / 2] = 0;
    }
    cfft_rec(P, L, xr, xi, 1, yr, yi, +1, tr, ti);
    for (int j = 0; j < L; ++j) r[j] = yr[j];
}

int orc_rfftf_many(int L, long count, double* data) {
    if (L < 1 || count < 0) return -1;
    if (L == 1) return 0;
    cplan P;
    cplan_init(&P, L);
#pragma omp parallel num_threads(g_threads)
    {
        real* w = (real*)malloc(sizeof(real) * 7 * L);
        real* v = w + 6 * L;
#pragma omp for schedule(static)
        for (long c = 0; c < count; ++c) {
            for (int j = 0; j < L; ++j) v[j] = data[c * L + j];
            rfftf_one(&P, L, v, w);
            for (int j = 0; j < L; ++j) data[c * L + j] = (double)v[j];
        }
        free(w);
    }
    cplan_free(&P);
    return 0;
}

int orc_rfftb_many(int L, long count, double* data) {
    if (L < 1 || count < 0) return -1;
    if (L == 1) return 0;
    cplan P;
    cplan_init(&P, L);
#pragma omp parallel num_threads(g_threads)
    {
        real* w = (real*)malloc(sizeof(real) * 7 * L);
        real* v = w + 6 * L;
#pragma omp for schedule(static)
        for (long c = 0; c < count; ++c) {
            for (int j = 0; j < L; ++j) v[j] = data[c * L + j];
            rfftb_one(&P, L, v, w);
            for (int j = 0; j < L; ++j) data[c * L + j] = (double)v[j];
        }
        free(w);
    }
    cplan_free(&P);
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* Symmetric tridiagonal eigenproblem: implicit QL with Wilkinson shifts (Bowdler, Martin, */
/* Reinsch and Wilkinson, Numer. Math. 11 (1968) 293-306), accumulating Givens rotations.  */
/* ------------------------------------------------------------------------------------ */
static int tql(int n, real* d, const real* e_in, real* z) {
    real* e = (real*)calloc((size_t)n + 1, sizeof(real));
    for (int i = 0; i + 1 < n; ++i) e[i] = e_in[i];
    for (int i = 0; i < n * n; ++i) z[i] = 0;
    for (int i = 0; i < n; ++i) z[i + (size_t)n * i] = 1;
    const real eps = REPS;
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        for (;;) {
            int m;
            for (m = l; m < n - 1; ++m) {
                real dd = RFABS(d[m]) + RFABS(d[m + 1]);
                if (RFABS(e[m]) <= eps * dd) break;
            }
            if (m == l) break;
            if (++iter > 60) {
                free(e);
                return -1;
            }
            real g = (d[l + 1] - d[l]) / (2 * e[l]);
            real r = RHYPOT(g, (real)1);
            g = d[m] - d[l] + e[l] / (g + (g >= 0 ? RFABS(r) : -RFABS(r)));
            real s = 1, c = 1, p = 0;
            int i;
            int early = 0;
            for (i = m - 1; i >= l; --i) {
                real f = s * e[i];
                real b = c * e[i];
                r = RHYPOT(f, g);
                e[i + 1] = r;
                if (r == 0) {
                    d[i + 1] -= p;
                    e[m] = 0;
                    early = 1;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                for (int k = 0; k < n; ++k) {
                    real zf = z[k + (size_t)n * (i + 1)];
                    real zi = z[k + (size_t)n * i];
                    z[k + (size_t)n * (i + 1)] = s * zi + c * zf;
                    z[k + (size_t)n * i] = c * zi - s * zf;
                }
            }
            if (early) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0;
        }
    }
    free(e);
    /* sort ascending, carrying eigenvectors */
    for (int i = 0; i < n - 1; ++i) {
        int k = i;
        for (int j = i + 1; j < n; ++j)
            if (d[j] < d[k]) k = j;
        if (k != i) {
            real t = d[i];
            d[i] = d[k];
            d[k] = t;
            for (int r = 0; r < n; ++r) {
                real tz = z[r + (size_t)n * i];
                z[r + (size_t)n * i] = z[r + (size_t)n * k];
                z[r + (size_t)n * k] = tz;
            }
        }
    }
    return 0;
}

int orc_tridiag_eig(int n, double* d, const double* e, double* z) {
    real* dr = (real*)malloc(sizeof(real) * n);
    real* er = (real*)calloc((size_t)(n > 1 ? n : 1), sizeof(real));
    real* zr = (real*)malloc(sizeof(real) * (size_t)n * n);
    for (int i = 0; i < n; ++i) dr[i] = d[i];
    for (int i = 0; i + 1 < n; ++i) er[i] = e[i];
    int rc = tql(n, dr, er, zr);
    for (int i = 0; i < n; ++i) d[i] = (double)dr[i];
    for (size_t i = 0; i < (size_t)n * n; ++i) z[i] = (double)zr[i];
    free(dr);
    free(er);
    free(zr);
    return rc;
}

/* ------------------------------------------------------------------------------------ */
/* Grid metrics of a wall-bounded stretched axis (grid.f:45-58)                           */
/* ------------------------------------------------------------------------------------ */
void orc_metrics_from_faces(int nx, const double* xu_in, double* ap, double* au) {
    /* Fortran xu(1..nx): xu(1..nx-1) given, xu(nx) extrapolated (grid.f:49). */
    double* xu = (double*)malloc(sizeof(double) * (nx + 1));
    for (int i = 1; i <= nx - 1; ++i) xu[i] = xu_in[i - 1];
    int ix1 = 2, ix2 = nx - 1;
    xu[nx] = 2.0 * xu[ix2] - xu[ix2 - 1];
    for (int i = ix1; i <= ix2; ++i) au[i - 1] = 2.0 / (xu[i + 1] - xu[i - 1]);
    au[0] = 2.0 / (-3.0 * xu[1] + 4.0 * xu[2] - xu[3]);
    au[ix2 - 1] = 2.0 / (3.0 * xu[ix2] - 4.0 * xu[ix2 - 1] + xu[ix2 - 2]);
    au[nx - 1] = au[ix2 - 1];
    for (int i = ix1; i <= ix2; ++i) ap[i - 1] = 1.0 / (xu[i] - xu[i - 1]);
    ap[0] = ap[ix1 - 1];
    ap[nx - 1] = ap[ix2 - 1];
    free(xu);
}

/* ------------------------------------------------------------------------------------ */
/* Stencils (timeindeg.f)                                                                 */
/* ------------------------------------------------------------------------------------ */
void orc_divergence(int nx, int ny, int nz, const double* us, const double* vs, const double* ws,
                    const double* ap, const double* bp, const double* cp, const double* rp,
                    const double* ru, double* div) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int K = 2; K <= nz - 1; ++K)
        for (int J = 2; J <= ny - 1; ++J)
            for (int I = 2; I <= nx - 1; ++I) {
                double a = ap[I - 1] * (ru[I - 1] * us[IDX(I, J, K)] - ru[I - 2] * us[IDX(I - 1, J, K)]) / rp[I - 1];
                double b = bp[J - 1] * (vs[IDX(I, J, K)] - vs[IDX(I, J - 1, K)]) / rp[I - 1];
                double c = cp[K - 1] * (ws[IDX(I, J, K)] - ws[IDX(I, J, K - 1)]);
                div[IDX(I, J, K)] = (a + b + c);
            }
}

void orc_correct(int nx, int ny, int nz, const double* us, const double* vs, const double* ws,
                 double* uo, double* vo, double* wo, const double* dp,
                 const double* au, const double* bv, const double* cw, const double* rp,
                 double alfxdt, int mp, int lp, int np) {
    int ieu = (mp == 1) ? nx - 2 : nx - 1;
    int jev = (lp == 1) ? ny - 2 : ny - 1;
    int kew = (np == 1) ? nz - 2 : nz - 1;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int K = 2; K <= nz - 1; ++K) {
        for (int J = 2; J <= ny - 1; ++J)
            for (int I = 2; I <= ieu; ++I)
                uo[IDX(I, J, K)] = us[IDX(I, J, K)] - au[I - 1] * (dp[IDX(I + 1, J, K)] - dp[IDX(I, J, K)]) * alfxdt;
        for (int J = 2; J <= jev; ++J)
            for (int I = 2; I <= nx - 1; ++I)
                vo[IDX(I, J, K)] = vs[IDX(I, J, K)] - bv[J - 1] * (dp[IDX(I, J + 1, K)] - dp[IDX(I, J, K)]) * alfxdt / rp[I - 1];
        if (K <= kew)
            for (int J = 2; J <= ny - 1; ++J)
                for (int I = 2; I <= nx - 1; ++I)
                    wo[IDX(I, J, K)] = ws[IDX(I, J, K)] - cw[K - 1] * (dp[IDX(I, J, K + 1)] - dp[IDX(I, J, K)]) * alfxdt;
    }
}

/* ------------------------------------------------------------------------------------ */
/* Operator coefficients of DIRECT, IPRES=1 form (direct.f:321-352, 370-378)              */
/* ------------------------------------------------------------------------------------ */
typedef struct {
    int M, N, L;
    real *am, *bm, *cm, *r2inv; /* x1 (X): length M */
    real *an, *bn, *cn, *cpz;   /* x2 (Z): length N */
    real* ak;                   /* modified wavenumbers, length L (AK(1..L)) */
} coefs;

static int coefs_build(coefs* C, int nx, int ny, int nz, double delysq, const double* ap, const double* au,
                       const double* cp, const double* cw, const double* rp, const double* ru, int mp, int np) {
    int M = nx - 2, L = ny - 2, N = nz - 2;
    C->M = M;
    C->N = N;
    C->L = L;
    C->am = (real*)malloc(sizeof(real) * M);
    C->bm = (real*)malloc(sizeof(real) * M);
    C->cm = (real*)malloc(sizeof(real) * M);
    C->r2inv = (real*)malloc(sizeof(real) * M);
    C->an = (real*)malloc(sizeof(real) * N);
    C->bn = (real*)malloc(sizeof(real) * N);
    C->cn = (real*)malloc(sizeof(real) * N);
    C->cpz = (real*)malloc(sizeof(real) * N);
    C->ak = (real*)malloc(sizeof(real) * L);
    /* AK(K) = 2(1 - cos(2 pi k'/L))/DELYSQ (direct.f:321-331), evaluated as 4 sin^2(pi k'/L)/DELYSQ */
    const real pi = 4 * RATAN((real)1);
    for (int k = 1; k <= L; ++k) {
        int kw = (k <= L / 2) ? (k - 1) : (L - k + 1);
        real sn = RSIN(pi * (real)kw / (real)L);
        C->ak[k - 1] = 4 * sn * sn / (real)delysq;
    }
    for (int i = 1; i <= M; ++i) {
        C->am[i - 1] = (real)ap[i] / (real)rp[i] * (real)au[i - 1] * (real)ru[i - 1];
        C->cm[i - 1] = (real)ap[i] / (real)rp[i] * (real)au[i] * (real)ru[i];
        C->r2inv[i - 1] = 1 / ((real)rp[i] * (real)rp[i]);
    }
    if (mp == 1) {
        C->am[0] = 0;
        C->cm[M - 1] = 0;
    }
    for (int i = 0; i < M; ++i) C->bm[i] = -C->am[i] - C->cm[i];
    for (int k = 1; k <= N; ++k) {
        C->an[k - 1] = (real)cp[k] * (real)cw[k - 1];
        C->cn[k - 1] = (real)cp[k] * (real)cw[k];
        C->cpz[k - 1] = cp[k];
    }
    if (np == 1) {
        C->an[0] = 0;
        C->cn[N - 1] = 0;
    }
    for (int k = 0; k < N; ++k) C->bn[k] = -C->an[k] - C->cn[k];
    return 0;
}

static void coefs_free(coefs* C) {
    free(C->am);
    free(C->bm);
    free(C->cm);
    free(C->r2inv);
    free(C->an);
    free(C->bn);
    free(C->cn);
    free(C->cpz);
    free(C->ak);
}

int orc_laplacian(int nx, int ny, int nz, double delysq, const double* ap, const double* au,
                  const double* cp, const double* cw, const double* rp, const double* ru,
                  int mp, int lp, int np, const double* phi, double* out) {
    (void)lp; /* DIRECT's y direction is periodic whatever LP says (LP only picks its ghost rule) */
    if ((mp != 0 && mp != 1) || (np != 0 && np != 1)) return -2;
    coefs C;
    coefs_build(&C, nx, ny, nz, delysq, ap, au, cp, cw, rp, ru, mp, np);
    int M = C.M, L = C.L, N = C.N;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int k = 1; k <= N; ++k)
        for (int j = 1; j <= L; ++j) {
            int jm = (j == 1) ? L : j - 1, jp = (j == L) ? 1 : j + 1;
            for (int i = 1; i <= M; ++i) {
                int I = i + 1, J = j + 1, K = k + 1;
                double p0 = phi[IDX(I, J, K)];
                double sx = (double)C.bm[i - 1] * p0;
                /* MP=0: AM(1) couples x1 cell 1 to cell M and CM(M) cell M to cell 1 */
                if (i > 1) sx += (double)C.am[i - 1] * phi[IDX(I - 1, J, K)];
                else if (mp == 0) sx += (double)C.am[0] * phi[IDX(M + 1, J, K)];
                if (i < M) sx += (double)C.cm[i - 1] * phi[IDX(I + 1, J, K)];
                else if (mp == 0) sx += (double)C.cm[M - 1] * phi[IDX(2, J, K)];
                double sy = (phi[IDX(I, jp + 1, K)] - 2.0 * p0 + phi[IDX(I, jm + 1, K)]) / delysq * (double)C.r2inv[i - 1];
                double sz = (double)C.bn[k - 1] * p0;
                if (k > 1) sz += (double)C.an[k - 1] * phi[IDX(I, J, K - 1)];
                else if (np == 0) sz += (double)C.an[0] * phi[IDX(I, J, N + 1)];
                if (k < N) sz += (double)C.cn[k - 1] * phi[IDX(I, J, K + 1)];
                else if (np == 0) sz += (double)C.cn[N - 1] * phi[IDX(I, J, 2)];
                out[IDX(I, J, K)] = sx + sy + sz;
            }
        }
    coefs_free(&C);
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* Dense symmetric eigenproblem by cyclic Jacobi rotations (Golub & Van Loan, Matrix      */
/* Computations, 8.5.2-8.5.3), for the cyclic (periodic) operators whose symmetric form   */
/* is not tridiagonal. A (n*n, A[r + n*c]) is destroyed; lam ascending, V column m =      */
/* eigenvector m.                                                                          */
/* ------------------------------------------------------------------------------------ */
static int jacobi_eig(int n, real* A, real* lam, real* V) {
    for (int i = 0; i < n * n; ++i) V[i] = 0;
    for (int i = 0; i < n; ++i) V[i + (size_t)n * i] = 1;
    real fro = 0;
    for (int i = 0; i < n * n; ++i) fro += A[i] * A[i];
    fro = RSQRT(fro);
    for (int sweep = 0; sweep < 100; ++sweep) {
        real off = 0;
        for (int c = 0; c < n; ++c)
            for (int r = 0; r < n; ++r)
                if (r != c) off += A[r + (size_t)n * c] * A[r + (size_t)n * c];
        if (RSQRT(off) <= REPS * fro * 1e-2 || off == 0) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                real apq = A[p + (size_t)n * q];
                if (apq == 0) continue;
                real app = A[p + (size_t)n * p], aqq = A[q + (size_t)n * q];
                real theta = (aqq - app) / (2 * apq);
                real t = (theta >= 0 ? 1 : -1) / (RFABS(theta) + RSQRT(theta * theta + 1));
                real c = 1 / RSQRT(t * t + 1), s = t * c;
                for (int k = 0; k < n; ++k) { /* A <- A J (columns p, q) */
                    real akp = A[k + (size_t)n * p], akq = A[k + (size_t)n * q];
                    A[k + (size_t)n * p] = c * akp - s * akq;
                    A[k + (size_t)n * q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) { /* A <- J^T A (rows p, q) */
                    real apk = A[p + (size_t)n * k], aqk = A[q + (size_t)n * k];
                    A[p + (size_t)n * k] = c * apk - s * aqk;
                    A[q + (size_t)n * k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    real vkp = V[k + (size_t)n * p], vkq = V[k + (size_t)n * q];
                    V[k + (size_t)n * p] = c * vkp - s * vkq;
                    V[k + (size_t)n * q] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < n; ++i) lam[i] = A[i + (size_t)n * i];
    for (int i = 0; i < n - 1; ++i) {
        int k = i;
        for (int j = i + 1; j < n; ++j)
            if (lam[j] < lam[k]) k = j;
        if (k != i) {
            real t = lam[i];
            lam[i] = lam[k];
            lam[k] = t;
            for (int r = 0; r < n; ++r) {
                real tz = V[r + (size_t)n * i];
                V[r + (size_t)n * i] = V[r + (size_t)n * k];
                V[r + (size_t)n * k] = tz;
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* Separable transform of one axis. The axis operator T (tridiagonal; with NP/MP=0 also   */
/* the corner couplings lo[0] -> x(n-1) and up[n-1] -> x(0)) is symmetric under weights   */
/* wt: wt(k) T(k,l) = wt(l) T(l,k). H = W^{1/2} T W^{-1/2} = U diag(lam) U^T; forward     */
/* xh = U^T W^{1/2} x, inverse x = W^{-1/2} U xh. The null pair (lam = 0, U ~ W^{1/2}) of */
/* an operator with zero row sums is set exactly. x2 axis: wt = 1/CP (T = diag(CP) S);    */
/* x1 axis: wt = RP/AP.                                                                    */
/* ------------------------------------------------------------------------------------ */
typedef struct {
    int N;
    real* lam;  /* N */
    real* U;    /* N*N, column m = eigenvector m */
    real* sq;   /* sqrt(wt) */
    int null_m;   /* index of the zero eigenvalue (-1: none) */
} zeig;

static int axis_eig(zeig* Z, int n, const real* lo, const real* dg, const real* up, const real* wt, int cyc,
                    int has_null) {
    Z->N = n;
    Z->lam = (real*)malloc(sizeof(real) * n);
    Z->U = (real*)malloc(sizeof(real) * (size_t)n * n);
    Z->sq = (real*)malloc(sizeof(real) * n);
    for (int k = 0; k < n; ++k) Z->sq[k] = RSQRT(wt[k]);
    int rc = 0;
    if (!cyc || n < 3) {
        real* e = (real*)malloc(sizeof(real) * (n > 1 ? n - 1 : 1));
        for (int k = 0; k < n; ++k) Z->lam[k] = dg[k];
        /* H(k,k+1) = T(k,k+1) sqrt(wt(k)/wt(k+1)) */
        for (int k = 0; k + 1 < n; ++k) e[k] = up[k] * Z->sq[k] / Z->sq[k + 1];
        if (cyc && n == 2) e[0] += lo[0] * Z->sq[0] / Z->sq[1]; /* both couplings join one entry */
        rc = tql(n, Z->lam, e, Z->U);
        free(e);
    } else {
        real* A = (real*)calloc((size_t)n * n, sizeof(real));
        for (int k = 0; k < n; ++k) {
            A[k + (size_t)n * k] = dg[k];
            if (k + 1 < n) {
                real h = up[k] * Z->sq[k] / Z->sq[k + 1];
                A[k + (size_t)n * (k + 1)] = h;
                A[(k + 1) + (size_t)n * k] = h;
            }
        }
        /* corners: H(0,n-1) = T(0,n-1) sqrt(wt0/wt_{n-1}), H(n-1,0) = T(n-1,0) sqrt(wt_{n-1}/wt0):
         * equal when the operator is symmetric under wt (the caller checked); their mean */
        real h = (real)0.5 * (lo[0] * Z->sq[0] / Z->sq[n - 1] + up[n - 1] * Z->sq[n - 1] / Z->sq[0]);
        A[(size_t)n * (n - 1)] += h;
        A[(n - 1)] += h;
        rc = jacobi_eig(n, A, Z->lam, Z->U);
        free(A);
    }
    if (rc) return rc;
    Z->null_m = -1;
    if (has_null) {
        int m0 = 0;
        for (int m = 1; m < n; ++m)
            if (RFABS(Z->lam[m]) < RFABS(Z->lam[m0])) m0 = m;
        real nrm = 0;
        for (int k = 0; k < n; ++k) nrm += wt[k];
        nrm = RSQRT(nrm);
        real sgn = 0;
        for (int k = 0; k < n; ++k) sgn += Z->U[k + (size_t)n * m0];
        sgn = sgn >= 0 ? 1 : -1;
        for (int k = 0; k < n; ++k) Z->U[k + (size_t)n * m0] = sgn * Z->sq[k] / nrm;
        Z->lam[m0] = 0;
        Z->null_m = m0;
    }
    return 0;
}

static void zeig_free(zeig* Z) {
    free(Z->lam);
    free(Z->U);
    free(Z->sq);
}

/* x2 axis of DIRECT's operator: T = diag(CP) S, weights 1/CP (direct.f:370-378). */
static int zeig_build(zeig* Z, const coefs* C, int np) {
    real* wt = (real*)malloc(sizeof(real) * C->N);
    for (int k = 0; k < C->N; ++k) wt[k] = 1 / C->cpz[k];
    int rc = axis_eig(Z, C->N, C->an, C->bn, C->cn, wt, np == 0, 1);
    free(wt);
    /* the forward transform divides by sq = sqrt(CP) (T = diag(CP) S): store sqrt(CP) */
    for (int k = 0; k < C->N; ++k) Z->sq[k] = RSQRT(C->cpz[k]);
    if (Z->null_m >= 0) {
        real nrm = 0;
        for (int k = 0; k < C->N; ++k) nrm += 1 / C->cpz[k];
        nrm = RSQRT(nrm);
        real sgn = 0;
        for (int k = 0; k < C->N; ++k) sgn += Z->U[k + (size_t)C->N * Z->null_m];
        sgn = sgn >= 0 ? 1 : -1;
        for (int k = 0; k < C->N; ++k) Z->U[k + (size_t)C->N * Z->null_m] = sgn / (Z->sq[k] * nrm);
    }
    return rc;
}

/* Thomas elimination for a(i)x(i-1)+b(i)x(i)+c(i)x(i+1)=d(i); pin==1 replaces row 1 by
 * x(1)=0 (the singular mode; cf. direct.f:849-852). a[0] and c[M-1] are not read. */
static void thomas(int M, const real* a, const real* b, const real* c, real* d, int pin, real* cp) {
    real bb = b[0], cc = c[0], dd = d[0];
    if (pin) {
        bb = 1;
        cc = 0;
        dd = 0;
    }
    cp[0] = cc / bb;
    d[0] = dd / bb;
    for (int i = 1; i < M; ++i) {
        real den = b[i] - a[i] * cp[i - 1];
        cp[i] = c[i] / den;
        d[i] = (d[i] - a[i] * d[i - 1]) / den;
    }
    for (int i = M - 2; i >= 0; --i) d[i] -= cp[i] * d[i + 1];
}

/* Cyclic tridiagonal system (periodic line: row 0 couples to x(M-1) through a[0], row M-1
 * to x(0) through c[M-1]) by the Sherman-Morrison correction of two tridiagonal solves
 * (the method of numrec.f:754-781, CYCLIC): T = T' + u v^T with u = (g, 0, ..., 0, c[M-1]),
 * v = (1, 0, ..., 0, a[0]/g), g = -b(0). work: 3*M. */
static void cyclic_solve(int M, const real* a, const real* b, const real* c, real* d, real* work) {
    real* bb = work;
    real* z = work + M;
    real* cp = work + 2 * M;
    const real alpha = c[M - 1], beta = a[0], g = -b[0];
    for (int i = 0; i < M; ++i) bb[i] = b[i];
    bb[0] = b[0] - g;
    bb[M - 1] = b[M - 1] - alpha * beta / g;
    thomas(M, a, bb, c, d, 0, cp);
    for (int i = 0; i < M; ++i) z[i] = 0;
    z[0] = g;
    z[M - 1] = alpha;
    thomas(M, a, bb, c, z, 0, cp);
    const real fact = (d[0] + beta * d[M - 1] / g) / (1 + z[0] + beta * z[M - 1] / g);
    for (int i = 0; i < M; ++i) d[i] -= fact * z[i];
}

/* One line of the separable solve: tridiagonal (walls) or cyclic (periodic) with diagonal
 * b, the singular line pinned at x(0) = 0 (its corner couplings then act on x(0) = 0). */
static void line_solve(int M, const real* a, const real* b, const real* c, real* d, int cyc, int pin, real* work) {
    if (cyc && !pin && M >= 3) cyclic_solve(M, a, b, c, d, work);
    else thomas(M, a, b, c, d, pin, work);
}

/* Solve one y-mode by the x2 eigenbasis. rhs layout: work[i + M*k] (x fastest). */
static void mode_solve_xfast(const coefs* C, const zeig* Z, int jmode, int mp, real* work, real* tmp) {
    int M = C->M, N = C->N;
    real akj = C->ak[jmode / 2]; /* AK(J/2+1), J 1-based (direct.f:536) */
    /* forward z transform: tmp[i + M*m] = sum_k U(k,m) work[i+M*k] / sq(k) */
    for (int m = 0; m < N; ++m)
        for (int i = 0; i < M; ++i) tmp[i + (size_t)M * m] = 0.0;
    for (int k = 0; k < N; ++k) {
        real s = 1 / Z->sq[k];
        for (int m = 0; m < N; ++m) {
            real u = Z->U[k + (size_t)N * m] * s;
            const real* src = work + (size_t)M * k;
            real* dst = tmp + (size_t)M * m;
            for (int i = 0; i < M; ++i) dst[i] += u * src[i];
        }
    }
    /* x solves */
    real* diag = (real*)malloc(sizeof(real) * M * 4);
    real* lw = diag + M;
    for (int m = 0; m < N; ++m) {
        for (int i = 0; i < M; ++i) diag[i] = C->bm[i] - akj * C->r2inv[i] + Z->lam[m];
        int pin = (jmode == 1 && m == Z->null_m);
        line_solve(M, C->am, diag, C->cm, tmp + (size_t)M * m, mp == 0, pin, lw);
    }
    free(diag);
    /* inverse z transform */
    for (int k = 0; k < N; ++k) {
        real* dst = work + (size_t)M * k;
        for (int i = 0; i < M; ++i) dst[i] = 0;
        for (int m = 0; m < N; ++m) {
            real u = Z->U[k + (size_t)N * m] * Z->sq[k];
            const real* src = tmp + (size_t)M * m;
            for (int i = 0; i < M; ++i) dst[i] += u * src[i];
        }
    }
    /* zero-mode plain-mean removal (direct.f:894-898) */
    if (jmode == 1) {
        real s = 0;
        for (size_t t = 0; t < (size_t)M * N; ++t) s += work[t];
        s /= (real)M * (real)N;
        for (size_t t = 0; t < (size_t)M * N; ++t) work[t] -= s;
    }
}

/* Solve one y-mode by the x1 eigenbasis of that mode's x1 operator (AM, BM - AK/RP^2, CM;
 * symmetric under the weights RP/AP), then x2 lines (tridiagonal, or cyclic with NP=0):
 * for x2 operators with no symmetric form (NP=0 with CW(1) != CW(N+1)). */
static int mode_solve_x1basis(const coefs* C, const real* wx, int jmode, int mp, int np, real* work, real* tmp) {
    int M = C->M, N = C->N;
    real akj = C->ak[jmode / 2];
    real* dg = (real*)malloc(sizeof(real) * M);
    for (int i = 0; i < M; ++i) dg[i] = C->bm[i] - akj * C->r2inv[i];
    zeig X;
    int rc = axis_eig(&X, M, C->am, dg, C->cm, wx, mp == 0, jmode == 1);
    free(dg);
    if (rc) {
        zeig_free(&X);
        return rc;
    }
    /* forward: tmp[k + N*m] = sum_i U(i,m) sqrt(wx(i)) work[i + M*k] */
    for (size_t t = 0; t < (size_t)M * N; ++t) tmp[t] = 0;
    for (int k = 0; k < N; ++k)
        for (int m = 0; m < M; ++m) {
            real s = 0;
            for (int i = 0; i < M; ++i) s += X.U[i + (size_t)M * m] * X.sq[i] * work[i + (size_t)M * k];
            tmp[k + (size_t)N * m] = s;
        }
    real* diag = (real*)malloc(sizeof(real) * N * 4);
    real* lw = diag + N;
    for (int m = 0; m < M; ++m) {
        for (int k = 0; k < N; ++k) diag[k] = C->bn[k] + X.lam[m];
        int pin = (jmode == 1 && m == X.null_m);
        line_solve(N, C->an, diag, C->cn, tmp + (size_t)N * m, np == 0, pin, lw);
    }
    free(diag);
    for (int k = 0; k < N; ++k)
        for (int i = 0; i < M; ++i) {
            real s = 0;
            for (int m = 0; m < M; ++m) s += X.U[i + (size_t)M * m] * tmp[k + (size_t)N * m];
            work[i + (size_t)M * k] = s / X.sq[i];
        }
    zeig_free(&X);
    if (jmode == 1) {
        real s = 0;
        for (size_t t = 0; t < (size_t)M * N; ++t) s += work[t];
        s /= (real)M * (real)N;
        for (size_t t = 0; t < (size_t)M * N; ++t) work[t] -= s;
    }
    return 0;
}

int orc_mode_solve(int m, int n, int l, int jmode, double delysq, const double* ap,
                   const double* au, const double* cp, const double* cw, const double* rp,
                   const double* ru, double* rhs) {
    int nx = m + 2, ny = l + 2, nz = n + 2;
    coefs C;
    coefs_build(&C, nx, ny, nz, delysq, ap, au, cp, cw, rp, ru, 1, 1);
    zeig Z;
    if (zeig_build(&Z, &C, 1)) {
        coefs_free(&C);
        return -3;
    }
    real* work = (real*)malloc(sizeof(real) * m * n);
    real* tmp = (real*)malloc(sizeof(real) * m * n);
    for (int i = 0; i < m; ++i)
        for (int k = 0; k < n; ++k) work[i + (size_t)m * k] = rhs[k + (size_t)n * i];
    mode_solve_xfast(&C, &Z, jmode, 1, work, tmp);
    for (int i = 0; i < m; ++i)
        for (int k = 0; k < n; ++k) rhs[k + (size_t)n * i] = work[i + (size_t)m * k];
    free(work);
    free(tmp);
    zeig_free(&Z);
    coefs_free(&C);
    return 0;
}

/* Is a periodic axis operator symmetric under its weights? (wt(0) T(0,n-1) = wt(n-1) T(n-1,0)) */
static int corners_symmetric(int n, const real* lo, const real* up, const real* wt) {
    real x = wt[0] * lo[0], y = wt[n - 1] * up[n - 1];
    real s = RFABS(x) > RFABS(y) ? RFABS(x) : RFABS(y);
    return RFABS(x - y) <= (real)1e-12 * s;
}

/* ------------------------------------------------------------------------------------ */
/* DIRECT (direct.f:2-1118 for MYSIZE=1, IPRES=1 operator, BLKTRI's BC flags)             */
/* ------------------------------------------------------------------------------------ */
int orc_direct_path(int nx, int ny, int nz, double delysq, const double* ap, const double* au,
                    const double* cp, const double* cw, const double* rp, const double* ru,
                    int mp, int lp, int np, double* dp, int path) {
    if ((mp != 0 && mp != 1) || (lp != 0 && lp != 1) || (np != 0 && np != 1)) return -2;
    if (nx < 4 || ny < 3 || nz < 4) return -1;
    coefs C;
    coefs_build(&C, nx, ny, nz, delysq, ap, au, cp, cw, rp, ru, mp, np);
    int M = C.M, L = C.L, N = C.N;
    real* wz = (real*)malloc(sizeof(real) * N);
    real* wx = (real*)malloc(sizeof(real) * M);
    for (int k = 0; k < N; ++k) wz[k] = 1 / C.cpz[k];
    for (int i = 0; i < M; ++i) wx[i] = (real)rp[i + 1] / (real)ap[i + 1];
    const int z_ok = np == 1 || corners_symmetric(N, C.an, C.cn, wz);
    int x_ok = mp == 1 || corners_symmetric(M, C.am, C.cm, wx);
    for (int i = 0; i < M; ++i) x_ok = x_ok && wx[i] > 0;
    if (path == 0) path = z_ok ? 1 : 2;
    if ((path == 1 && !z_ok) || (path == 2 && !x_ok)) {
        free(wz);
        free(wx);
        coefs_free(&C);
        return -2;
    }
    zeig Z;
    if (path == 1 && zeig_build(&Z, &C, np)) {
        free(wz);
        free(wx);
        coefs_free(&C);
        return -3;
    }
    /* spectral buffer: spec[i + M*(k + N*j)] (mode-major, x fastest) */
    real* spec = (real*)malloc(sizeof(real) * (size_t)M * N * L);
    cplan P;
    cplan_init(&P, L);
    /* forward y-FFT of every interior pencil (direct.f:620-629) */
#pragma omp parallel num_threads(g_threads)
    {
        real* pen = (real*)malloc(sizeof(real) * 7 * L);
        real* w = pen + L;
#pragma omp for schedule(static) collapse(2)
        for (int k = 0; k < N; ++k)
            for (int i = 0; i < M; ++i) {
                for (int j = 0; j < L; ++j) pen[j] = dp[IDX(i + 2, j + 2, k + 2)];
                if (L > 1) rfftf_one(&P, L, pen, w);
                for (int j = 0; j < L; ++j) spec[i + (size_t)M * (k + (size_t)N * j)] = pen[j];
            }
        free(pen);
    }
    /* per-mode solves (direct.f:710-981) */
    int bad = 0;
#pragma omp parallel num_threads(g_threads)
    {
        real* tmp = (real*)malloc(sizeof(real) * (size_t)M * N);
#pragma omp for schedule(dynamic, 1)
        for (int j = 0; j < L; ++j) {
            if (path == 1) mode_solve_xfast(&C, &Z, j + 1, mp, spec + (size_t)M * N * j, tmp);
            else if (mode_solve_x1basis(&C, wx, j + 1, mp, np, spec + (size_t)M * N * j, tmp)) {
#pragma omp atomic write
                bad = 1;
            }
        }
        free(tmp);
    }
    /* DP=0, /L, inverse FFT (direct.f:1035-1085) */
    size_t tot = (size_t)nx * ny * nz;
    for (size_t t = 0; t < tot; ++t) dp[t] = 0.0;
    real invL = 1 / (real)L;
#pragma omp parallel num_threads(g_threads)
    {
        real* pen = (real*)malloc(sizeof(real) * 7 * L);
        real* w = pen + L;
#pragma omp for schedule(static) collapse(2)
        for (int k = 0; k < N; ++k)
            for (int i = 0; i < M; ++i) {
                for (int j = 0; j < L; ++j) pen[j] = spec[i + (size_t)M * (k + (size_t)N * j)] * invL;
                if (L > 1) rfftb_one(&P, L, pen, w);
                for (int j = 0; j < L; ++j) dp[IDX(i + 2, j + 2, k + 2)] = (double)pen[j];
            }
        free(pen);
    }
    /* ghosts (direct.f:1090-1106): x cyclic (MP=0) or copied (MP=1), then y periodic (LP=0)
     * or copied (LP=1); z ghost planes stay zero */
    for (int K = 1; K <= nz; ++K)
        for (int J = 1; J <= ny; ++J) {
            dp[IDX(1, J, K)] = mp == 0 ? dp[IDX(nx - 1, J, K)] : dp[IDX(2, J, K)];
            dp[IDX(nx, J, K)] = mp == 0 ? dp[IDX(2, J, K)] : dp[IDX(nx - 1, J, K)];
        }
    for (int K = 1; K <= nz; ++K)
        for (int I = 1; I <= nx; ++I) {
            dp[IDX(I, 1, K)] = lp == 0 ? dp[IDX(I, ny - 1, K)] : dp[IDX(I, 2, K)];
            dp[IDX(I, ny, K)] = lp == 0 ? dp[IDX(I, 2, K)] : dp[IDX(I, ny - 1, K)];
        }
    cplan_free(&P);
    free(spec);
    if (path == 1) zeig_free(&Z);
    free(wz);
    free(wx);
    coefs_free(&C);
    return bad ? -3 : 0;
}

int orc_direct(int nx, int ny, int nz, double delysq, const double* ap, const double* au,
               const double* cp, const double* cw, const double* rp, const double* ru,
               int mp, int lp, int np, double* dp) {
    return orc_direct_path(nx, ny, nz, delysq, ap, au, cp, cw, rp, ru, mp, lp, np, dp, 0);
}

/* ------------------------------------------------------------------------------------ */
/* Per-step velocity noise of the resident time loop (workload definition, not a         */
/* reference function). Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random      */
/* numbers: as easy as 1, 2, 3", SC'11): ten rounds of two 32x32->64 multiplies with the  */
/* published multipliers and Weyl key increments. Gaussians by Box-Muller.                 */
/* ------------------------------------------------------------------------------------ */
void orc_philox4x32_10(const unsigned* ctr, const unsigned* key, unsigned* out) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* Two standard normals from two 32-bit words: u = (w + 1/2) 2^-32 in (0,1),
 * r = sqrt(-2 ln u0), g0 = r cos(2 pi u1), g1 = r sin(2 pi u1). */
static void box_muller(uint32_t a, uint32_t b, double* g0, double* g1) {
    const double u0 = ((double)a + 0.5) * 2.3283064365386962890625e-10;
    const double u1 = ((double)b + 0.5) * 2.3283064365386962890625e-10;
    const double r = sqrt(-2.0 * log(u0));
    const double t = 6.283185307179586476925 * u1;
    *g0 = r * cos(t);
    *g1 = r * sin(t);
}

void orc_normals(unsigned long cell, unsigned step, unsigned seed, double* g) {
    const unsigned ctr[4] = {(unsigned)(cell & 0xffffffffu), (unsigned)(cell >> 32), step, 0u};
    const unsigned key[2] = {seed, 0u};
    unsigned w[4];
    double g3;
    orc_philox4x32_10(ctr, key, w);
    box_muller(w[0], w[1], &g[0], &g[1]);
    box_muller(w[2], w[3], &g[2], &g3);
}

void orc_perturb(int nx, int ny, int nz, const double* u, const double* v, const double* w, double* us,
                 double* vs, double* ws, double sigma, unsigned seed, unsigned step) {
    const size_t plane = (size_t)nx * ny;
    const size_t n = plane * nz;
    memcpy(us, u, n * sizeof(double));
    memcpy(vs, v, n * sizeof(double));
    memcpy(ws, w, n * sizeof(double));
    const int L = ny - 2;
    for (int K = 2; K <= nz - 1; ++K)
        for (int J = 1; J <= ny; ++J)
            for (int I = 2; I <= nx - 1; ++I) {
                /* ghost rows J=1 and J=NY carry the noise of their periodic images J=NY-1, J=2 */
                const int Js = (J == 1) ? L + 1 : ((J == ny) ? 2 : J);
                const size_t o = IDX(I, Js, K);
                const size_t t = IDX(I, J, K);
                double g[3];
                orc_normals((unsigned long)o, step, seed, g);
                if (I <= nx - 2) us[t] = u[o] + sigma * g[0]; /* U faces I=2..NX-2 */
                vs[t] = v[o] + sigma * g[1];
                if (K <= nz - 2) ws[t] = w[o] + sigma * g[2]; /* W faces K=2..NZ-2 */
            }
}

/* ------------------------------------------------------------------------------------ */
/* Pressure update after CORRECT, Eddy6.f:2772-2846, statement by statement.              */
/* ------------------------------------------------------------------------------------ */
int orc_pressure_update(int nx, int ny, int nz, double* p, const double* dp, double alfxdt, double ru1,
                        double dely, const double* ap, const double* au, const double* rp, const double* ru,
                        const double* rvimplx, const double* rvimply, int implx, int imply, int mp, int lp,
                        int itype1, int itype5, int itype6) {
    if (nx < 4 || ny < 4 || nz < 3) return -1;
    const int ix1 = 2, ix2 = nx - 1, jy1 = 2, jy2 = ny - 1, kz1 = 2, kz2 = nz - 1;
    const size_t n = (size_t)nx * ny * nz;
    /* P = P + DP (:2774) */
    for (size_t t = 0; t < n; ++t) p[t] = p[t] + dp[t];
    /* implicit y correction (:2777-2785), I = 2..IX2, J = 2..JY2, K = 2..KZ2 */
    if (imply == 1)
        for (int K = 2; K <= kz2; ++K)
            for (int J = 2; J <= jy2; ++J)
                for (int I = 2; I <= ix2; ++I) {
                    const double r = dely * rp[I - 1];
                    p[IDX(I, J, K)] = p[IDX(I, J, K)] - 0.5 * alfxdt * ru1 / (r * r) *
                                                            (dp[IDX(I, J - 1, K)] - 2. * dp[IDX(I, J, K)] +
                                                             dp[IDX(I, J + 1, K)]) *
                                                            rvimply[I - 1];
                }
    /* implicit x correction (:2787-2796) */
    if (implx == 1)
        for (int K = 2; K <= kz2; ++K)
            for (int J = 2; J <= jy2; ++J)
                for (int I = 2; I <= ix2; ++I)
                    p[IDX(I, J, K)] = p[IDX(I, J, K)] -
                                      0.5 * alfxdt * ru1 * ap[I - 1] / rp[I - 1] *
                                          ((dp[IDX(I + 1, J, K)] - dp[IDX(I, J, K)]) * au[I - 1] * ru[I - 1] -
                                           (dp[IDX(I, J, K)] - dp[IDX(I - 1, J, K)]) * au[I - 2] * ru[I - 2]) *
                                          rvimplx[I - 1];
    /* PREF = sum(P(IX1, JY1:JY2, 2)) / (NY-2); P = P - PREF (:2801-2806) */
    double pref = 0.0;
    for (int J = jy1; J <= jy2; ++J) pref += p[IDX(ix1, J, 2)];
    pref = pref / (double)(ny - 2);
    for (size_t t = 0; t < n; ++t) p[t] = p[t] - pref;
    /* x boundary (:2813-2824) */
    for (int K = 1; K <= nz; ++K)
        for (int J = 1; J <= ny; ++J) {
            if (mp == 0) {
                p[IDX(1, J, K)] = p[IDX(ix2, J, K)];
                p[IDX(nx, J, K)] = p[IDX(ix1, J, K)];
            } else {
                p[IDX(1, J, K)] = -p[IDX(ix1, J, K)];
                p[IDX(nx, J, K)] = -p[IDX(ix2, J, K)];
            }
        }
    if (itype1 == 300)
        for (int K = 1; K <= nz; ++K)
            for (int J = 1; J <= ny; ++J) {
                int js = J + (ny - 2) / 2; /* JSYM, grid.f:467-470 */
                if (js > jy2) js -= ny - 2;
                p[IDX(1, J, K)] = p[IDX(ix1, js, K)];
            }
    /* y boundary (:2829-2835) */
    for (int K = 1; K <= nz; ++K)
        for (int I = 1; I <= nx; ++I) {
            if (lp == 0) {
                p[IDX(I, 1, K)] = p[IDX(I, jy2, K)];
                p[IDX(I, ny, K)] = p[IDX(I, jy1, K)];
            } else {
                p[IDX(I, 1, K)] = p[IDX(I, jy1, K)];
                p[IDX(I, ny, K)] = p[IDX(I, jy2, K)];
            }
        }
    /* z boundary (:2843-2847) */
    for (int J = 1; J <= ny; ++J)
        for (int I = 1; I <= nx; ++I) {
            if (itype5 == 500) p[IDX(I, J, 1)] = p[IDX(I, J, kz2)];
            if (itype6 == 500) p[IDX(I, J, nz)] = p[IDX(I, J, kz1)];
            if (itype5 != 0 && itype5 != 500) p[IDX(I, J, 1)] = p[IDX(I, J, kz1)];
            if (itype6 != 0 && itype6 != 500) p[IDX(I, J, nz)] = p[IDX(I, J, kz2)];
        }
    return 0;
}
