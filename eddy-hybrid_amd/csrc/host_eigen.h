// host_eigen.h -- symmetric eigen-decompositions on the host (plan set-up of the dense x2
// transform and of the x1 eigenbases): tridiagonal matrices directly, and dense ones (the
// periodic operators, whose corner couplings leave the tridiagonal form) after a Householder
// reduction to tridiagonal form.
//
// Implicit symmetric QR steps with Wilkinson shifts (Golub & Van Loan, Matrix Computations,
// 4th ed., Algorithms 8.3.2 and 8.3.3) on the diagonal d and off-diagonal e directly, the
// Givens rotations accumulated into the eigenvector matrix. A step on the unreduced block
// [lo, hi] chases one bulge down the block; an off-diagonal entry below
// eps * (|d_i| + |d_i+1|) splits the matrix. O(n^2) per eigenvalue, the vectors exactly
// orthogonal up to rounding (products of rotations).
#pragma once

#include <cmath>
#include <vector>

namespace eddy {

// d (n diagonal entries), e (n-1 off-diagonal entries, e[i] couples i and i+1) are
// overwritten. On return w holds the eigenvalues and qt (row-major n x n) the eigenvectors
// as ROWS: qt[m*n + k] = component k of eigenvector m. Returns 0, or -1 if a block did not
// converge in 60 steps.
inline int tridiag_eigen(int n, std::vector<double> d, std::vector<double> e, std::vector<double>& qt,
                         std::vector<double>& w) {
    qt.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) qt[(size_t)i * n + i] = 1.0;
    const double eps = 2.220446049250313e-16;
    auto small = [&](int i) { return std::fabs(e[i]) <= eps * (std::fabs(d[i]) + std::fabs(d[i + 1])); };
    // Q <- Q G on columns (p, p+1) of Q, i.e. rows p, p+1 of qt
    auto rotate = [&](int p, double c, double s) {
        double* a = &qt[(size_t)p * n];
        double* b = a + n;
        for (int k = 0; k < n; ++k) {
            const double x = a[k], y = b[k];
            a[k] = c * x - s * y;
            b[k] = s * x + c * y;
        }
    };
    int hi = n - 1, steps = 0;
    while (hi > 0) {
        if (small(hi - 1)) {  // d[hi] has converged
            e[hi - 1] = 0.0;
            --hi;
            steps = 0;
            continue;
        }
        if (++steps > 60) return -1;
        int lo = hi - 1;
        while (lo > 0 && !small(lo - 1)) --lo;
        // Wilkinson shift: the eigenvalue of the trailing 2x2 block nearer d[hi]
        const double del = 0.5 * (d[hi - 1] - d[hi]);
        const double eh = e[hi - 1];
        const double mu = d[hi] - eh * eh / (del + std::copysign(std::hypot(del, eh), del == 0.0 ? 1.0 : del));
        double x = d[lo] - mu, z = e[lo];
        for (int p = lo; p < hi; ++p) {
            const int q = p + 1;
            // rotation G on (p, q) with c*x - s*z = r, s*x + c*z = 0
            const double h = std::hypot(x, z);
            double c = 1.0, s = 0.0;
            if (h > 0.0) {
                c = x / h;
                s = -z / h;
            }
            if (p > lo) e[p - 1] = h;  // the bulge at (q, p-1) is annihilated
            // G^T [[a, b], [b, cc]] G
            const double a = d[p], b = e[p], cc = d[q];
            d[p] = a * c * c - 2.0 * b * c * s + cc * s * s;
            d[q] = a * s * s + 2.0 * b * c * s + cc * c * c;
            e[p] = (a - cc) * c * s + b * (c * c - s * s);
            rotate(p, c, s);
            if (q < hi) {  // the next bulge at (q+1, p) and the next rotation's pair
                x = e[p];
                z = -s * e[q];
                e[q] = c * e[q];
            }
        }
    }
    w = d;
    return 0;
}

// Symmetric tridiagonal matrix plus the corner couplings of a periodic axis (entries (0, n-1)
// and (n-1, 0) equal to `corner`), as a dense row-major matrix: Householder reduction to
// tridiagonal form A = Q T Q^T (Golub & Van Loan, Algorithm 8.3.1, with the reflectors
// accumulated backwards into Q), then tridiag_eigen on T; the eigenvectors of A are Q times
// those of T. Same output convention as tridiag_eigen. O(n^3).
inline int cyclic_eigen(int n, const std::vector<double>& d0, const std::vector<double>& e0, double corner,
                        std::vector<double>& qt, std::vector<double>& w) {
    std::vector<double> A((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) {
        A[(size_t)i * n + i] = d0[i];
        if (i + 1 < n) {
            A[(size_t)i * n + i + 1] = e0[i];
            A[(size_t)(i + 1) * n + i] = e0[i];
        }
    }
    if (n > 2) {
        A[n - 1] += corner;
        A[(size_t)(n - 1) * n] += corner;
    } else if (n == 2) {
        A[1] += corner;
        A[2] += corner;
    }
    std::vector<std::vector<double>> vs;  // reflector vectors v_k (entries k+1..n-1) and betas
    std::vector<double> betas;
    std::vector<double> v, p, wv;
    for (int k = 0; k + 2 < n; ++k) {
        const int r = n - k - 1;  // length of the column part below the diagonal
        v.assign(r, 0.0);
        double nx = 0.0;
        for (int i = 0; i < r; ++i) {
            v[i] = A[(size_t)(k + 1 + i) * n + k];
            nx += v[i] * v[i];
        }
        nx = std::sqrt(nx);
        double beta = 0.0, alpha = v[0];
        if (nx > 0.0) {
            alpha = v[0] >= 0.0 ? -nx : nx;
            v[0] -= alpha;
            double vv = 0.0;
            for (int i = 0; i < r; ++i) vv += v[i] * v[i];
            beta = vv > 0.0 ? 2.0 / vv : 0.0;
        }
        if (beta != 0.0) {
            // A22 <- H A22 H with H = I - beta v v^T: p = beta A22 v, w = p - (beta/2)(p.v) v
            p.assign(r, 0.0);
            for (int i = 0; i < r; ++i) {
                const double* row = &A[(size_t)(k + 1 + i) * n + k + 1];
                double s = 0.0;
                for (int j = 0; j < r; ++j) s += row[j] * v[j];
                p[i] = beta * s;
            }
            double pv = 0.0;
            for (int i = 0; i < r; ++i) pv += p[i] * v[i];
            wv.assign(r, 0.0);
            for (int i = 0; i < r; ++i) wv[i] = p[i] - 0.5 * beta * pv * v[i];
            for (int i = 0; i < r; ++i) {
                double* row = &A[(size_t)(k + 1 + i) * n + k + 1];
                for (int j = 0; j < r; ++j) row[j] -= v[i] * wv[j] + wv[i] * v[j];
            }
        }
        A[(size_t)(k + 1) * n + k] = alpha;
        A[(size_t)k * n + k + 1] = alpha;
        for (int i = 1; i < r; ++i) {
            A[(size_t)(k + 1 + i) * n + k] = 0.0;
            A[(size_t)k * n + k + 1 + i] = 0.0;
        }
        vs.push_back(v);
        betas.push_back(beta);
    }
    std::vector<double> d(n), e(n > 1 ? n - 1 : 0);
    for (int i = 0; i < n; ++i) {
        d[i] = A[(size_t)i * n + i];
        if (i + 1 < n) e[i] = A[(size_t)(i + 1) * n + i];
    }
    std::vector<double> vt;
    if (tridiag_eigen(n, d, e, vt, w)) return -1;
    // Q = H_0 H_1 ... H_{n-3}, accumulated backwards into a row-major Q
    std::vector<double> Q((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) Q[(size_t)i * n + i] = 1.0;
    for (int k = (int)vs.size() - 1; k >= 0; --k) {
        const double beta = betas[k];
        if (beta == 0.0) continue;
        const std::vector<double>& vk = vs[k];
        const int r = (int)vk.size();
        // rows k+1.. of Q: Q_sub -= beta v (v^T Q_sub)
        std::vector<double> t(n, 0.0);
        for (int i = 0; i < r; ++i) {
            const double* row = &Q[(size_t)(k + 1 + i) * n];
            for (int j = 0; j < n; ++j) t[j] += vk[i] * row[j];
        }
        for (int i = 0; i < r; ++i) {
            double* row = &Q[(size_t)(k + 1 + i) * n];
            const double f = beta * vk[i];
            for (int j = 0; j < n; ++j) row[j] -= f * t[j];
        }
    }
    qt.assign((size_t)n * n, 0.0);
    for (int m = 0; m < n; ++m)
        for (int i = 0; i < n; ++i) {
            const double* qrow = &Q[(size_t)i * n];
            const double* vrow = &vt[(size_t)m * n];
            double s = 0.0;
            for (int k = 0; k < n; ++k) s += qrow[k] * vrow[k];
            qt[(size_t)m * n + i] = s;
        }
    return 0;
}

}  // namespace eddy
