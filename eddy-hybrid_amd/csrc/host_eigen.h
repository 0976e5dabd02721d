// host_eigen.h -- eigen-decomposition of a symmetric tridiagonal matrix on the host (plan
// set-up of the dense x2 transform).
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

}  // namespace eddy
