// zgemm.hip -- the dense transform of a stretched-x2 grid as batched matrix products, in two
// forms.
//
// x2 form: the x2 part of DIRECT's operator (AN/BN/CN, direct.f:370-378) is diagonalised by
// the eigenvectors of its generalized symmetric form (plan set-up, on the host). Applying
// them is a plain matrix product per y-mode j: the slab S_j (M x N, x1 fastest, leading
// dimension M) times the N x N eigenvector matrix, then the x1 lines are solved.
// x1 form: the x1 operator of each y-mode (AM/BM - AK_j/RP^2/CM, direct.f:344-352, 536) is
// diagonalised instead, S_j^T times that mode's M x M basis, then the x2 lines are solved.
// A product costs 2 x (the diagonalised length) flops per element, so the plan picks the
// shorter axis (the production grid: M = 300 against N = 1024). Either way it is one
// strided-batched rocBLAS GEMM over the y-modes (out of place, S <-> S2) on the FP64 matrix
// cores. Around the products, two small kernels apply what the fused kernel of small N does
// inside its tiles:
//   forward: the solvability shift of the zero y-mode (L * VOLB/VOL subtracted from slab 0);
//   inverse: the plain-mean correction of direct.f:894-898, delta added to the null mode's
//            line of slab 0 (delta from the line solve's line sums, as k_zdense / k_zdct_inv).
#include <rocblas/rocblas.h>

#include "kernels.h"

namespace eddy {

template <typename T>
__global__ void __launch_bounds__(256) k_zgemm_shift(T* S0, long n, const double* red, double L) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) S0[e] -= T(L * red[RED_VOLB] / red[RED_VOL]);
}

template <typename T>
__global__ void __launch_bounds__(256) k_zgemm_delta(T* X0, const T* linesum, const T* wsum, int M, int N, int null_m) {
    __shared__ double part[4];
    double s = 0.0;
    for (int m = threadIdx.x; m < N; m += blockDim.x) s += (double)wsum[m] * (double)linesum[m];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    const double tot = part[0] + part[1] + part[2] + part[3];
    const T delta = T(-tot / ((double)wsum[null_m] * (double)M));
    for (int i = threadIdx.x; i < M; i += blockDim.x) X0[i + (size_t)M * null_m] += delta;
}

static rocblas_operation op(bool t) { return t ? rocblas_operation_transpose : rocblas_operation_none; }

// C_b = op(A_b) op(B_b) for b < nb (column-major, strided batches)
static rocblas_status gemm(rocblas_handle h, bool ta, bool tb, int m, int n, int k, const double* A, int lda,
                           long sa, const double* B, int ldb, long sb, double* C, int ldc, long sc, int nb) {
    const double one = 1.0, zero = 0.0;
    return rocblas_dgemm_strided_batched(h, op(ta), op(tb), m, n, k, &one, A, lda, sa, B, ldb, sb, &zero, C, ldc, sc,
                                         nb);
}
static rocblas_status gemm(rocblas_handle h, bool ta, bool tb, int m, int n, int k, const float* A, int lda, long sa,
                           const float* B, int ldb, long sb, float* C, int ldc, long sc, int nb) {
    const float one = 1.0f, zero = 0.0f;
    return rocblas_sgemm_strided_batched(h, op(ta), op(tb), m, n, k, &one, A, lda, sa, B, ldb, sb, &zero, C, ldc, sc,
                                         nb);
}
static rocblas_status gemm(rocblas_handle h, int M, int N, const double* A, const double* B, double* C, int nb) {
    return gemm(h, false, true, M, N, N, A, M, (long)M * N, B, N, 0, C, M, (long)M * N, nb);
}
static rocblas_status gemm(rocblas_handle h, int M, int N, const float* A, const float* B, float* C, int nb) {
    return gemm(h, false, true, M, N, N, A, M, (long)M * N, B, N, 0, C, M, (long)M * N, nb);
}

// Forward: S2_j = S_j Qf for modes [j0, j0 + nj), with Qf(k, m) = qf[k*N + m] (row-major, so
// the GEMM takes it transposed from column-major storage).
template <typename T>
int launch_zgemm_fwd(hipStream_t st, void* blas, const Dims& d, const Tables<T>& tb, T* S, T* S2,
                     const double* shift_red, int j0, int nj) {
    const size_t slab = (size_t)d.M * d.N;
    if (shift_red && j0 == 0)
        hipLaunchKernelGGL(k_zgemm_shift<T>, dim3((unsigned)((slab + 255) / 256)), dim3(256), 0, st, S, (long)slab,
                           shift_red, (double)d.L);
    const rocblas_status r = gemm((rocblas_handle)blas, d.M, d.N, S + slab * j0, tb.qf, S2 + slab * j0, nj);
    return r == rocblas_status_success ? 0 : -1;
}

// Inverse: S_j = S2_j Qi for modes [j0, j0 + nj), Qi(m, k) = qiT[m*N + k]; slab 0 first gets
// the plain-mean correction of its null mode.
template <typename T>
int launch_zgemm_inv(hipStream_t st, void* blas, const Dims& d, const Tables<T>& tb, T* S2, T* S, const T* linesum,
                     int j0, int nj) {
    const size_t slab = (size_t)d.M * d.N;
    if (j0 == 0)
        hipLaunchKernelGGL(k_zgemm_delta<T>, dim3(1), dim3(256), 0, st, S2, linesum, tb.wsum, d.M, d.N, d.null_m);
    const rocblas_status r = gemm((rocblas_handle)blas, d.M, d.N, S2 + slab * j0, tb.qiT, S + slab * j0, nj);
    return r == rocblas_status_success ? 0 : -1;
}

// The x1 form (x2 stretched, x1 shorter): slot j's x1 eigenbasis is tz.qf + j*M*M (column-
// major QF_j(i, m), X_j(m) = sum_i QF_j(i, m) x(i)) and tz.qiT + j*M*M (QI_j(i, m),
// x(i) = sum_m QI_j(i, m) X_j(m)). Forward: S2_j (N x M, x2 fastest) = S_j^T QF_j, so that
// the x2 lines of each x1 mode are contiguous for the line solve; inverse: S_j = QI_j S2_j^T.
template <typename T>
int launch_xgemm_fwd(hipStream_t st, void* blas, const Dims& d, const Tables<T>& tz, T* S, T* S2,
                     const double* shift_red, int j0, int nj) {
    const size_t slab = (size_t)d.M * d.N, mm = (size_t)d.M * d.M;
    if (shift_red && j0 == 0)
        hipLaunchKernelGGL(k_zgemm_shift<T>, dim3((unsigned)((slab + 255) / 256)), dim3(256), 0, st, S, (long)slab,
                           shift_red, (double)d.L);
    const rocblas_status r = gemm((rocblas_handle)blas, true, false, d.N, d.M, d.M, S + slab * j0, d.M, (long)slab,
                                  tz.qf + mm * j0, d.M, (long)mm, S2 + slab * j0, d.N, (long)slab, nj);
    return r == rocblas_status_success ? 0 : -1;
}

// Inverse of the x1 form; slab 0 first gets the plain-mean correction on its null x1 mode (the
// x2 line of that mode), with the line sums of the x2 solve and the column sums tz.wsum of QI_0.
template <typename T>
int launch_xgemm_inv(hipStream_t st, void* blas, const Dims& d, const Tables<T>& tz, T* S2, T* S, const T* linesum,
                     int null_x, int j0, int nj) {
    const size_t slab = (size_t)d.M * d.N, mm = (size_t)d.M * d.M;
    if (j0 == 0)
        hipLaunchKernelGGL(k_zgemm_delta<T>, dim3(1), dim3(256), 0, st, S2, linesum, tz.wsum, d.N, d.M, null_x);
    const rocblas_status r = gemm((rocblas_handle)blas, false, true, d.M, d.N, d.M, tz.qiT + mm * j0, d.M, (long)mm,
                                  S2 + slab * j0, d.N, (long)slab, S + slab * j0, d.M, (long)slab, nj);
    return r == rocblas_status_success ? 0 : -1;
}

int zgemm_create(void** blas, hipStream_t st) {
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success) return -1;
    // deterministic sums (no atomics-based split reductions), the plan's stream
    if (rocblas_set_atomics_mode(h, rocblas_atomics_not_allowed) != rocblas_status_success ||
        rocblas_set_stream(h, st) != rocblas_status_success) {
        rocblas_destroy_handle(h);
        return -1;
    }
    *blas = h;
    return 0;
}

int zgemm_set_stream(void* blas, hipStream_t st) {
    return rocblas_set_stream((rocblas_handle)blas, st) == rocblas_status_success ? 0 : -1;
}

void zgemm_destroy(void* blas) {
    if (blas) rocblas_destroy_handle((rocblas_handle)blas);
}

template int launch_zgemm_fwd<double>(hipStream_t, void*, const Dims&, const Tables<double>&, double*, double*,
                                      const double*, int, int);
template int launch_zgemm_fwd<float>(hipStream_t, void*, const Dims&, const Tables<float>&, float*, float*,
                                     const double*, int, int);
template int launch_zgemm_inv<double>(hipStream_t, void*, const Dims&, const Tables<double>&, double*, double*,
                                      const double*, int, int);
template int launch_zgemm_inv<float>(hipStream_t, void*, const Dims&, const Tables<float>&, float*, float*,
                                     const float*, int, int);
template int launch_xgemm_fwd<double>(hipStream_t, void*, const Dims&, const Tables<double>&, double*, double*,
                                      const double*, int, int);
template int launch_xgemm_fwd<float>(hipStream_t, void*, const Dims&, const Tables<float>&, float*, float*,
                                     const double*, int, int);
template int launch_xgemm_inv<double>(hipStream_t, void*, const Dims&, const Tables<double>&, double*, double*,
                                      const double*, int, int, int);
template int launch_xgemm_inv<float>(hipStream_t, void*, const Dims&, const Tables<float>&, float*, float*,
                                     const float*, int, int, int);

}  // namespace eddy
