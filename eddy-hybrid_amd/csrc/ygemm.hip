// ygemm.hip -- the periodic-axis real DFT as a dense product on the FP32 matrix cores
// (v_mfma_f32_16x16x4_f32), the alternative to the radix FFT of yfft.hip for the FP32
// transforms of the defect-corrected solve.
//
// Forward (fftpack RFFTF packing, fftpack.f:140-240): c(p) = sum_j F(p, j) x(j) with
//   F(0, j) = 1, F(2m-1, j) = cos(2 pi m j/L), F(2m, j) = -sin(2 pi m j/L), F(L-1, j) = (-1)^j.
// Inverse with DIRECT's 1/L (RFFTB / L, direct.f:1035-1106): x(j) = sum_p G(j, p) c(p) with
//   G(j, 0) = 1/L, G(j, 2m-1) = 2 cos/L, G(j, 2m) = -2 sin/L, G(j, L-1) = (-1)^j / L.
// The matrices are stored transposed (FT[j*L + p] = F(p, j), GT[p*L + j] = G(j, p)) so a
// wave's A-operand loads are 64-byte rows.
//
// A workgroup (4 waves) owns 32 consecutive x1 pencils at one x2 index: the 32 x L input
// block is staged in LDS as FP32, each wave accumulates L/64 row blocks of 16 outputs for the
// two 16-pencil column blocks, the 256 x 256 matrix streams from L2 (256 KiB, shared by every
// workgroup of an XCD). 2 L flops per element: compute-bound on MFMA, where the radix FFT is
// bound by HBM.
//
// MFMA 16x16x4 f32 operand layout (CDNA3/CDNA4): lane l supplies A(l%16, l/16) and
// B(l/16, l%16); the result D(4*(l/16) + r, l%16) is in accumulator register r.
#include "batch_device.h"
#include "kernels.h"
#include "reduce_device.h"
#include "stencil_device.h"

namespace eddy {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kGemmPencils = 32;  // pencils per workgroup (two 16-wide column blocks)
constexpr int kGemmMaxBlk = 8;    // row blocks per wave: L <= 4 * 16 * 8 = 512

struct GArgs {
    const double* src;  // forward: RHS (YF_DIRECT) or RHS of the residual (YF_RESIDUAL)
    const double* phi;  // forward YF_RESIDUAL: current solution
    Op64 op;
    double* part;
    double* red;
    int shift;          // forward: subtract red[RED_VOLB]/red[RED_VOL] from src
    const float* mat;   // FT (forward) or GT (inverse), L x L
    float* S;           // forward output / inverse input (fftpack packing along j)
    double* dp;         // inverse output
    Dims d;
};

// acc[b][c] += A(rows of block) * B(4 rows of X at j0, column block c)
__device__ __forceinline__ f32x4 mfma4(float a, float b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Dense product of one staged tile: out(r, n) = sum_j mat(r, j) X(j, n) for r < L, n < 32,
// X in LDS as X[j * 32 + n]. Row block b of wave w is 16 * (w + 4 b).
__device__ __forceinline__ void gemm_tile(const float* __restrict__ mat, const float* X, int L, int lane, int wave,
                                          f32x4 (&acc)[kGemmMaxBlk][2]) {
#pragma unroll
    for (int b = 0; b < kGemmMaxBlk; ++b) acc[b][0] = acc[b][1] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int nblk = L / 16;
    const int col = lane & 15, kq = lane >> 4;
    for (int j0 = 0; j0 < L; j0 += 4) {
        const int jj = j0 + kq;
        const float b0 = X[jj * kGemmPencils + col];
        const float b1 = X[jj * kGemmPencils + 16 + col];
        const float* mrow = mat + (size_t)jj * L + col;
#pragma unroll
        for (int b = 0; b < kGemmMaxBlk; ++b) {
            const int rb = wave + 4 * b;
            if (rb < nblk) {
                const float a = mrow[16 * rb];
                acc[b][0] = mfma4(a, b0, acc[b][0]);
                acc[b][1] = mfma4(a, b1, acc[b][1]);
            }
        }
    }
}

template <int MODE>
__global__ void __launch_bounds__(256) k_ygemm_fwd(GArgs a) {
    extern __shared__ float X[];  // L x 32 input block
    const Dims d = a.d;
    const int L = d.L;
    const int ntx = gridDim.x;
    const int tile = xcd_tile(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    const int i0 = (tile % ntx) * kGemmPencils;
    const int k = tile / ntx;
    const int tid = threadIdx.x;
    double shift = 0.0;
    if (a.shift) shift = a.red[RED_VOLB] / a.red[RED_VOL];
    double rsq = 0.0;
    const int total = kGemmPencils * L;
    constexpr int U = 8;
    for (int e0 = tid; e0 < total; e0 += 256 * U) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = e0 + u * 256;
            const int ii = e & (kGemmPencils - 1);
            const int j = e >> 5;
            const int gi = i0 + ii;
            v[u] = 0.0;
            if (e < total && gi < d.M) {
                if (MODE == YF_DIRECT) {
                    v[u] = a.src[sidx(d, gi + 1, j + 1, k + 1)] - shift;
                } else {
                    v[u] = (a.src[sidx(d, gi + 1, j + 1, k + 1)] - shift) - lap_at(d, a.op, a.phi, gi, j, k);
                    rsq += v[u] * v[u];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = e0 + u * 256;
            if (e < total) X[e] = (float)v[u];
        }
    }
    if (MODE == YF_RESIDUAL) block_partial<RED_OP_SUM>(rsq, a.part, linear_bid());
    __syncthreads();
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: scalar branches
    f32x4 acc[kGemmMaxBlk][2];
    gemm_tile(a.mat, X, L, lane, wave, acc);
    const int nblk = L / 16;
    const size_t js = (size_t)d.M * (size_t)d.N;
    float* outK = a.S + (size_t)d.M * (size_t)k;
#pragma unroll
    for (int b = 0; b < kGemmMaxBlk; ++b) {
        const int rb = wave + 4 * b;
        if (rb >= nblk) continue;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int gi = i0 + 16 * c + (lane & 15);
            if (gi >= d.M) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int p = 16 * rb + 4 * (lane >> 4) + r;
                outK[gi + js * (size_t)p] = acc[b][c][r];
            }
        }
    }
}

template <int MODE>
__global__ void __launch_bounds__(256) k_ygemm_inv(GArgs a) {
    extern __shared__ float X[];  // L x 32 spectral block
    const Dims d = a.d;
    const int L = d.L;
    const int ntx = gridDim.x;
    const int tile = xcd_tile(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    const int i0 = (tile % ntx) * kGemmPencils;
    const int k = tile / ntx;
    const int tid = threadIdx.x;
    const size_t js = (size_t)d.M * (size_t)d.N;
    const float* inK = a.S + (size_t)d.M * (size_t)k;
    const int total = kGemmPencils * L;
    constexpr int U = 8;
    for (int e0 = tid; e0 < total; e0 += 256 * U) {
        float v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = e0 + u * 256;
            const int gi = i0 + (e & (kGemmPencils - 1));
            const int p = e >> 5;
            v[u] = (e < total && gi < d.M) ? inK[gi + js * (size_t)p] : 0.f;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = e0 + u * 256;
            if (e < total) X[e] = v[u];
        }
    }
    __syncthreads();
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: scalar branches
    f32x4 acc[kGemmMaxBlk][2];
    gemm_tile(a.mat, X, L, lane, wave, acc);
    const int nblk = L / 16;
    const int K = k + 1;
    double psum = 0.0;
#pragma unroll
    for (int b = 0; b < kGemmMaxBlk; ++b) {
        const int rb = wave + 4 * b;
        if (rb >= nblk) continue;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int gi = i0 + 16 * c + (lane & 15);
            if (gi >= d.M) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int j = 16 * rb + 4 * (lane >> 4) + r;
                const size_t o = sidx(d, gi + 1, j + 1, K);
                const double v = (double)acc[b][c][r];
                if (MODE == YI_STORE) {
                    a.dp[o] = v;
                    psum += v;
                } else {
                    const double nv = a.dp[o] + v;
                    a.dp[o] = nv;
                    psum += nv;
                }
            }
        }
    }
    block_partial<RED_OP_SUM>(psum, a.part, linear_bid());
}

bool ygemm_supported(const Dims& d) { return d.L % 16 == 0 && d.L >= 16 && d.L <= 4 * 16 * kGemmMaxBlk; }

void launch_ygemm_fwd(hipStream_t st, const Dims& d, const float* FT, const FwdInput& in, const Op64& op,
                      const Reduce& rd, float* S) {
    GArgs a{};
    a.src = in.src;
    a.phi = in.phi;
    a.op = op;
    a.part = rd.part;
    a.red = rd.red;
    a.shift = in.shift;
    a.mat = FT;
    a.S = S;
    a.d = d;
    const dim3 grid(ceil_div(d.M, kGemmPencils), d.N);
    const size_t sh = (size_t)d.L * kGemmPencils * sizeof(float);
    if (in.mode == YF_RESIDUAL) {
        hipLaunchKernelGGL(k_ygemm_fwd<YF_RESIDUAL>, grid, dim3(256), sh, st, a);
        launch_fold(st, rd.part, (int)(grid.x * grid.y), rd.red + RED_RNORM2, RED_OP_SUM, 0);
    } else {
        hipLaunchKernelGGL(k_ygemm_fwd<YF_DIRECT>, grid, dim3(256), sh, st, a);
    }
}

void launch_ygemm_inv(hipStream_t st, const Dims& d, const float* GT, const float* S, double* dp, const Reduce& rd,
                      int mode) {
    GArgs a{};
    a.part = rd.part;
    a.red = rd.red;
    a.mat = GT;
    a.S = const_cast<float*>(S);
    a.dp = dp;
    a.d = d;
    const dim3 grid(ceil_div(d.M, kGemmPencils), d.N);
    const size_t sh = (size_t)d.L * kGemmPencils * sizeof(float);
    if (mode == YI_STORE) hipLaunchKernelGGL(k_ygemm_inv<YI_STORE>, grid, dim3(256), sh, st, a);
    else hipLaunchKernelGGL(k_ygemm_inv<YI_UPDATE>, grid, dim3(256), sh, st, a);
    launch_fold(st, rd.part, (int)(grid.x * grid.y), rd.red + RED_PHISUM, RED_OP_SUM, 0);
}

}  // namespace eddy
