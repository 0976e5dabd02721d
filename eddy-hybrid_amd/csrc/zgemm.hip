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
// shorter axis (the production grid: M = 300 against N = 1024). Either way it is one batched
// product over the y-modes (out of place, S <-> S2) on the matrix cores: the library's own
// MFMA kernel (k_mma below), or a strided-batched rocBLAS GEMM when the plan asks for it.
// The MFMA kernel applies, inside its operand loads, what the fused kernel of small N does
// inside its tiles (rocBLAS needs two small kernels for it):
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

// ---- the products on the matrix cores --------------------------------------------------------
//
// C_b(r, c) = sum_p Aop_b(r, p) Bop_b(p, c), r < R, c < Cn, p < P, b < nb, with C column-major
// (C[r + c*ldc]). Both operands are stored either contiguous along p ("p-major": Aop(r, p) =
// A[p + r*lda], Bop(p, c) = B[p + c*ldb]; the x1-form forward product) or contiguous along
// their own index ("x-major": A[r + p*lda], B[c + p*ldb]; every other product).
//
// A workgroup (4 waves, 2 x 2) owns a BR x BC tile of C, each wave a (BR/2) x (BC/2) block of
// 16 x 16 MFMA tiles (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32). The contraction runs in
// chunks of KC staged in LDS, in two buffers: the next chunk's global loads are in flight
// while the MFMAs of the current one run, one barrier per chunk.
// MFMA 16x16x4 lane layout: lane l supplies A(l%16, l/16) and B(l/16, l%16); accumulator
// element v of lane l is D(4*(l/16) + v, l%16) for f32 and D(l/16 + 4v, l%16) for f64 (the
// f64 rows interleave across the four lane groups; checked on the device). The MFMA takes
// Bop^T as its A input and Aop^T as its B input, so lane x + 16q holds C(r0 + x, c0 + row(q,
// v)): 16 lanes store 16 consecutive rows of a column of C. The MFMA's k index 4s + q (step s of a chunk, lane
// group q) stands for chunk offset q*(KC/4) + s -- a permutation of the contraction, exact
// in any order -- so that a lane reads KC/4 consecutive p of a p-major operand at once.
template <typename T>
struct MmArgs {
    const T* A;
    const T* B;
    T* C;
    long sA, sB, sC;  // batch strides (elements)
    int lda, ldb, ldc;
    int R, Cn, P, nb;
    int ntr, ntc;     // tiles along r and along c
    int zb;           // batch of the zero y-mode (-1: not in this launch)
    // forward: operand A of the zero mode minus shiftL * red[RED_VOLB] / red[RED_VOL]
    const double* shift_red;
    double shiftL;
    // inverse: operand dop (0 A, 1 B) of the zero mode plus delta where p == dnull, delta =
    // -sum_{m < dn} wsum[m] linesum[m] / (wsum[dnull] * dnorm)
    const T* linesum;
    const T* wsum;
    int dn, dnull, dop;
    double dnorm;
};

template <typename T>
struct Mma;
template <>
struct Mma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    typedef double v16b __attribute__((ext_vector_type(2)));
    static __device__ __forceinline__ v4 run(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int q, int v) { return q + 4 * v; }
};
template <>
struct Mma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    typedef float v16b __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 run(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int q, int v) { return 4 * q + v; }
};

template <typename T>
struct MmShape {
    static constexpr int KC = sizeof(T) == 8 ? 16 : 32;  // contraction chunk: 128 B of p
    static constexpr int KS = KC / 4;                    // MFMA steps per chunk
    static constexpr int QP = KC + 16 / (int)sizeof(T);  // p-major LDS row pitch
    static constexpr int XPAD = sizeof(T) == 8 ? 4 : 2;  // x-major LDS row pad
};

// LDS: p-major operands as rows X[x][p'] (pitch QP: a lane's KS consecutive p' are one or two
// 16-byte reads, and the 16 rows a read touches fall in distinct banks); x-major operands as
// rows X[p'][x] (pitch BX + XPAD: the four lane groups of a read fall in distinct banks).
// Thread tid's LOADS slots of a chunk: p-major, p = tid % KC (every slot) and x = tid / KC +
// u * 256/KC; x-major, x = tid % BX and p = tid / BX + u * 256/BX. A slot is a 32-bit byte
// offset from the chunk's (uniform) base; rows past the operand's end read its last row
// instead (they feed only entries of C that are never stored), and only the last chunk checks
// the contraction bound (zeros past P).
template <typename T, bool PM, int BX>
struct MmOperand {
    typedef MmShape<T> S;
    static constexpr int ELEMS = PM ? BX * S::QP : S::KC * (BX + S::XPAD);
    static constexpr int LOADS = BX * S::KC / 256;  // global loads per thread and chunk
    static constexpr int XSTEP = PM ? 256 / S::KC : 0;
    static constexpr int PSTEP = PM ? 0 : 256 / BX;

    unsigned off[LOADS];  // byte offsets of the slots in chunk 0
    int p;                // contraction offset of slot 0 within a chunk
    int xt;               // tile row of slot 0
    size_t cstride;       // bytes from one chunk to the next

    __device__ __forceinline__ void init(int ld, int x0, int Xn, int tid) {
        if (PM) {
            p = tid % S::KC;
            xt = tid / S::KC;
#pragma unroll
            for (int u = 0; u < LOADS; ++u) {
                const int gx = min(x0 + xt + u * XSTEP, Xn - 1);
                off[u] = (unsigned)(((size_t)p + (size_t)gx * ld) * sizeof(T));
            }
            cstride = (size_t)S::KC * sizeof(T);
        } else {
            p = tid / BX;
            xt = tid % BX;
            const int gx = min(x0 + xt, Xn - 1);
#pragma unroll
            for (int u = 0; u < LOADS; ++u) off[u] = (unsigned)(((size_t)gx + (size_t)(p + u * PSTEP) * ld) * sizeof(T));
            cstride = (size_t)S::KC * ld * sizeof(T);
        }
    }
    __device__ __forceinline__ int slot_p(int u) const { return p + u * PSTEP; }
    // global -> registers, chunk c
    __device__ __forceinline__ void load(const T* X, int c, int P, T (&r)[LOADS]) const {
        const char* base = reinterpret_cast<const char*>(X) + (size_t)c * cstride;
        const int p0 = c * S::KC;
        if (p0 + S::KC <= P) {
#pragma unroll
            for (int u = 0; u < LOADS; ++u) r[u] = *reinterpret_cast<const T*>(base + off[u]);
        } else {
#pragma unroll
            for (int u = 0; u < LOADS; ++u)
                r[u] = (p0 + slot_p(u) < P) ? *reinterpret_cast<const T*>(base + off[u]) : T(0);
        }
    }
    // the zero mode's adjustments: every element minus `shift` (the other operand is zero past
    // P), and `delta` where p == dnull
    __device__ __forceinline__ void adjust(int c, T shift, T delta, int dnull, T (&r)[LOADS]) const {
        const int p0 = c * S::KC;
#pragma unroll
        for (int u = 0; u < LOADS; ++u) r[u] = r[u] - shift + (p0 + slot_p(u) == dnull ? delta : T(0));
    }
    __device__ __forceinline__ void store(T* lds, const T (&r)[LOADS]) const {
#pragma unroll
        for (int u = 0; u < LOADS; ++u) {
            if (PM) lds[(xt + u * XSTEP) * S::QP + p] = r[u];
            else lds[(p + u * PSTEP) * (BX + S::XPAD) + xt] = r[u];
        }
    }
    // lane (x, q): the KS values of row xr (within the tile) at chunk offsets q*KS .. q*KS+KS-1
    static __device__ __forceinline__ void read(const T* lds, int xr, int q, T (&v)[S::KS]) {
        if (PM) {
            typedef typename Mma<T>::v16b V;
            constexpr int W = 16 / sizeof(T);
            const T* row = lds + xr * S::QP + q * S::KS;
#pragma unroll
            for (int i = 0; i < S::KS / W; ++i) {
                const V t = *reinterpret_cast<const V*>(row + i * W);
#pragma unroll
                for (int j = 0; j < W; ++j) v[i * W + j] = t[j];
            }
        } else {
#pragma unroll
            for (int s = 0; s < S::KS; ++s) v[s] = lds[(q * S::KS + s) * (BX + S::XPAD) + xr];
        }
    }
};

template <typename T, bool PM, int BR, int BC>
__global__ void __launch_bounds__(256) k_mma(MmArgs<T> a) {
    typedef MmShape<T> S;
    typedef MmOperand<T, PM, BR> OA;
    typedef MmOperand<T, PM, BC> OB;
    typedef typename Mma<T>::v4 v4;
    constexpr int NR = BR / 32, NC = BC / 32;  // 16 x 16 tiles per wave along r and c
    extern __shared__ __align__(16) unsigned char mm_lds[];
    T* lds = reinterpret_cast<T*>(mm_lds);  // 2 x (OA::ELEMS + OB::ELEMS)
    __shared__ double red4[4];

    const int tid = threadIdx.x;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int ct = tile % a.ntc;
    const int rest = tile / a.ntc;
    const int rt = rest % a.ntr;
    const int b = rest / a.ntr;
    const int r0 = rt * BR, c0 = ct * BC;
    const T* A = a.A + (size_t)b * a.sA;
    const T* B = a.B + (size_t)b * a.sB;
    T* C = a.C + (size_t)b * a.sC;

    // the zero mode: shift and delta (the same for every workgroup of that batch)
    const bool zm = (b == a.zb);
    T shift = T(0), delta = T(0);
    if (zm) {
        if (a.shift_red) shift = T(a.shiftL * a.shift_red[RED_VOLB] / a.shift_red[RED_VOL]);
        if (a.linesum) {
            double s = 0.0;
            for (int m = tid; m < a.dn; m += 256) s += (double)a.wsum[m] * (double)a.linesum[m];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
            if ((tid & 63) == 0) red4[tid >> 6] = s;
            __syncthreads();
            const double tot = red4[0] + red4[1] + red4[2] + red4[3];
            delta = T(-tot / ((double)a.wsum[a.dnull] * a.dnorm));
        }
    }
    const T deltaA = (a.dop == 0) ? delta : T(0), deltaB = (a.dop == 1) ? delta : T(0);

    const int lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: scalar branches
    const int x = lane & 15, q = lane >> 4;
    const int wr = (wv & 1) * (BR / 2), wc = (wv >> 1) * (BC / 2);
    // 16 x 16 tiles of the wave inside C: the partial tiles at the ends of R and Cn run no
    // MFMAs for their blocks that lie wholly outside (wave-uniform)
    const int nri = min(NR, max(0, (a.R - r0 - wr + 15) / 16));
    const int nci = min(NC, max(0, (a.Cn - c0 - wc + 15) / 16));

    v4 acc[NR][NC];
#pragma unroll
    for (int i = 0; i < NR; ++i)
#pragma unroll
        for (int j = 0; j < NC; ++j) acc[i][j] = v4{T(0), T(0), T(0), T(0)};

    OA oa;
    OB ob;
    oa.init(a.lda, r0, a.R, tid);
    ob.init(a.ldb, c0, a.Cn, tid);
    // chunk kc + 1 is loaded into registers while chunk kc's MFMAs run from LDS, then goes to
    // the other LDS buffer; one barrier per chunk
    const int nk = (a.P + S::KC - 1) / S::KC;
    T ra[OA::LOADS], rb[OB::LOADS];
    auto put = [&](int kc) {
        if (zm) {
            oa.adjust(kc, shift, deltaA, a.dnull, ra);
            ob.adjust(kc, T(0), deltaB, a.dnull, rb);
        }
        T* nA = lds + (kc & 1) * (OA::ELEMS + OB::ELEMS);
        oa.store(nA, ra);
        ob.store(nA + OA::ELEMS, rb);
    };
    oa.load(A, 0, a.P, ra);
    ob.load(B, 0, a.P, rb);
    put(0);
    __syncthreads();
    for (int kc = 0; kc < nk; ++kc) {
        const bool more = kc + 1 < nk;
        if (more) {
            oa.load(A, kc + 1, a.P, ra);
            ob.load(B, kc + 1, a.P, rb);
        }
        const T* As = lds + (kc & 1) * (OA::ELEMS + OB::ELEMS);
        const T* Bs = As + OA::ELEMS;
        T av[NR][S::KS], bv[NC][S::KS];
#pragma unroll
        for (int i = 0; i < NR; ++i) OA::read(As, wr + 16 * i + x, q, av[i]);
#pragma unroll
        for (int j = 0; j < NC; ++j) OB::read(Bs, wc + 16 * j + x, q, bv[j]);
        if (nri == NR && nci == NC) {
#pragma unroll
            for (int s = 0; s < S::KS; ++s)
#pragma unroll
                for (int i = 0; i < NR; ++i)
#pragma unroll
                    for (int j = 0; j < NC; ++j) acc[i][j] = Mma<T>::run(bv[j][s], av[i][s], acc[i][j]);
        } else {
#pragma unroll
            for (int s = 0; s < S::KS; ++s)
#pragma unroll
                for (int i = 0; i < NR; ++i)
#pragma unroll
                    for (int j = 0; j < NC; ++j)
                        if (i < nri && j < nci) acc[i][j] = Mma<T>::run(bv[j][s], av[i][s], acc[i][j]);
        }
        if (more) put(kc + 1);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        const int gr = r0 + wr + 16 * i + x;
        if (gr >= a.R) continue;
#pragma unroll
        for (int j = 0; j < NC; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int gc = c0 + wc + 16 * j + Mma<T>::row(q, v);
                if (gc < a.Cn) C[(size_t)gr + (size_t)gc * a.ldc] = acc[i][j][v];
            }
    }
}

// ---- FP64 products with the operand tiles staged by LDS DMA ----------------------------------
//
// The same product and lane layout as k_mma (BR x BC tiles of 64 or 128, 4 waves of
// (BR/2) x (BC/2), KC = 16), with the chunks copied from global memory straight into LDS by
// global_load_lds_dwordx4 (lane l's 16 bytes land at the instruction's LDS base + 16 l): no
// registers hold a chunk in flight, and two LDS stages let chunk kc+1 load while chunk kc's
// MFMAs run, one barrier per chunk. Each wave copies 1 KiB of an operand's stage region per
// instruction, BX/32 instructions per operand and chunk. LDS rows are XOR-swizzled in 16-byte
// units so that the MFMA operand reads are conflict-free: p-major rows (BX x 128 B), chunk c
// of row x at c ^ ((x >> 1) & 7); x-major rows (16 x 8 BX bytes), chunk c of row p at
// c ^ (((p >> 2) & 1) << 3).
// The copies cannot zero-fill: rows past the operand's end read its last row pair, contraction
// indices past P read index P-2 (p-major) or row P-1 (x-major), and the reads of the last
// chunk zero every p >= P; the zero mode's shift and delta are applied on the same reads.
// Taken when every 16-byte pair lies inside one row and is aligned (even leading dimensions,
// P and operand lengths, 16-byte aligned bases and batch strides).
// one chunk of a BX-wide FP64 operand tile: BX x 16 values, 128 bytes per 64 of them
template <int BX>
constexpr int dma_tile_bytes() { return BX * 128; }

template <bool PM, int BX>
struct DmaOperand {
    static constexpr int NI = BX / 32;  // copy instructions per wave and chunk (4 waves, 1 KiB each)
    static constexpr int LPR = BX / 2;  // x-major: lanes per row of BX values
    // lane's global element offset for instruction j of its wave in chunk 0 (rows clamped)
    long base[NI];
    int pj[NI];  // p-major: 2c (pair start within the chunk); x-major: row p
    int ld, P;

    __device__ __forceinline__ void init(int ld_, int x0, int Xn, int P_, int w, int l) {
        ld = ld_;
        P = P_;
#pragma unroll
        for (int j = 0; j < NI; ++j) {
            const int g = w * NI + j;  // 1 KiB piece of the tile's stage region
            if (PM) {
                const int xr = g * 8 + (l >> 3);
                const int c = (l & 7) ^ ((xr >> 1) & 7);
                const int gx = min(x0 + xr, Xn - 1);
                pj[j] = 2 * c;
                base[j] = (long)gx * ld;
            } else {
                const int pr = g * (64 / LPR) + l / LPR;
                const int c = (l % LPR) ^ (((pr >> 2) & 1) << 3);
                const int gx = min(x0 + 2 * c, Xn - 2);
                pj[j] = pr;
                base[j] = gx;
            }
        }
    }
    // instruction j's source for the chunk at p0
    __device__ __forceinline__ const double* src(const double* X, int j, int p0) const {
        if (PM) return X + base[j] + min(p0 + pj[j], P - 2);
        return X + base[j] + (long)min(p0 + pj[j], P - 1) * ld;
    }
};

template <bool PM, int BX>
__device__ __forceinline__ void dma_read(const char* stage, int xr, int q, double (&v)[4]) {
    if (PM) {
        // lane (x, q) reads p = 4q .. 4q+3 of row xr: chunks 2q and 2q+1
        typedef double d2 __attribute__((ext_vector_type(2)));
        const int sw = (xr >> 1) & 7;
        const d2 lo = *reinterpret_cast<const d2*>(stage + xr * 128 + (((2 * q) ^ sw) << 4));
        const d2 hi = *reinterpret_cast<const d2*>(stage + xr * 128 + (((2 * q + 1) ^ sw) << 4));
        v[0] = lo[0];
        v[1] = lo[1];
        v[2] = hi[0];
        v[3] = hi[1];
    } else {
        // lane (x, q) reads element xr of rows p = 4q .. 4q+3 (rows of BX values)
        const int pos = ((xr >> 1) ^ ((q & 1) << 3)) << 4;
#pragma unroll
        for (int s = 0; s < 4; ++s)
            v[s] = *reinterpret_cast<const double*>(stage + (4 * q + s) * (BX * 8) + pos + ((xr & 1) << 3));
    }
}

// s_waitcnt with only the vector-memory count: vmcnt <= n (n < 16), expcnt and lgkmcnt free
template <int N>
__device__ __forceinline__ void wait_vm() {
    __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8));
}

// BR x BC tiles (64 or 128 each), two LDS stages: the copies of chunk kc+1 run while chunk
// kc's MFMAs do
template <bool PM, int BR, int BC>
__global__ void __launch_bounds__(256) k_mma_dma(MmArgs<double> a) {
    typedef Mma<double>::v4 v4;
    constexpr int NR = BR / 32, NC = BC / 32;
    constexpr int ABYTES = dma_tile_bytes<BR>(), STAGE = ABYTES + dma_tile_bytes<BC>();
    // the two stages and nothing else (64 x 64: 32 KiB, room for 5 workgroups per CU); the
    // zero mode's four partial sums use stage 0 before its first copy
    extern __shared__ __align__(16) unsigned char mm_lds[];
    double* red4 = reinterpret_cast<double*>(mm_lds);
    const int tid = threadIdx.x;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int ct = tile % a.ntc;
    const int rt = (tile / a.ntc) % a.ntr;
    const int b = tile / (a.ntr * a.ntc);
    const int r0 = rt * BR, c0 = ct * BC;
    const double* A = a.A + (size_t)b * a.sA;
    const double* B = a.B + (size_t)b * a.sB;
    double* C = a.C + (size_t)b * a.sC;

    const bool zm = (b == a.zb);
    double shift = 0.0, delta = 0.0;
    if (zm) {
        if (a.shift_red) shift = a.shiftL * a.shift_red[RED_VOLB] / a.shift_red[RED_VOL];
        if (a.linesum) {
            double s = 0.0;
            for (int m = tid; m < a.dn; m += 256) s += a.wsum[m] * a.linesum[m];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
            if ((tid & 63) == 0) red4[tid >> 6] = s;
            __syncthreads();
            const double tot = red4[0] + red4[1] + red4[2] + red4[3];
            delta = -tot / (a.wsum[a.dnull] * a.dnorm);
            __syncthreads();  // every wave has read red4 before the copies overwrite it
        }
    }
    const double deltaA = (a.dop == 0) ? delta : 0.0, deltaB = (a.dop == 1) ? delta : 0.0;

    const int lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: scalar branches
    const int x = lane & 15, q = lane >> 4;
    const int wr = (wv & 1) * (BR / 2), wc = (wv >> 1) * (BC / 2);
    const int nri = min(NR, max(0, (a.R - r0 - wr + 15) / 16));
    const int nci = min(NC, max(0, (a.Cn - c0 - wc + 15) / 16));

    DmaOperand<PM, BR> oa;
    DmaOperand<PM, BC> ob;
    oa.init(a.lda, r0, a.R, a.P, wv, lane);
    ob.init(a.ldb, c0, a.Cn, a.P, wv, lane);
    typedef __attribute__((address_space(3))) void* lptr;
    auto issue = [&](int c) {
        char* st = reinterpret_cast<char*>(mm_lds) + (c & 1) * STAGE;
        const int p0 = c * 16;
#pragma unroll
        for (int j = 0; j < DmaOperand<PM, BR>::NI; ++j)
            __builtin_amdgcn_global_load_lds((const void*)oa.src(A, j, p0),
                                             (lptr)(st + (wv * DmaOperand<PM, BR>::NI + j) * 1024), 16, 0, 0);
#pragma unroll
        for (int j = 0; j < DmaOperand<PM, BC>::NI; ++j)
            __builtin_amdgcn_global_load_lds((const void*)ob.src(B, j, p0),
                                             (lptr)(st + ABYTES + (wv * DmaOperand<PM, BC>::NI + j) * 1024), 16, 0,
                                             0);
    };

    v4 acc[NR][NC];
#pragma unroll
    for (int i = 0; i < NR; ++i)
#pragma unroll
        for (int j = 0; j < NC; ++j) acc[i][j] = v4{0.0, 0.0, 0.0, 0.0};

    const int nk = (a.P + 15) / 16;
    issue(0);
    for (int kc = 0; kc < nk; ++kc) {
        // this wave's copies of chunk kc are done; after the barrier every wave's are, and
        // every wave has finished reading the stage the next copy overwrites. A bare s_barrier:
        // __syncthreads' fence would also wait for the copies still in flight (none here, but
        // the same holds with more stages). The empty asm statements keep the compiler from
        // moving LDS accesses across it.
        wait_vm<0>();
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        if (kc + 1 < nk) issue(kc + 1);
        const char* st = reinterpret_cast<const char*>(mm_lds) + (kc & 1) * STAGE;
        double av[NR][4], bv[NC][4];
#pragma unroll
        for (int i = 0; i < NR; ++i) dma_read<PM, BR>(st, wr + 16 * i + x, q, av[i]);
#pragma unroll
        for (int j = 0; j < NC; ++j) dma_read<PM, BC>(st + ABYTES, wc + 16 * j + x, q, bv[j]);
        const int p0 = kc * 16;
        if (zm || p0 + 16 > a.P) {
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int p = p0 + 4 * q + s;
                const bool ok = p < a.P;
#pragma unroll
                for (int i = 0; i < NR; ++i) av[i][s] = ok ? av[i][s] - shift + (p == a.dnull ? deltaA : 0.0) : 0.0;
#pragma unroll
                for (int j = 0; j < NC; ++j) bv[j][s] = ok ? bv[j][s] + (p == a.dnull ? deltaB : 0.0) : 0.0;
            }
        }
        if (nri == NR && nci == NC) {
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int i = 0; i < NR; ++i)
#pragma unroll
                    for (int j = 0; j < NC; ++j) acc[i][j] = Mma<double>::run(bv[j][s], av[i][s], acc[i][j]);
        } else {
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int i = 0; i < NR; ++i)
#pragma unroll
                    for (int j = 0; j < NC; ++j)
                        if (i < nri && j < nci) acc[i][j] = Mma<double>::run(bv[j][s], av[i][s], acc[i][j]);
        }
    }
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        const int gr = r0 + wr + 16 * i + x;
        if (gr >= a.R) continue;
#pragma unroll
        for (int j = 0; j < NC; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int gc = c0 + wc + 16 * j + Mma<double>::row(q, v);
                if (gc < a.Cn) C[(size_t)gr + (size_t)gc * a.ldc] = acc[i][j][v];
            }
    }
}

template <bool PM, int BR = 64, int BC = 64>
static void launch_mma_dma(hipStream_t st, MmArgs<double> a) {
    a.ntr = ceil_div(a.R, BR);
    a.ntc = ceil_div(a.Cn, BC);
    const long nwg = (long)a.ntr * a.ntc * a.nb;
    const size_t sh = 2 * (size_t)(dma_tile_bytes<BR>() + dma_tile_bytes<BC>());
    (void)hipFuncSetAttribute((const void*)k_mma_dma<PM, BR, BC>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)sh);
    hipLaunchKernelGGL((k_mma_dma<PM, BR, BC>), dim3((unsigned)nwg), dim3(256), sh, st, a);
}

// Every 16-byte copy of the DMA kernel aligned and inside one row of its operand.
static bool dma_ok(const MmArgs<double>& a) {
    auto al = [](const void* p) { return ((uintptr_t)p & 15) == 0; };
    return al(a.A) && al(a.B) && (a.lda % 2 == 0) && (a.ldb % 2 == 0) && (a.sA % 2 == 0) && (a.sB % 2 == 0) &&
           (a.P % 2 == 0) && (a.R % 2 == 0) && (a.Cn % 2 == 0) && a.P >= 2 && a.R >= 2 && a.Cn >= 2;
}

template <typename T, bool PM, int BR, int BC>
static void launch_mma_tiles(hipStream_t st, MmArgs<T> a) {
    a.ntr = ceil_div(a.R, BR);
    a.ntc = ceil_div(a.Cn, BC);
    const long nwg = (long)a.ntr * a.ntc * a.nb;
    const size_t sh = 2 * (MmOperand<T, PM, BR>::ELEMS + MmOperand<T, PM, BC>::ELEMS) * sizeof(T);
    (void)hipFuncSetAttribute((const void*)k_mma<T, PM, BR, BC>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    hipLaunchKernelGGL((k_mma<T, PM, BR, BC>), dim3((unsigned)nwg), dim3(256), sh, st, a);
}

// 64 x 64 tiles; FP64 with the DMA-staged kernel where its copies are aligned
template <typename T>
static void launch_mma(hipStream_t st, bool pmajor, const MmArgs<T>& a) {
    if constexpr (std::is_same<T, double>::value) {
        if (dma_ok(a)) {
            if (pmajor) launch_mma_dma<true>(st, a);
            else launch_mma_dma<false>(st, a);
            return;
        }
    }
    if (pmajor) launch_mma_tiles<T, true, 64, 64>(st, a);
    else launch_mma_tiles<T, false, 64, 64>(st, a);
}

template <typename T>
static MmArgs<T> mm_args(const T* A, int lda, long sA, const T* B, int ldb, long sB, T* C, int ldc, long sC, int R,
                         int Cn, int P, int nb, int j0) {
    MmArgs<T> a{};
    a.A = A;
    a.B = B;
    a.C = C;
    a.sA = sA;
    a.sB = sB;
    a.sC = sC;
    a.lda = lda;
    a.ldb = ldb;
    a.ldc = ldc;
    a.R = R;
    a.Cn = Cn;
    a.P = P;
    a.nb = nb;
    a.zb = (j0 == 0) ? 0 : -1;
    a.dnull = -1;
    return a;
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
// the product takes it transposed from column-major storage). blas: a rocBLAS handle, or null
// for the MFMA kernel.
template <typename T>
int launch_zgemm_fwd(hipStream_t st, void* blas, const Dims& d, const Tables<T>& tb, T* S, T* S2,
                     const double* shift_red, int j0, int nj) {
    const size_t slab = (size_t)d.M * d.N;
    if (!blas) {
        MmArgs<T> a = mm_args<T>(S + slab * j0, d.M, (long)slab, tb.qf, d.N, 0, S2 + slab * j0, d.M, (long)slab, d.M,
                                 d.N, d.N, nj, j0);
        a.shift_red = shift_red;
        a.shiftL = (double)d.L;
        launch_mma<T>(st, false, a);
        return 0;
    }
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
    if (!blas) {
        MmArgs<T> a = mm_args<T>(S2 + slab * j0, d.M, (long)slab, tb.qiT, d.N, 0, S + slab * j0, d.M, (long)slab, d.M,
                                 d.N, d.N, nj, j0);
        a.linesum = linesum;
        a.wsum = tb.wsum;
        a.dn = d.N;
        a.dnull = d.null_m;
        a.dop = 0;
        a.dnorm = (double)d.M;
        launch_mma<T>(st, false, a);
        return 0;
    }
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
    if (!blas) {
        MmArgs<T> a = mm_args<T>(S + slab * j0, d.M, (long)slab, tz.qf + mm * j0, d.M, (long)mm, S2 + slab * j0, d.N,
                                 (long)slab, d.N, d.M, d.M, nj, j0);
        a.shift_red = shift_red;
        a.shiftL = (double)d.L;
        launch_mma<T>(st, true, a);
        return 0;
    }
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
    if (!blas) {
        MmArgs<T> a = mm_args<T>(tz.qiT + mm * j0, d.M, (long)mm, S2 + slab * j0, d.N, (long)slab, S + slab * j0, d.M,
                                 (long)slab, d.M, d.N, d.M, nj, j0);
        a.linesum = linesum;
        a.wsum = tz.wsum;
        a.dn = d.M;
        a.dnull = null_x;
        a.dop = 1;
        a.dnorm = (double)d.N;
        launch_mma<T>(st, false, a);
        return 0;
    }
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
