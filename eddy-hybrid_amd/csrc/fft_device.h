// fft_device.h -- batched complex FFTs in LDS (Stockham autosort, radix 2/3/4/5/8) and the
// real-FFT packing used by the y-transform (fftpack convention, src/fftpack.f:140-240) and
// the x2 DCT.
//
// A workgroup holds `P` independent length-n complex sequences in ONE LDS buffer (split re/im
// arrays, row stride `ld`) and transforms them in place: each radix pass loads every
// butterfly into registers, synchronises, then stores. Twiddles come from a table
// tw[t] = exp(-2 pi i t / n), t < n (conjugated for the inverse).
#pragma once

#include "common.h"

namespace eddy {

// LDS rows of the FFT buffers are padded by one element per 8 (index x -> x + x/8): the
// Stockham writes stride by the radix and by 8*R, which would otherwise land many lanes of a
// wave on one LDS bank. A row holding indices 0..n needs fft_ld(n) elements.
__device__ __forceinline__ int pidx(int x) { return x + (x >> 3); }
__host__ __device__ inline int fft_ld(int n) { return n + (n >> 3) + 2; }

template <typename T>
struct Cx {
    T re, im;
};

template <typename T>
__device__ __forceinline__ Cx<T> cmul(Cx<T> a, Cx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// multiply by dir*i  (dir = -1 forward, +1 inverse)
template <typename T>
__device__ __forceinline__ Cx<T> mul_di(Cx<T> a, T dir) {
    return {-dir * a.im, dir * a.re};
}

template <typename T>
__device__ __forceinline__ void dft2(Cx<T>* v) {
    Cx<T> a = v[0], b = v[1];
    v[0] = {a.re + b.re, a.im + b.im};
    v[1] = {a.re - b.re, a.im - b.im};
}

template <typename T>
__device__ __forceinline__ void dft4(Cx<T>* v, T dir) {
    Cx<T> s02 = {v[0].re + v[2].re, v[0].im + v[2].im};
    Cx<T> d02 = {v[0].re - v[2].re, v[0].im - v[2].im};
    Cx<T> s13 = {v[1].re + v[3].re, v[1].im + v[3].im};
    Cx<T> d13 = mul_di(Cx<T>{v[1].re - v[3].re, v[1].im - v[3].im}, dir);
    v[0] = {s02.re + s13.re, s02.im + s13.im};
    v[2] = {s02.re - s13.re, s02.im - s13.im};
    v[1] = {d02.re + d13.re, d02.im + d13.im};
    v[3] = {d02.re - d13.re, d02.im - d13.im};
}

template <typename T>
__device__ __forceinline__ void dft3(Cx<T>* v, T dir) {
    const T h = T(0.86602540378443864676);  // sin(2 pi/3)
    Cx<T> t = {v[1].re + v[2].re, v[1].im + v[2].im};
    Cx<T> u = {v[1].re - v[2].re, v[1].im - v[2].im};
    Cx<T> a = {v[0].re - T(0.5) * t.re, v[0].im - T(0.5) * t.im};
    Cx<T> b = mul_di(Cx<T>{h * u.re, h * u.im}, dir);
    v[0] = {v[0].re + t.re, v[0].im + t.im};
    v[1] = {a.re + b.re, a.im + b.im};
    v[2] = {a.re - b.re, a.im - b.im};
}

template <typename T>
__device__ __forceinline__ void dft5(Cx<T>* v, T dir) {
    const T c1 = T(0.30901699437494742410);   // cos(2 pi/5)
    const T c2 = T(-0.80901699437494742410);  // cos(4 pi/5)
    const T s1 = T(0.95105651629515357212);   // sin(2 pi/5)
    const T s2 = T(0.58778525229247312917);   // sin(4 pi/5)
    Cx<T> t1 = {v[1].re + v[4].re, v[1].im + v[4].im};
    Cx<T> t2 = {v[2].re + v[3].re, v[2].im + v[3].im};
    Cx<T> u1 = {v[1].re - v[4].re, v[1].im - v[4].im};
    Cx<T> u2 = {v[2].re - v[3].re, v[2].im - v[3].im};
    Cx<T> a1 = {v[0].re + c1 * t1.re + c2 * t2.re, v[0].im + c1 * t1.im + c2 * t2.im};
    Cx<T> a2 = {v[0].re + c2 * t1.re + c1 * t2.re, v[0].im + c2 * t1.im + c1 * t2.im};
    Cx<T> b1 = mul_di(Cx<T>{s1 * u1.re + s2 * u2.re, s1 * u1.im + s2 * u2.im}, dir);
    Cx<T> b2 = mul_di(Cx<T>{s2 * u1.re - s1 * u2.re, s2 * u1.im - s1 * u2.im}, dir);
    v[0] = {v[0].re + t1.re + t2.re, v[0].im + t1.im + t2.im};
    v[1] = {a1.re + b1.re, a1.im + b1.im};
    v[4] = {a1.re - b1.re, a1.im - b1.im};
    v[2] = {a2.re + b2.re, a2.im + b2.im};
    v[3] = {a2.re - b2.re, a2.im - b2.im};
}

template <typename T>
__device__ __forceinline__ void dft8(Cx<T>* v, T dir) {
    const T r2 = T(0.70710678118654752440);
    Cx<T> e[4] = {v[0], v[2], v[4], v[6]};
    Cx<T> o[4] = {v[1], v[3], v[5], v[7]};
    dft4(e, dir);
    dft4(o, dir);
    // w8^k = exp(dir * 2 pi i k / 8)
    Cx<T> w1 = {r2, dir * r2};
    Cx<T> w3 = {-r2, dir * r2};
    Cx<T> o1 = cmul(o[1], w1);
    Cx<T> o2 = mul_di(o[2], dir);
    Cx<T> o3 = cmul(o[3], w3);
    v[0] = {e[0].re + o[0].re, e[0].im + o[0].im};
    v[4] = {e[0].re - o[0].re, e[0].im - o[0].im};
    v[1] = {e[1].re + o1.re, e[1].im + o1.im};
    v[5] = {e[1].re - o1.re, e[1].im - o1.im};
    v[2] = {e[2].re + o2.re, e[2].im + o2.im};
    v[6] = {e[2].re - o2.re, e[2].im - o2.im};
    v[3] = {e[3].re + o3.re, e[3].im + o3.im};
    v[7] = {e[3].re - o3.re, e[3].im - o3.im};
}

// Exact quotient and remainder of small non-negative integers (x < 2^22, d >= 1) through a
// float reciprocal rd = 1/d and one correction step: a few VALU ops instead of the ~25 of a
// generic 32-bit division.
__device__ __forceinline__ int udiv_small(int x, int d, float rd, int& rem) {
    int q = (int)((float)x * rd);
    int r = x - q * d;
    if (r < 0) {
        --q;
        r += d;
    } else if (r >= d) {
        ++q;
        r -= d;
    }
    rem = r;
    return q;
}

// In-place complex FFT of P sequences of length fp.n held in ONE buffer (split re/im, row
// stride ld). Each radix pass: every thread loads at most one butterfly (up to 8 points, the
// radix chosen at run time with predicated loads/stores so that all radices share the same
// registers), the workgroup synchronises, then every thread stores. Needs
// P * n / (smallest radix of the plan) <= nthr. The caller synchronises before (data
// ready); the result is visible on return.
template <typename T>
__device__ void cfft_lds_ip(T* __restrict__ xr, T* __restrict__ xi, int ld, int P, const FftPlan& fp,
                            const T* twr, const T* twi, T dir, int tid, int nthr) {
    const int n = fp.n;
    int Ns = 1;
    for (int p = 0; p < fp.count; ++p) {
        const int R = fp.rad[p];
        const int nb = n / R;
        const int tstep = n / (Ns * R);
        const bool active = tid < P * nb;
        Cx<T> v[8];
        int orow = 0, obase = 0;
        if (active) {
            int jj, k;
            const int pen = udiv_small(tid, nb, 1.0f / (float)nb, jj);
            const int blk = udiv_small(jj, Ns, 1.0f / (float)Ns, k);
            const T* pr = xr + pen * ld;
            const T* pi = xi + pen * ld;
#pragma unroll
            for (int r = 0; r < 8; ++r)
                if (r < R) v[r] = {pr[pidx(jj + r * nb)], pi[pidx(jj + r * nb)]};
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < 8; ++r) {
                    if (r < R) {
                        const int t = k * r * tstep;
                        // table holds exp(-2 pi i t/n); forward (dir=-1) uses it as is, inverse conj
                        Cx<T> w = {twr[t], (dir < 0) ? twi[t] : -twi[t]};
                        v[r] = cmul(v[r], w);
                    }
                }
            }
            switch (R) {
                case 2: dft2(v); break;
                case 3: dft3(v, dir); break;
                case 4: dft4(v, dir); break;
                case 5: dft5(v, dir); break;
                default: dft8(v, dir); break;
            }
            orow = pen;
            obase = blk * Ns * R + k;
        }
        __syncthreads();
        if (active) {
            T* qr = xr + orow * ld;
            T* qi = xi + orow * ld;
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                if (r < R) {
                    qr[pidx(obase + r * Ns)] = v[r].re;
                    qi[pidx(obase + r * Ns)] = v[r].im;
                }
            }
        }
        __syncthreads();
        Ns *= R;
    }
}

// ------------------------------------------------------------------ compile-time radix plans
// The same in-place passes with the length and the radices fixed at compile time, for the
// power-of-two plans radix_plan makes with radices up to 8 (n = 16 ... 512). With n, R and the
// span Ns constants, the butterfly's pencil, block and twiddle indices are shifts and masks,
// and every LDS address of a butterfly is one computed base plus constant offsets: for
// power-of-two strides the padded index pidx(base + r*s) equals pidx(base) + r*s + (r*s >> 3)
// whenever base is aligned as the pass aligns it (the loads: base = jj < s or s % 8 == 0; the
// stores: base = blk*Ns*R + k with k < Ns), so the offsets fold into the LDS instructions.
__device__ __forceinline__ constexpr int poff(int x) { return x + (x >> 3); }

template <typename T, int n, int R, int Ns>
__device__ __forceinline__ void cfft_pass_static(T* __restrict__ xr, T* __restrict__ xi, int ld, int P,
                                                 const T* twr, const T* twi, T dir, int tid) {
    constexpr int nb = n / R;
    constexpr int tstep = n / (Ns * R);
    const bool active = tid < P * nb;
    Cx<T> v[R];
    int oaddr = 0;
    if (active) {
        const int pen = tid / nb;  // powers of two: shifts and masks
        const int jj = tid - pen * nb;
        const int blk = jj / Ns;
        const int k = jj - blk * Ns;
        const int ia = pen * ld + pidx(jj);
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = {xr[ia + poff(r * nb)], xi[ia + poff(r * nb)]};
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                const int t = k * (r * tstep);
                // table holds exp(-2 pi i t/n); forward (dir=-1) uses it as is, inverse conj
                Cx<T> w = {twr[t], (dir < 0) ? twi[t] : -twi[t]};
                v[r] = cmul(v[r], w);
            }
        }
        if constexpr (R == 2) dft2(v);
        else if constexpr (R == 4) dft4(v, dir);
        else dft8(v, dir);
        oaddr = pen * ld + pidx(blk * Ns * R + k);
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            xr[oaddr + poff(r * Ns)] = v[r].re;
            xi[oaddr + poff(r * Ns)] = v[r].im;
        }
    }
    __syncthreads();
}

template <typename T, int n, int Ns, int R, int... Rest>
__device__ __forceinline__ void cfft_passes_static(T* xr, T* xi, int ld, int P, const T* twr, const T* twi, T dir,
                                                   int tid) {
    cfft_pass_static<T, n, R, Ns>(xr, xi, ld, P, twr, twi, dir, tid);
    if constexpr (sizeof...(Rest) > 0) cfft_passes_static<T, n, Ns * R, Rest...>(xr, xi, ld, P, twr, twi, dir, tid);
}

template <int... Rs>
__device__ __forceinline__ bool plan_is(const FftPlan& fp) {
    constexpr int rs[] = {Rs...};
    if (fp.count != (int)sizeof...(Rs)) return false;
#pragma unroll
    for (int p = 0; p < (int)sizeof...(Rs); ++p)
        if (fp.rad[p] != rs[p]) return false;
    return true;
}

// cfft_lds_ip with the power-of-two plans of radix_plan (largest radix 8) taken by the
// compile-time passes; every other plan by the run-time passes. Same contract.
template <typename T>
__device__ __forceinline__ void cfft_lds(T* __restrict__ xr, T* __restrict__ xi, int ld, int P, const FftPlan& fp,
                                         const T* twr, const T* twi, T dir, int tid, int nthr) {
    if (fp.n == 128 && plan_is<8, 4, 4>(fp)) cfft_passes_static<T, 128, 1, 8, 4, 4>(xr, xi, ld, P, twr, twi, dir, tid);
    else if (fp.n == 64 && plan_is<8, 8>(fp)) cfft_passes_static<T, 64, 1, 8, 8>(xr, xi, ld, P, twr, twi, dir, tid);
    else if (fp.n == 256 && plan_is<8, 8, 4>(fp))
        cfft_passes_static<T, 256, 1, 8, 8, 4>(xr, xi, ld, P, twr, twi, dir, tid);
    else if (fp.n == 512 && plan_is<8, 8, 8>(fp))
        cfft_passes_static<T, 512, 1, 8, 8, 8>(xr, xi, ld, P, twr, twi, dir, tid);
    else if (fp.n == 32 && plan_is<8, 4>(fp)) cfft_passes_static<T, 32, 1, 8, 4>(xr, xi, ld, P, twr, twi, dir, tid);
    else if (fp.n == 16 && plan_is<4, 4>(fp)) cfft_passes_static<T, 16, 1, 4, 4>(xr, xi, ld, P, twr, twi, dir, tid);
    else cfft_lds_ip<T>(xr, xi, ld, P, fp, twr, twi, dir, tid, nthr);
}

// Smallest radix of a plan (bounds the butterflies per pass of cfft_lds_ip).
__host__ __device__ inline int min_radix(const FftPlan& fp) {
    int r = 8;
    for (int p = 0; p < fp.count; ++p) r = fp.rad[p] < r ? fp.rad[p] : r;
    return fp.count ? r : 8;
}

// Workgroup size of the transform kernels: 512 threads hold 16 pencils of a length-256 real
// transform at one butterfly per thread and pass, so every HBM row access of a tile covers a
// full 128-byte line.
constexpr int kFftThreads = 512;

}  // namespace eddy
