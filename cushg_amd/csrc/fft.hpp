// fft.hpp -- register/LDS Stockham FFT for one 1D transform of length N
// spread over TPR = N/8 threads (gfx950, 64-wide wavefronts).
//
// Replaces cufftExecC2C (src/headers/Solver.h:159-183): unnormalised DFT,
// forward sign -1, inverse +1.
//
// Data distribution ("natural"): thread tp (0..TPR-1) holds x[tp + q*TPR]
// in v[q], q = 0..7.  The transform runs as Stockham autosort passes with
// radices {8, M, 8} (M = N/64, omitted when N = 64):
//   pass 1 (span 1, radix 8): registers -> radix-8 DFT -> LDS
//   pass 2 (span 8, radix M): LDS -> twiddle -> radix-M DFT -> LDS
//   pass 3 (span N/8, radix 8): LDS -> twiddle -> radix-8 DFT -> registers
// and leaves the result in the same natural distribution, so a pointwise
// operation and the next transform can follow without data movement.
// Twiddles come from a table tw[m] = exp(-2 pi i m/N) computed in FP64.
// The pass-2 stage always has exactly 64 butterflies.
//
// FP32 accuracy: a constant rounded to FP32 (sqrt(1/2), sqrt(3)/2, a
// twiddle) carries a fixed relative error of up to 3e-8 that every transform
// applies again to the same modes, so over thousands of SSFM slices it
// accumulates coherently (a systematic power drift of ~3e-8 per FFT round
// trip with plain FP32 constants).  Here every product with a constant is
// correctly rounded: the radix constants are split into hi + lo FP32 parts
// applied by one FMA (t*c = fma(t, hi, t*lo)), and twiddles are applied in
// FP64; what remains is data-dependent rounding that does not accumulate
// coherently.
#pragma once

#include "common.hpp"

namespace shg {

template <int N> struct FftShape {
    static_assert(N % 64 == 0 && N >= 64 && N <= 512, "N must be 64*M, M in {1,2,3,4,6,8}");
    static constexpr int TPR = N / 8;             // threads per transform
    static constexpr int M = N / 64;              // middle radix (1: no middle pass)
    static constexpr int LDS = N + N / 8 + 1;     // padded stride of one transform in LDS
};

__device__ __forceinline__ constexpr int fpad(int i) { return i + (i >> 3); }

// Twiddle tables are FP64 for both precisions.  An FP32 point is multiplied
// by its FP64 twiddle in FP64 and rounded once: the product is correctly
// rounded, with no fixed per-twiddle error (an FP32 twiddle, even split
// into hi + lo parts, leaves one: the lo correction is absorbed when the two
// partial products have similar size, e.g. at 30/60 degree angles).  FP64
// FMAs issue at the rate of (non-packed) FP32 ones on gfx950.
template <typename C> struct TwTraits;
template <> struct TwTraits<float2> { using TW = double2; };
template <> struct TwTraits<double2> { using TW = double2; };
template <typename C> using tw_t = typename TwTraits<C>::TW;

template <bool INV> __device__ __forceinline__ float2 twmul(float2 x, const double2* __restrict__ tw, int m) {
    const double2 w = tw[m];
    const double wy = INV ? -w.y : w.y;
    const double xr = x.x, xi = x.y;
    return make_float2((float)(xr * w.x - xi * wy), (float)(xr * wy + xi * w.x));
}
template <bool INV> __device__ __forceinline__ double2 twmul(double2 x, const double2* __restrict__ tw, int m) {
    double2 w = tw[m];
    if (INV) w.y = -w.y;
    return cmul(x, w);
}

// t * c for a constant c = hi + lo (FP32: compensated; FP64: plain)
template <typename R> __device__ __forceinline__ R mulk(R t, double c);
template <> __device__ __forceinline__ float mulk<float>(float t, double c) {
    const float hi = (float)c;
    const float lo = (float)(c - (double)hi);
    return __builtin_fmaf(t, hi, t * lo);
}
template <> __device__ __forceinline__ double mulk<double>(double t, double c) { return t * c; }

template <typename C, bool INV> __device__ __forceinline__ C rot_mi(C a) {  // * (-i) forward, * (+i) inverse
    return INV ? mul_i(a) : mul_mi(a);
}

template <typename C, bool INV> __device__ __forceinline__ void dft2(C* x) {
    C a = x[0];
    x[0] = a + x[1];
    x[1] = a - x[1];
}

template <typename C, bool INV> __device__ __forceinline__ void dft3(C* x) {
    using R = decltype(C{}.x);
    const R c = R(-0.5);
    C t1 = x[1] + x[2];
    C t2 = axpy(c, t1, x[0]);
    const C d = x[1] - x[2];
    C t3 = mk<C>(mulk<R>(d.x, 0.86602540378443864676), mulk<R>(d.y, 0.86602540378443864676));
    // forward: y1 = t2 - i t3, y2 = t2 + i t3
    C it3 = rot_mi<C, INV>(t3);
    x[0] = x[0] + t1;
    x[1] = t2 + it3;
    x[2] = t2 - it3;
}

template <typename C, bool INV> __device__ __forceinline__ void dft4(C* x) {
    C a = x[0] + x[2], b = x[0] - x[2];
    C c = x[1] + x[3], d = rot_mi<C, INV>(x[1] - x[3]);
    x[0] = a + c;
    x[2] = a - c;
    x[1] = b + d;
    x[3] = b - d;
}

template <typename C, bool INV> __device__ __forceinline__ void dft6(C* x) {
    using R = decltype(C{}.x);
    C e[3] = {x[0], x[2], x[4]};
    C o[3] = {x[1], x[3], x[5]};
    dft3<C, INV>(e);
    dft3<C, INV>(o);
    // W6^k forward = exp(-i pi k / 3)
    // w1 = 1/2 -+ i sqrt3/2, w2 = -1/2 -+ i sqrt3/2 (forward sign -)
    const double s3 = INV ? 0.86602540378443864676 : -0.86602540378443864676;
    const R h = R(0.5);
    C o1 = mk<C>(h * o[1].x - mulk<R>(o[1].y, s3), h * o[1].y + mulk<R>(o[1].x, s3));
    C o2 = mk<C>(-h * o[2].x - mulk<R>(o[2].y, s3), -h * o[2].y + mulk<R>(o[2].x, s3));
    x[0] = e[0] + o[0];
    x[3] = e[0] - o[0];
    x[1] = e[1] + o1;
    x[4] = e[1] - o1;
    x[2] = e[2] + o2;
    x[5] = e[2] - o2;
}

template <typename C, bool INV> __device__ __forceinline__ void dft8(C* x) {
    using R = decltype(C{}.x);
    C e[4] = {x[0], x[2], x[4], x[6]};
    C o[4] = {x[1], x[3], x[5], x[7]};
    dft4<C, INV>(e);
    dft4<C, INV>(o);
    const double r = 0.70710678118654752440;
    // W8^1 = (1 - i)/sqrt2 forward, W8^2 = -i, W8^3 = (-1 - i)/sqrt2
    C o1 = mk<C>(mulk<R>(o[1].x + (INV ? -o[1].y : o[1].y), r), mulk<R>(o[1].y + (INV ? o[1].x : -o[1].x), r));
    C o2 = rot_mi<C, INV>(o[2]);
    C o3 = mk<C>(mulk<R>(-o[3].x + (INV ? -o[3].y : o[3].y), r), mulk<R>(-o[3].y + (INV ? o[3].x : -o[3].x), r));
    x[0] = e[0] + o[0];
    x[4] = e[0] - o[0];
    x[1] = e[1] + o1;
    x[5] = e[1] - o1;
    x[2] = e[2] + o2;
    x[6] = e[2] - o2;
    x[3] = e[3] + o3;
    x[7] = e[3] - o3;
}

template <typename C, bool INV, int R> __device__ __forceinline__ void dftR(C* x) {
    if constexpr (R == 2) dft2<C, INV>(x);
    else if constexpr (R == 3) dft3<C, INV>(x);
    else if constexpr (R == 4) dft4<C, INV>(x);
    else if constexpr (R == 6) dft6<C, INV>(x);
    else if constexpr (R == 8) dft8<C, INV>(x);
}

// One length-N transform held in v[8] by thread tp of its group.  `lds`
// points to this transform's private LDS buffer (FftShape<N>::LDS entries).
// Every thread of the workgroup must call it the same number of times
// (it contains workgroup barriers).
// NF independent transforms (e.g. the pump and SH rows) sharing the same
// barriers: v[f] is held by thread tp, buffer f at lds + f*ld.  A thread
// with act = false only takes part in the barriers (it neither reads nor
// writes `lds` and leaves v unchanged).
template <typename C, int N, bool INV, int NF>
__device__ __forceinline__ void fft_multi(C (&v)[NF][8], C* lds, int ld, const tw_t<C>* __restrict__ tw, int tp,
                                          bool act = true) {
    constexpr int TPR = FftShape<N>::TPR;
    constexpr int M = FftShape<N>::M;
    // pass 1: span 1, radix 8, butterfly t = tp reads x[tp + j*TPR] = v[j]
    if (act) {
#pragma unroll
        for (int f = 0; f < NF; f++) {
            dft8<C, INV>(v[f]);
#pragma unroll
            for (int q = 0; q < 8; q++) lds[f * ld + fpad(8 * tp + q)] = v[f][q];
        }
    }
    __syncthreads();
    if constexpr (M > 1) {
        // pass 2: span 8, radix M, 64 butterflies, twiddle W_{8M}^{jk} = W_N^{8jk}
        constexpr int PER = (64 + TPR - 1) / TPR;
        C xs[NF][PER][M];
#pragma unroll
        for (int f = 0; f < NF; f++)
#pragma unroll
            for (int i = 0; i < PER; i++) {
                const int b = tp + i * TPR;
                if (act && (PER * TPR == 64 || b < 64)) {
                    const int k = b & 7;
#pragma unroll
                    for (int j = 0; j < M; j++) {
                        C x = lds[f * ld + fpad(b + 64 * j)];
                        if (j > 0) x = twmul<INV>(x, tw, 8 * j * k);
                        xs[f][i][j] = x;
                    }
                }
            }
        __syncthreads();
#pragma unroll
        for (int f = 0; f < NF; f++)
#pragma unroll
            for (int i = 0; i < PER; i++) {
                const int b = tp + i * TPR;
                if (act && (PER * TPR == 64 || b < 64)) {
                    const int k = b & 7;
                    dftR<C, INV, M>(xs[f][i]);
                    const int base = (b - k) * M + k;
#pragma unroll
                    for (int q = 0; q < M; q++) lds[f * ld + fpad(base + 8 * q)] = xs[f][i][q];
                }
            }
        __syncthreads();
    }
    // pass 3: span N/8, radix 8, butterfly t = tp, twiddle W_N^{j tp}
    if (act) {
#pragma unroll
        for (int f = 0; f < NF; f++) {
#pragma unroll
            for (int j = 0; j < 8; j++) {
                C x = lds[f * ld + fpad(tp + j * TPR)];
                if (j > 0) x = twmul<INV>(x, tw, j * tp);
                v[f][j] = x;
            }
            dft8<C, INV>(v[f]);
        }
    }
    __syncthreads();  // the buffers may be reused by the next call
}

template <typename C, int N, bool INV>
__device__ __forceinline__ void fft_dist(C (&v)[8], C* lds, const tw_t<C>* __restrict__ tw, int tp,
                                         bool act = true) {
    fft_multi<C, N, INV, 1>(reinterpret_cast<C(&)[1][8]>(v), lds, 0, tw, tp, act);
}

// Twiddles W^(j m), j = 1, 2, ..., as FP64 powers of the table entry
// W^m = tw[m] (one table read per pass instead of one per point; each power
// carries a relative error of about j FP64 ulps, far below the FP32 and FP64
// rounding of the transform itself).  mul(x) multiplies by the next power.
template <bool INV> struct TwPow {
    double2 w, p;
    __device__ __forceinline__ TwPow(const double2* __restrict__ tw, int m) {
        w = tw[m];
        if (INV) w.y = -w.y;
        p = w;
    }
    __device__ __forceinline__ float2 mul(float2 x) {
        const double xr = x.x, xi = x.y;
        const float2 r = make_float2((float)(xr * p.x - xi * p.y), (float)(xr * p.y + xi * p.x));
        p = cmul(p, w);
        return r;
    }
    __device__ __forceinline__ double2 mul(double2 x) {
        const double2 r = cmul(x, p);
        p = cmul(p, w);
        return r;
    }
};

// One transform (either direction) in the arithmetic of fft_multi (radices
// {8, M, 8}, the same butterflies and products) with fewer barriers and
// twiddle reads:
//  * the middle pass writes each butterfly's outputs back into the slots it
//    read (no barrier between its reads and writes), and the last pass reads
//    them from there;
//  * twiddles are FP64 powers of one table entry per pass (TwPow).
// With tp = 8u + w, the physical slots (before padding) are: pass 1 writes
// 8 tp + q; middle butterfly b reads/writes b + 64 j; pass 3 reads
// 8 j + w + 64 u (a relabelling of fft_multi's Stockham slots: the values
// and the arithmetic are the same).  Two barriers, and a third at the end
// when END (the buffer may then be reused at once); without it, a caller
// that writes `lds` afterwards must synchronise first, unless each thread
// writes only the slots it read in pass 3 (as fft_fwd_op_inv does).  A
// thread with act = false only takes part in the barriers.
template <typename C, int N, bool INV, bool END>
__device__ __forceinline__ void fft_pow(C (&v)[8], C* lds, const double2* __restrict__ tw, int tp, bool act = true) {
    constexpr int TPR = FftShape<N>::TPR;
    constexpr int M = FftShape<N>::M;
    constexpr int PER = (64 + TPR - 1) / TPR;
    const int w = tp & 7, u = tp >> 3;
    if (act) {
        dft8<C, INV>(v);
#pragma unroll
        for (int q = 0; q < 8; q++) lds[fpad(8 * tp + q)] = v[q];
    }
    __syncthreads();
    if constexpr (M > 1) {
        if (act) {
#pragma unroll
            for (int i = 0; i < PER; i++) {
                const int b = tp + i * TPR;
                if (PER * TPR == 64 || b < 64) {
                    C xs[M];
                    TwPow<INV> t(tw, 8 * (b & 7));
#pragma unroll
                    for (int j = 0; j < M; j++) {
                        const C a = lds[fpad(b + 64 * j)];
                        xs[j] = j ? t.mul(a) : a;
                    }
                    dftR<C, INV, M>(xs);
#pragma unroll
                    for (int q = 0; q < M; q++) lds[fpad(b + 64 * q)] = xs[q];
                }
            }
        }
        __syncthreads();
    }
    if (act) {
        TwPow<INV> t(tw, tp);
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const C a = lds[fpad(8 * j + w + 64 * u)];
            v[j] = j ? t.mul(a) : a;
        }
        dft8<C, INV>(v);
    }
    if (END) __syncthreads();
}

// Forward transform (fft_pow), a pointwise operation op(v) on the
// natural distribution, inverse transform: the sequence of the column pass.
// The inverse transform's first pass writes each thread's outputs into the
// slots that thread read in the forward transform's last pass, so no
// barrier separates the two transforms.  Inverse slots (tp = 8u + w):
// pass 1 writes 8 q + w + 64 u; middle butterfly b reads/writes
// 8 (b & 7) + (b >> 3) + 64 j; pass 3 reads 8 w + j + 64 u.  Four barriers
// for the pair; none at the end: a caller that touches `lds` afterwards
// must synchronise first.
template <typename C, int N, typename Op>
__device__ __forceinline__ void fft_fwd_op_inv(C (&v)[8], C* lds, const double2* __restrict__ tw, int tp, Op op) {
    constexpr int TPR = FftShape<N>::TPR;
    constexpr int M = FftShape<N>::M;
    constexpr int PER = (64 + TPR - 1) / TPR;
    const int w = tp & 7, u = tp >> 3;
    fft_pow<C, N, false, false>(v, lds, tw, tp);
    op(v);
    dft8<C, true>(v);
#pragma unroll
    for (int q = 0; q < 8; q++) lds[fpad(8 * q + w + 64 * u)] = v[q];
    __syncthreads();
    if constexpr (M > 1) {
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int b = tp + i * TPR;
            if (PER * TPR == 64 || b < 64) {
                const int s0 = 8 * (b & 7) + (b >> 3);
                C xs[M];
                TwPow<true> t(tw, 8 * (b & 7));
#pragma unroll
                for (int j = 0; j < M; j++) {
                    const C a = lds[fpad(s0 + 64 * j)];
                    xs[j] = j ? t.mul(a) : a;
                }
                dftR<C, true, M>(xs);
#pragma unroll
                for (int q = 0; q < M; q++) lds[fpad(s0 + 64 * q)] = xs[q];
            }
        }
        __syncthreads();
    }
    TwPow<true> t(tw, tp);
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const C a = lds[fpad(8 * w + j + 64 * u)];
        v[j] = j ? t.mul(a) : a;
    }
    dft8<C, true>(v);
}

}  // namespace shg
