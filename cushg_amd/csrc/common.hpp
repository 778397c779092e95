// common.hpp -- shared host/device helpers of the SHG library (gfx950).
//
// Complex values are float2 / double2 (interleaved re, im).  The complex
// algebra replaces src/headers/Operators.h:13-192 of the reference.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "../../include/shg.h"

// Floating-point contraction only within one source expression: a*b + c
// written as one expression becomes one FMA in every kernel that inlines it,
// and nothing else is fused.  The backend's context-dependent fusion (the
// default for HIP) would let two kernels running the same arithmetic -- the
// row pass and its wide form, say -- round differently.  Multiply-adds
// spread over helper calls are written with fma() explicitly (axpy).
#pragma clang fp contract(on)

namespace shg {

// src/headers/Crystal.h:19-21 (um, ps, W, V)
constexpr double kPi = 3.14159265358979323846;
constexpr double kC = 299792458.0 * 1e6 / 1e12;
constexpr double kEps0 = 8.8541878128e-12 * 1e12 / 1e6;

template <typename T> struct cpx_t;
template <> struct cpx_t<float> { using type = float2; };
template <> struct cpx_t<double> { using type = double2; };
template <typename T> using cpx = typename cpx_t<T>::type;

// Load / store through the global address space (global_load / global_store
// instead of the flat forms a generic pointer gets).  Global loads complete
// in issue order on the vector-memory counter alone, so a load issued early
// stays in flight while the wait for an earlier one is satisfied; a flat
// access also counts on the LDS counter, whose waits are all-or-nothing.
template <typename T> __device__ __forceinline__ T ld_global(const T* p) {
    if constexpr (std::is_arithmetic_v<T>) {
        return *(const __attribute__((address_space(1))) T*)p;
    } else {  // an 8-byte HIP vector type: one 64-bit load
        static_assert(sizeof(T) == 8, "ld_global: scalar or 8-byte vector");
        return __builtin_bit_cast(T, *(const __attribute__((address_space(1))) unsigned long long*)p);
    }
}
template <typename T> __device__ __forceinline__ void st_global(T* p, T v) {
    *(__attribute__((address_space(1))) T*)p = v;
}

template <typename C> __host__ __device__ __forceinline__ C mk(decltype(C{}.x) a, decltype(C{}.x) b) {
    C c;
    c.x = a;
    c.y = b;
    return c;
}
// Complex product.  (The HIP vector types' own operator* is elementwise;
// +, - and scalar * of those types are the complex ones.)
template <typename C> __host__ __device__ __forceinline__ C cmul(C a, C b) {
    C c;
    c.x = a.x * b.x - a.y * b.y;
    c.y = a.x * b.y + a.y * b.x;
    return c;
}
template <typename C> __host__ __device__ __forceinline__ C conjc(C a) { a.y = -a.y; return a; }
// multiply by +i or -i
template <typename C> __host__ __device__ __forceinline__ C mul_i(C a) { C r; r.x = -a.y; r.y = a.x; return r; }
template <typename C> __host__ __device__ __forceinline__ C mul_mi(C a) { C r; r.x = a.y; r.y = -a.x; return r; }
template <typename C> __host__ __device__ __forceinline__ decltype(C{}.x) abs2(C a) { return a.x * a.x + a.y * a.y; }
// y + a x, one FMA per component
__host__ __device__ __forceinline__ double2 axpy(double a, double2 x, double2 y) {
    return make_double2(fma(a, x.x, y.x), fma(a, x.y, y.y));
}
__host__ __device__ __forceinline__ float2 axpy(float a, float2 x, float2 y) {
    return make_float2(fmaf(a, x.x, y.x), fmaf(a, x.y, y.y));
}

// Extraordinary index n(lambda, T).  MgO:sPPLT per src/headers/Crystal.h:54-70
// (Bruner et al.); MgO:PPLN per Gayer et al., Appl. Phys. B 91, 343 (2008).
template <typename R>
__host__ __device__ __forceinline__ R refr_index(int crystal, R L, R T) {
    const R L2 = L * L;
    if (crystal == SHG_CRYSTAL_MGO_PPLN) {
        const R f = (T - R(24.5)) * (T + R(570.82));
        const R g = R(0.2020) + R(6.113e-8) * f;
        const R n2 = R(5.756) + R(2.860e-6) * f + (R(0.0983) + R(4.700e-8) * f) / (L2 - g * g) +
                     (R(189.32) + R(1.516e-4) * f) / (L2 - R(12.52) * R(12.52)) - R(1.32e-2) * L2;
        return sqrt(n2);
    }
    const R tk = T + R(273.15);
    const R b = R(3.483933e-8) * tk * tk;
    const R c = R(1.607839e-8) * tk * tk;
    const R cc = R(0.185087) + c;
    const R n2 = R(4.502483) + (R(0.007294) + b) / (L2 - cc * cc) +
                 R(0.073423) / (L2 - R(0.199595) * R(0.199595)) +
                 R(0.001) / (L2 - R(7.99724) * R(7.99724)) + R(-0.02357) * L2;
    return sqrt(n2);
}

// dk(T) [1/um] in FP64 -- src/headers/PhaseMatching.h:59,80
__host__ __device__ __forceinline__ double delta_k(int crystal, double lp, double Lambda, double ea,
                                                   double eb, double T) {
    const double ls = 0.5 * lp;
    const double np = refr_index<double>(crystal, lp, T);
    const double ns = refr_index<double>(crystal, ls, T);
    const double d = T - 25.0;
    return 2.0 * kPi * (2.0 * np / lp - ns / ls + 1.0 / (Lambda * (1.0 + ea * d + eb * d * d)));
}

// exp(i ph) for an FP64 phase: ph / (2 pi) is reduced in FP64 to a
// fraction of a turn r in [-1/2, 1/2], rounded once to FP32, and the
// hardware sine and cosine of that fraction (v_sin_f32 / v_cos_f32 take
// their argument in turns) give the components.  Maximum error 1.9e-7 per
// component; the magnitude is low by 3e-8 on average and the phase
// unbiased (checked against FP64 sin/cos over 4.2 million phases up to
// |ph| = 600), which scales a coupling term by 1 - 3e-8: far below the FP32
// fields' own rounding.
__device__ __forceinline__ float2 expi_fast(double ph) {
    const double t = ph * (0.5 / kPi);
    const float r = (float)(t - rint(t));
    return make_float2(__builtin_amdgcn_cosf(r), __builtin_amdgcn_sinf(r));
}

// exp(i ph) in FP64: ph / (2 pi) reduced to the nearest quarter turn n/4
// and a remainder a with |a| <= pi/4, the Taylor series of sin and cos
// through a^15 and a^16 in FP64 (truncation error below 5e-17), rotated by
// n quarter turns.
__device__ __forceinline__ double2 expi_fast64(double ph) {
    const double t = ph * (0.5 / kPi);
    const double n = rint(4.0 * t);
    const double a = (t - 0.25 * n) * (2.0 * kPi);
    const double x2 = a * a;
    double s = 1.0 / 1307674368000.0;  // 1/15!
    s = s * -x2 + 1.0 / 6227020800.0;  // 1/13!
    s = s * -x2 + 1.0 / 39916800.0;
    s = s * -x2 + 1.0 / 362880.0;
    s = s * -x2 + 1.0 / 5040.0;
    s = s * -x2 + 1.0 / 120.0;
    s = s * -x2 + 1.0 / 6.0;
    s = a * (1.0 - x2 * s);
    double c = 1.0 / 20922789888000.0;  // 1/16!
    c = c * -x2 + 1.0 / 87178291200.0;  // 1/14!
    c = c * -x2 + 1.0 / 479001600.0;
    c = c * -x2 + 1.0 / 3628800.0;
    c = c * -x2 + 1.0 / 40320.0;
    c = c * -x2 + 1.0 / 720.0;
    c = c * -x2 + 1.0 / 24.0;
    c = c * -x2 + 0.5;
    c = 1.0 - x2 * c;
    const int q = (int)(n - 4.0 * floor(0.25 * n));
    const double re = (q & 1) ? s : c;
    const double im = (q & 1) ? c : s;
    return make_double2((q == 1 || q == 2) ? -re : re, (q >= 2) ? -im : im);
}

// exp(i t) for a small angle: the Taylor series without range reduction
// when |t| <= 1/8 -- FP64 through t^9 / t^10, FP32 through t^5 / t^6
// (truncation error below 1e-17 and 1e-10) -- else the reduced forms above.
template <typename C> __device__ __forceinline__ C expi_small(double t);
template <> __device__ __forceinline__ double2 expi_small<double2>(double t) {
    if (!(fabs(t) <= 0.125)) return expi_fast64(t);
    const double x2 = t * t;
    double s = 1.0 / 362880.0;
    s = s * -x2 + 1.0 / 5040.0;
    s = s * -x2 + 1.0 / 120.0;
    s = s * -x2 + 1.0 / 6.0;
    s = t * (1.0 - x2 * s);
    double c = 1.0 / 3628800.0;
    c = c * -x2 + 1.0 / 40320.0;
    c = c * -x2 + 1.0 / 720.0;
    c = c * -x2 + 1.0 / 24.0;
    c = c * -x2 + 0.5;
    c = 1.0 - x2 * c;
    return make_double2(c, s);
}
template <> __device__ __forceinline__ float2 expi_small<float2>(double t) {
    if (!(fabs(t) <= 0.125)) return expi_fast(t);
    const float a = (float)t;
    const float x2 = a * a;
    const float s = a * (1.0f - x2 * (1.0f / 6.0f - x2 * (1.0f / 120.0f)));
    const float c = 1.0f - x2 * (0.5f - x2 * (1.0f / 24.0f - x2 * (1.0f / 720.0f)));
    return make_float2(c, s);
}

}  // namespace shg
