// dev_common.hpp -- device helpers shared by the gfx950 kernels: complex
// arithmetic, precision-dispatched math, the Philox-4x32-10 counter RNG and
// sub-wavefront group reductions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rbmgpu {

template <typename R>
struct cplx {
    R re, im;
};

template <typename R>
__host__ __device__ inline cplx<R> mk(R a, R b) {
    return cplx<R>{a, b};
}
template <typename R>
__host__ __device__ inline cplx<R> operator+(cplx<R> a, cplx<R> b) {
    return {a.re + b.re, a.im + b.im};
}
template <typename R>
__host__ __device__ inline cplx<R> operator-(cplx<R> a, cplx<R> b) {
    return {a.re - b.re, a.im - b.im};
}
template <typename R>
__host__ __device__ inline cplx<R> operator*(cplx<R> a, cplx<R> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename R>
__host__ __device__ inline cplx<R> operator*(R s, cplx<R> a) {
    return {s * a.re, s * a.im};
}
template <typename R>
__host__ __device__ inline cplx<R> conj(cplx<R> a) {
    return {a.re, -a.im};
}
template <typename R>
__host__ __device__ inline R norm2(cplx<R> a) {
    return a.re * a.re + a.im * a.im;
}
template <typename R, typename S>
__host__ __device__ inline cplx<R> cvt(cplx<S> a) {
    return {(R)a.re, (R)a.im};
}

// ------------------------------------------------------------------ math
template <typename R>
struct M;
template <>
struct M<double> {
    static __device__ double exp(double x) { return ::exp(x); }
    static __device__ double log(double x) { return ::log(x); }
    static __device__ double log2(double x) { return ::log2(x); }
    static __device__ double exp2(double x) { return ::exp2(x); }
    static __device__ double log1p(double x) { return ::log1p(x); }
    static __device__ double expm1(double x) { return ::expm1(x); }
    static __device__ double atan2(double y, double x) { return ::atan2(y, x); }
    static __device__ void sincos(double x, double* s, double* c) { ::sincos(x, s, c); }
    static __device__ double fabs(double x) { return ::fabs(x); }
    static __device__ double copysign(double a, double b) { return ::copysign(a, b); }
    static __device__ double frexp(double x, int* e) { return ::frexp(x, e); }
    static __device__ double ldexp(double x, int e) { return ::ldexp(x, e); }
};
template <>
struct M<float> {
    static __device__ float exp(float x) { return __expf(x); }
    static __device__ float log(float x) { return __logf(x); }
    // v_log_f32 / v_exp_f32 (base 2, about 1 ulp)
    static __device__ float log2(float x) { return __builtin_amdgcn_logf(x); }
    static __device__ float exp2(float x) { return __builtin_amdgcn_exp2f(x); }
    static __device__ float log1p(float x) { return ::log1pf(x); }
    static __device__ float expm1(float x) { return ::expm1f(x); }
    static __device__ float atan2(float y, float x) { return ::atan2f(y, x); }
    static __device__ void sincos(float x, float* s, float* c) { ::sincosf(x, s, c); }
    static __device__ float fabs(float x) { return ::fabsf(x); }
    static __device__ float copysign(float a, float b) { return ::copysignf(a, b); }
    static __device__ float frexp(float x, int* e) { return ::frexpf(x, e); }
    static __device__ float ldexp(float x, int e) { return ::ldexpf(x, e); }
};

// tanh(x + iy) = (sinh 2x + i sin 2y) / (cosh 2x + cos 2y); dividing by
// 2 cosh^2 x gives (th + i s c sech^2 x) / (c^2 + th^2 s^2) with th = tanh x,
// s = sin y, c = cos y: the denominator is a sum of non-negative terms (no
// cancellation near the poles x = 0, y = pi/2 + k pi) and nothing overflows.
// th and sech^2 come from em = expm1(-2|x|): th = -em / (2 + em),
// sech^2 = 4 (1 + em) / (2 + em)^2.
//
// FP32 evaluates the same formula with cheaper pieces: expm1 by its Taylor
// polynomial for |2x| < 1/4 (truncation < 5e-8 relative) and exp - 1
// beyond, sin/cos by sincos_f below, reciprocals by v_rcp_f32 (1 ulp).
//
// sincos_f: y = n pi/2 + x with n = rint(2y/pi) and |x| <= pi/4 (two-term
// Cody-Waite split of pi/2), then the Taylor series of sin to x^9 and of
// cos to x^10 (truncation < 2e-9 on |x| <= pi/4) and the quadrant swap.
__device__ inline void sincos_f(float y, float& s, float& c) {
    float n = __builtin_rintf(y * 0.63661977236758134f);
    float x = __builtin_fmaf(-n, 1.57079637050628662f, y);  // pi/2 rounded to float
    x = __builtin_fmaf(n, 4.371139006309477e-8f, x);      // (that float) - pi/2
    float x2 = x * x;
    float sp = x * (1.f + x2 * (-1.f / 6 + x2 * (1.f / 120 + x2 * (-1.f / 5040 + x2 * (1.f / 362880)))));
    float cp = 1.f + x2 * (-0.5f + x2 * (1.f / 24 + x2 * (-1.f / 720 + x2 * (1.f / 40320 + x2 * (-1.f / 3628800)))));
    int q = (int)n & 3;
    float ss = (q & 1) ? cp : sp, cc = (q & 1) ? sp : cp;
    s = (q & 2) ? -ss : ss;
    c = ((q + 1) & 2) ? -cc : cc;
}

__device__ inline cplx<float> ctanh_fast(cplx<float> z) {
    float ax = __builtin_fabsf(z.re), t = -2.f * ax;
    float ps = t * (1.f + t * (0.5f + t * (1.f / 6 + t * (1.f / 24 + t * (1.f / 120 + t * (1.f / 720))))));
    float em = ax < 0.125f ? ps : __expf(t) - 1.f;
    float q = __builtin_amdgcn_rcpf(2.f + em);
    float th = __builtin_copysignf(-em * q, z.re);
    float sech2 = 4.f * (1.f + em) * q * q;
    float s, c;
    sincos_f(z.im, s, c);
    float inv = __builtin_amdgcn_rcpf(c * c + th * th * s * s);
    return {th * inv, s * c * sech2 * inv};
}

template <typename R>
__device__ inline cplx<R> ctanh(cplx<R> z) {
    if constexpr (sizeof(R) == 4) return ctanh_fast(z);
    R em = M<R>::expm1(-2 * M<R>::fabs(z.re));
    R q = 1 / (2 + em);
    R th = M<R>::copysign(-em * q, z.re);
    R sech2 = 4 * (1 + em) * q * q;
    R s, c;
    M<R>::sincos(z.im, &s, &c);
    R inv = 1 / (c * c + th * th * s * s);
    return {th * inv, s * c * sech2 * inv};
}

// ln cosh(x + iy) = ln cosh x + ln(cos y + i tanh(x) sin y), with
// ln cosh x = |x| - ln 2 + log1p(exp(-2|x|)); imaginary part is the
// principal argument.
template <typename R>
__device__ inline cplx<R> clncosh(cplx<R> z) {
    R a = M<R>::fabs(z.re);
    R em = M<R>::expm1(-2 * a);  // E - 1 without cancellation
    R th = M<R>::copysign(-em / (2 + em), z.re);
    R s, c;
    M<R>::sincos(z.im, &s, &c);
    s *= th;
    R re = a - (R)0.69314718055994530942 + M<R>::log1p(1 + em) + (R)0.5 * M<R>::log(s * s + c * c);
    return {re, M<R>::atan2(s, c)};
}

// complex exp
template <typename R>
__device__ inline cplx<R> cexp(cplx<R> z) {
    R e = M<R>::exp(z.re), s, c;
    M<R>::sincos(z.im, &s, &c);
    return {e * c, e * s};
}

// scale p to a mantissa with max(|re|,|im|) in [0.5, 1) and an exponent
template <typename R>
__device__ inline void renorm(cplx<R>& p, int& e) {
    R m = fmax(M<R>::fabs(p.re), M<R>::fabs(p.im));
    if (m == 0 || !(m < (R)INFINITY)) return;
    int k;
    M<R>::frexp(m, &k);
    p.re = M<R>::ldexp(p.re, -k);
    p.im = M<R>::ldexp(p.im, -k);
    e += k;
}

// ------------------------------------------------------------------- RNG
// Philox-4x32-10 (Salmon et al., SC'11) with the published constants.
struct u4 {
    uint32_t x, y, z, w;
};
__device__ inline u4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                            uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return {c0, c1, c2, c3};
}
// uniform in (0,1): 53 bits for double, 24 bits for float
template <typename R>
__device__ inline R uniform(uint32_t a, uint32_t b);
template <>
__device__ inline double uniform<double>(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6) + 0.5) * (1.0 / 9007199254740992.0);
}
template <>
__device__ inline float uniform<float>(uint32_t a, uint32_t) {
    return ((float)(a >> 8) + 0.5f) * (1.0f / 16777216.0f);
}
__device__ inline uint32_t bounded(uint32_t r, uint32_t n) {
    return (uint32_t)(((uint64_t)r * n) >> 32);
}

// ------------------------------------------------------ group reductions
// Butterfly reductions over aligned groups of L lanes.  Step O pairs every
// lane with a partner in the other half of its aligned 2O-lane block:
// O = 1, 2 by DPP quad permutes (xor 1, xor 2), O = 4, 8 by DPP row
// half-mirror / mirror (lane i <-> 7-i, 15-i), O = 16 by ds_swizzle xor 16,
// O = 32 by ds_bpermute.  After steps 1..O every lane of a 2O block holds
// the same value, so sums and products come out bit-identical in every lane
// of the group (a + b == b + a) and control flow that depends on them stays
// uniform within the group.
template <int O>
__device__ inline int partner(int v) {
    if constexpr (O == 1) return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);
    else if constexpr (O == 2) return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);
    else if constexpr (O == 4) return __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false);
    else if constexpr (O == 8) return __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, false);
    else if constexpr (O == 16) return __builtin_amdgcn_ds_swizzle(v, 0x401F);
    else return __shfl_xor(v, 32);
}
template <int O>
__device__ inline float partner(float v) {
    return __int_as_float(partner<O>(__float_as_int(v)));
}
template <int O>
__device__ inline double partner(double v) {
    int2 h = *reinterpret_cast<int2*>(&v);
    h.x = partner<O>(h.x);
    h.y = partner<O>(h.y);
    return *reinterpret_cast<double*>(&h);
}

template <int L, int O = 1, typename T>
__device__ inline T group_sum(T v) {
    if constexpr (O < L) {
        v += partner<O>(v);
        return group_sum<L, 2 * O>(v);
    } else {
        return v;
    }
}
template <int L, int O = 1>
__device__ inline float group_max(float v) {
    if constexpr (O < L) {
        v = fmaxf(v, partner<O>(v));
        return group_max<L, 2 * O>(v);
    } else {
        return v;
    }
}
template <int L, int O = 1, typename R>
__device__ inline cplx<R> group_csum(cplx<R> v) {
    if constexpr (O < L) {
        v.re += partner<O>(v.re);
        v.im += partner<O>(v.im);
        return group_csum<L, 2 * O>(v);
    } else {
        return v;
    }
}
// product of complex numbers over the group with exponent tracking; the
// result is returned as mantissa m and exponent e (value m * 2^e).  Each
// lane's mantissa is normalised once (max(|re|, |im|) in [0.5, 1), so |m|
// in [0.5, sqrt 2)); the log2(L) <= 6 butterfly products then stay within
// [2^-64, 2^32] and need no further scaling.  With FMA contraction the
// lanes of a group may differ in the last bit of the result, so callers use
// it as data only (never for control flow).
template <int L, int O = 1, typename R>
__device__ inline void group_cprod_steps(cplx<R>& m, int& e) {
    if constexpr (O < L) {
        cplx<R> q{partner<O>(m.re), partner<O>(m.im)};
        int qe = partner<O>(e);
        m = m * q;
        e += qe;
        group_cprod_steps<L, 2 * O>(m, e);
    }
}
template <int L, typename R>
__device__ inline void group_cprod(cplx<R>& m, int& e) {
    renorm(m, e);
    group_cprod_steps<L>(m, e);
}

}  // namespace rbmgpu
