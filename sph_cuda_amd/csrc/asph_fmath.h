// asph_fmath.h - single-precision exp / log / pow with identical results on the gfx950 device
// and on an x86-64 host with FMA.
//
// Why: the step's perturbation pressure p' = cs^2 (rho - rho_ref(z)) cancels catastrophically
// (rho - rho_ref ~ 1e-4 rho), so a 1-ulp difference in the sounding's exp/pow between two
// libraries shows up as a ~1e-3 relative difference in p' and in the accelerations. Host code
// and device kernels therefore both evaluate the sounding, the Exner function and the
// microphysics with the functions below, which use only IEEE-754 operations that are
// correctly rounded on both targets (+, -, *, /, fma, rint, exact power-of-two scaling).
// Accuracy against the exact functions is tested (<= 2 ulp for exp, <= 3 ulp for log on the
// ranges the physics uses).
//
// Methods: exp by Cody-Waite reduction x = n ln2 + r, |r| <= ln2/2, and a degree-7 Taylor
// polynomial in Horner form; log by x = 2^e m, m in [sqrt(1/2), sqrt(2)), ln m = 2 atanh(s),
// s = (m-1)/(m+1), odd series to s^9; pow(x,y) = exp(y ln x) for x > 0.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP_DEVICE_COMPILE__)
#define ASPH_HD __host__ __device__ __forceinline__
#else
#define ASPH_HD static inline
#endif

namespace asphm {

ASPH_HD uint32_t f2u(float x) { return __builtin_bit_cast(uint32_t, x); }
ASPH_HD float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }
ASPH_HD float fmaf_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
ASPH_HD float rintf_(float x) { return __builtin_rintf(x); }

// 2^k for k in [-126, 127] (exact).
ASPH_HD float pow2i(int k) { return u2f((uint32_t)(k + 127) << 23); }

ASPH_HD float fexp(float x) {
    if (x != x) return x;
    if (x > 88.7228394f) return u2f(0x7f800000u);  // +inf
    if (x < -103.972084f) return 0.0f;
    const float log2e = 1.44269502f;
    const float ln2_hi = 0.693145752f;      // 0x3f317200: 16 significant bits, n*ln2_hi exact
    const float ln2_lo = 1.42860677e-06f;   // ln2 - ln2_hi
    float n = rintf_(x * log2e);
    float r = fmaf_(-n, ln2_hi, x);
    r = fmaf_(-n, ln2_lo, r);
    float p = 1.98412701e-04f;              // 1/7!
    p = fmaf_(p, r, 1.38888892e-03f);       // 1/6!
    p = fmaf_(p, r, 8.33333377e-03f);       // 1/5!
    p = fmaf_(p, r, 4.16666679e-02f);       // 1/4!
    p = fmaf_(p, r, 1.66666672e-01f);       // 1/3!
    p = fmaf_(p, r, 0.5f);
    p = fmaf_(p, r, 1.0f);
    p = fmaf_(p, r, 1.0f);
    int k = (int)n;                          // k in [-150, 128]
    int k1 = k / 2;
    int k2 = k - k1;
    return (p * pow2i(k1)) * pow2i(k2);
}

ASPH_HD float flog(float x) {
    if (x != x) return x;
    if (x < 0.0f) return u2f(0x7fc00000u);  // NaN
    if (x == 0.0f) return u2f(0xff800000u); // -inf
    if (x == u2f(0x7f800000u)) return x;
    int e = 0;
    if (x < 1.17549435e-38f) {               // subnormal: scale into the normal range
        x *= 16777216.0f;
        e = -24;
    }
    uint32_t b = f2u(x);
    e += (int)((b >> 23) & 0xffu) - 127;
    float m = u2f((b & 0x007fffffu) | 0x3f800000u);  // [1, 2)
    if (m > 1.41421354f) {
        m *= 0.5f;
        e += 1;
    }
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float p = 0.111111112f;                  // 1/9
    p = fmaf_(p, s2, 0.142857149f);          // 1/7
    p = fmaf_(p, s2, 0.2f);                  // 1/5
    p = fmaf_(p, s2, 0.333333343f);          // 1/3
    float lm = 2.0f * fmaf_(s * s2, p, s);   // 2 atanh(s) = ln m
    const float ln2_hi = 0.693145752f;
    const float ln2_lo = 1.42860677e-06f;
    float fe = (float)e;
    return fmaf_(fe, ln2_hi, fmaf_(fe, ln2_lo, lm));
}

ASPH_HD float fpow(float x, float y) {
    if (x > 0.0f) return fexp(y * flog(x));
    if (x == 0.0f) return y > 0.0f ? 0.0f : (y == 0.0f ? 1.0f : u2f(0x7f800000u));
    return u2f(0x7fc00000u);
}

}  // namespace asphm
