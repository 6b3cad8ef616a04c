// ms_math.h — elementary functions written for this project, compiled identically
// for gfx950 device code (hipcc) and for host code (g++ / hipcc host pass).
//
// The tile rectangles, sort keys and per-pixel contributor indices of the splatting
// pipeline are integers derived from float arithmetic (exp of the log-scale, the
// covariance, the ceil of the 3-sigma radius, alpha thresholds). Integer outputs can
// only be bit-identical between the GPU path and a CPU restatement if both evaluate
// every float operation identically, so this header provides exp / log / sincos that
// use nothing but IEEE-754 +, *, fma, correctly rounded / and sqrt, rint and bit
// manipulation. Both sides must be built with -ffp-contract=off and without
// fast-math.
//
// Methods (from the mathematics, no vendor code):
//   ms_expf   Cody & Waite argument reduction x = n ln2 + r, |r| <= ln2/2 (two-part ln2
//             applied with fma), degree-7 Taylor polynomial of e^r by Horner's rule,
//             then scaling by 2^n through the exponent field.
//             (W. J. Cody, W. Waite, "Software Manual for the Elementary Functions", 1980.)
//   ms_logf   x = m 2^e with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s), s = (m-1)/(m+1),
//             odd series of atanh to s^11, plus e ln2 (two-part ln2).
//   ms_sincosf  reduction by k = rint(a 2/pi) with two-part pi/2 (fma), Taylor
//             polynomials of sin to r^9 and cos to r^10 on |r| <= pi/4, quadrant swap.
// Accuracy is checked against double-precision libm in tests/test_math.py.
#pragma once
#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MS_HD __host__ __device__ __forceinline__
#else
#define MS_HD static inline
#endif

MS_HD float ms_bits_to_float(uint32_t u) { return __builtin_bit_cast(float, u); }
MS_HD uint32_t ms_float_to_bits(float f) { return __builtin_bit_cast(uint32_t, f); }

// Exponent-field test: true unless x is +-Inf or NaN. Cannot be folded away by any
// floating-point optimisation setting.
MS_HD bool ms_finite(float x) { return (ms_float_to_bits(x) & 0x7F800000u) != 0x7F800000u; }

// float -> int32 with defined behaviour everywhere: truncation toward zero,
// saturation at the int32 range, NaN -> 0 (the semantics of v_cvt_i32_f32).
MS_HD int ms_f2i(float x) {
    if (!(x == x)) return 0;
    if (x >= 2147483520.0f) return 2147483647;
    if (x <= -2147483648.0f) return (-2147483647 - 1);
    return (int)x;
}

// Core of ms_expf without range handling: valid for |x| <= 87 (normal results), where it
// returns exactly what ms_expf returns.
MS_HD float ms_expf_core(float x) {
    const float n = rintf(x * 1.44269502162933349609375f);
    float r = fmaf(-n, 0.693147182464599609375f, x);
    r = fmaf(-n, -1.904654323148236e-09f, r);
    float p = 1.98412701e-04f;
    p = fmaf(p, r, 1.38888892e-03f);
    p = fmaf(p, r, 8.33333377e-03f);
    p = fmaf(p, r, 4.16666679e-02f);
    p = fmaf(p, r, 1.66666672e-01f);
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    return p * ms_bits_to_float((uint32_t)((int)n + 127) << 23);
}

MS_HD float ms_expf(float x) {
    if (!(x == x)) return x;
    if (x > 88.72283935546875f) return __builtin_inff();
    if (x < -103.97208404541015625f) return 0.0f;
    const float n = rintf(x * 1.44269502162933349609375f);          // x / ln2
    float r = fmaf(-n, 0.693147182464599609375f, x);                // ln2 high part
    r = fmaf(-n, -1.904654323148236e-09f, r);                       // ln2 low part
    float p = 1.98412701e-04f;                                      // 1/7!
    p = fmaf(p, r, 1.38888892e-03f);                                // 1/6!
    p = fmaf(p, r, 8.33333377e-03f);                                // 1/5!
    p = fmaf(p, r, 4.16666679e-02f);                                // 1/4!
    p = fmaf(p, r, 1.66666672e-01f);                                // 1/3!
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    int e = (int)n;
    if (e > 127) { p *= 2.0f; e -= 1; }
    if (e < -126) {                                                 // subnormal result
        p *= 5.42101086e-20f;                                       // 2^-64
        e += 64;
    }
    return p * ms_bits_to_float((uint32_t)(e + 127) << 23);
}

MS_HD float ms_logf(float x) {
    if (!(x == x)) return x;
    if (x < 0.0f) return __builtin_nanf("");
    if (x == 0.0f) return -__builtin_inff();
    if (!ms_finite(x)) return x;
    float scale_e = 0.0f;
    if (x < 1.17549435e-38f) { x *= 8388608.0f; scale_e = -23.0f; } // subnormal input
    const uint32_t u = ms_float_to_bits(x);
    int e = (int)((u >> 23) & 0xFFu) - 127;
    float m = ms_bits_to_float((u & 0x007FFFFFu) | 0x3F800000u);    // [1, 2)
    if (m > 1.41421353816986083984375f) { m *= 0.5f; e += 1; }       // [sqrt(1/2), sqrt(2)]
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    float p = 0.181818187f;                                          // 2/11
    p = fmaf(p, s2, 0.222222224f);                                   // 2/9
    p = fmaf(p, s2, 0.285714298f);                                   // 2/7
    p = fmaf(p, s2, 0.400000006f);                                   // 2/5
    p = fmaf(p, s2, 0.666666687f);                                   // 2/3
    const float logm = fmaf(p * s2, s, 2.0f * s);
    const float ef = (float)e + scale_e;
    return fmaf(ef, 0.693147182464599609375f, fmaf(ef, -1.904654323148236e-09f, logm));
}

MS_HD void ms_sincosf(float a, float* s_out, float* c_out) {
    const float k = rintf(a * 0.636619746685028076171875f);          // a * 2/pi
    float r = fmaf(-k, 1.57079637050628662109375f, a);               // pi/2 high
    r = fmaf(-k, -4.37113900018624283e-08f, r);                      // pi/2 low
    const float r2 = r * r;
    float ps = -2.50521079e-08f;                                     // -1/11!
    ps = fmaf(ps, r2, 2.75573188e-06f);                              // 1/9!
    ps = fmaf(ps, r2, -1.98412701e-04f);                             // -1/7!
    ps = fmaf(ps, r2, 8.33333377e-03f);                              // 1/5!
    ps = fmaf(ps, r2, -1.66666672e-01f);                             // -1/3!
    const float sn = fmaf(ps * r2, r, r);
    float pc = -2.75573188e-07f;                                     // -1/10!
    pc = fmaf(pc, r2, 2.48015876e-05f);                              // 1/8!
    pc = fmaf(pc, r2, -1.38888892e-03f);                             // -1/6!
    pc = fmaf(pc, r2, 4.16666679e-02f);                              // 1/4!
    pc = fmaf(pc, r2, -0.5f);
    const float cs = fmaf(pc, r2, 1.0f);
    const int q = ((int)k) & 3;
    float s, c;
    if (q == 0)      { s = sn;  c = cs;  }
    else if (q == 1) { s = cs;  c = -sn; }
    else if (q == 2) { s = -sn; c = -cs; }
    else             { s = -cs; c = sn;  }
    *s_out = s;
    *c_out = c;
}

// Logistic sigmoid in the reference's form 1 / (1 + exp(-x)).
MS_HD float ms_sigmoid(float x) { return 1.0f / (1.0f + ms_expf(-x)); }

// bfloat16 storage: round to nearest, ties to even, on the upper 16 bits of the float;
// NaN stays NaN (quiet), infinities stay infinite. Shared by the kernels and the oracle.
MS_HD uint16_t ms_bf16_bits(float x) {
    const uint32_t u = ms_float_to_bits(x);
    if ((u & 0x7F800000u) == 0x7F800000u) {
        if (u & 0x007FFFFFu) return (uint16_t)((u >> 16) | 0x0040u);
        return (uint16_t)(u >> 16);
    }
    const uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(r >> 16);
}
MS_HD float ms_bf16_float(uint16_t h) { return ms_bits_to_float((uint32_t)h << 16); }
MS_HD float ms_bf16_round(float x) { return ms_bf16_float(ms_bf16_bits(x)); }
