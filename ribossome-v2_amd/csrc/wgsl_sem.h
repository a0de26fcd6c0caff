/*
 * wgsl_sem.h - scalar semantics of the world-step arithmetic, one
 * implementation for the HIP kernels and for host code.
 *
 * The reference runs WGSL, whose exp/sin/pow accuracy is implementation
 * defined and whose float->int casts saturate.  Integer monomer counts are
 * decided by float comparisons, so the device kernels and any host
 * restatement must round every operation identically.  This header fixes
 * those operations:
 *   - rb_expf / rb_sinf: project-written, Cody-Waite range reduction followed
 *     by the Taylor polynomial of the reduced argument, evaluated by Horner
 *     with fused multiply-adds (one rounding each, identical on gfx950
 *     v_fma_f32 and x86 FMA3).  Tested against libm at <= 2 ulp.
 *   - round-half-to-even (WGSL round), floor, fract = x - floor(x).
 *   - saturating float -> u32 / i32 conversions (WGSL u32(f) / i32(f)).
 * Everything else is plain IEEE binary32 arithmetic, which both sides get
 * when compiled with -ffp-contract=off and correctly rounded divide/sqrt.
 */
#ifndef RIBO_WGSL_SEM_H
#define RIBO_WGSL_SEM_H

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define RB_FN __host__ __device__ __forceinline__
#else
#include <math.h>
#define RB_FN static inline
#endif

RB_FN float rb_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
RB_FN float rb_floor(float x) { return __builtin_floorf(x); }
/* WGSL round(): half to even. */
RB_FN float rb_rint(float x) { return __builtin_rintf(x); }
RB_FN float rb_fract(float x) { return x - __builtin_floorf(x); }
RB_FN float rb_sqrt(float x) { return __builtin_sqrtf(x); }
RB_FN float rb_abs(float x) { return __builtin_fabsf(x); }
/* WGSL min/max on finite operands. */
RB_FN float rb_min(float a, float b) { return b < a ? b : a; }
RB_FN float rb_max(float a, float b) { return a < b ? b : a; }
RB_FN float rb_clamp(float e, float lo, float hi) { return rb_min(rb_max(e, lo), hi); }
RB_FN uint32_t rb_minu(uint32_t a, uint32_t b) { return b < a ? b : a; }
RB_FN uint32_t rb_maxu(uint32_t a, uint32_t b) { return a < b ? b : a; }
RB_FN int32_t rb_clampi(int32_t e, int32_t lo, int32_t hi) { return e < lo ? lo : (e > hi ? hi : e); }
/* WGSL mix(a, b, t) = a * (1 - t) + b * t. */
RB_FN float rb_mix(float a, float b, float t) { return a * (1.0f - t) + b * t; }
/* WGSL length(vec2) = sqrt(dot(v, v)). */
RB_FN float rb_len2(float x, float y) { return __builtin_sqrtf(x * x + y * y); }
/* WGSL smoothstep(0, 1, x). */
RB_FN float rb_smoothstep01(float x) {
    float t = rb_clamp(x, 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

/* WGSL u32(f32): truncation toward zero, saturated to [0, 4294967040]
 * (the largest binary32 value below 2^32); NaN gives 0.  Written as a clamp
 * then a select so that GPU code has no branches (the clamp maps NaN to a
 * bound and the select replaces it). */
RB_FN uint32_t rb_f2u(float x) {
    const float c = fminf(fmaxf(x, 0.0f), 4294967040.0f);
    return x > 0.0f ? (uint32_t)c : 0u;
}
/* WGSL i32(f32): truncation toward zero, saturated to
 * [-2147483648, 2147483520]; NaN gives 0. */
RB_FN int32_t rb_f2i(float x) {
    const float c = fminf(fmaxf(x, -2147483648.0f), 2147483520.0f);
    return x == x ? (int32_t)c : 0;
}
/* Wrapping i32 add (WGSL integer arithmetic wraps). */
RB_FN int32_t rb_addi(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }

RB_FN uint32_t rb_as_u32(float f) {
    union { float f; uint32_t u; } c;
    c.f = f;
    return c.u;
}
RB_FN float rb_as_f32(uint32_t u) {
    union { uint32_t u; float f; } c;
    c.u = u;
    return c.f;
}
/* x * 2^k for an integer k in [-160, 128], exact whenever the result is a
 * normal or subnormal number (two power-of-two multiplies). */
RB_FN float rb_ldexp(float x, int k) {
    int k1 = k >> 1;
    int k2 = k - k1;
    float s1 = rb_as_f32((uint32_t)(k1 + 127) << 23);
    float s2 = rb_as_f32((uint32_t)(k2 + 127) << 23);
    return (x * s1) * s2;
}

/* exp(x) in binary32.
 * Method: Cody & Waite range reduction x = k ln2 + r with ln2 split in a
 * 15-bit head (k * head is exact for |k| <= 512) and a tail, |r| <= ln2/2,
 * then the degree-7 Taylor polynomial of exp(r) (truncation < 0.1 ulp) and
 * scaling by 2^k. */
RB_FN float rb_expf(float x) {
    if (x != x) return x;
    if (x > 88.72283935546875f) return __builtin_inff();
    if (x < -103.97208404541016f) return 0.0f;
    const float kf = rb_rint(x * 1.44269502162933349609375f);
    float r = rb_fma(kf, -0.693145751953125f, x);
    r = rb_fma(kf, -1.428606765330187e-06f, r);
    float p = 1.0f / 5040.0f;
    p = rb_fma(p, r, 1.0f / 720.0f);
    p = rb_fma(p, r, 1.0f / 120.0f);
    p = rb_fma(p, r, 1.0f / 24.0f);
    p = rb_fma(p, r, 1.0f / 6.0f);
    p = rb_fma(p, r, 0.5f);
    p = rb_fma(p, r, 1.0f);
    p = rb_fma(p, r, 1.0f);
    return rb_ldexp(p, (int)kf);
}

/* sin(x) in binary32 for |x| < 2^16 (the step evaluates it on [0, 16)).
 * Method: Cody & Waite reduction x = k pi/2 + r with pi/2 split in three
 * parts (the first two short enough that k * part is exact), then the
 * Taylor polynomials of sin r (to r^11) and cos r (to r^12) on
 * |r| <= pi/4, selected by the quadrant k mod 4. */
RB_FN float rb_sinf(float x) {
    if (x != x) return x;
    const float kf = rb_rint(x * 0.636619746685028076171875f);
    float r = rb_fma(kf, -1.5703125f, x);
    r = rb_fma(kf, -4.837512969970703125e-04f, r);
    r = rb_fma(kf, -7.549789948768648e-08f, r);
    const float r2 = r * r;
    float s = -1.0f / 39916800.0f;
    s = rb_fma(s, r2, 1.0f / 362880.0f);
    s = rb_fma(s, r2, -1.0f / 5040.0f);
    s = rb_fma(s, r2, 1.0f / 120.0f);
    s = rb_fma(s, r2, -1.0f / 6.0f);
    s = rb_fma(s * r2, r, r);
    float c = 1.0f / 479001600.0f;
    c = rb_fma(c, r2, -1.0f / 3628800.0f);
    c = rb_fma(c, r2, 1.0f / 40320.0f);
    c = rb_fma(c, r2, -1.0f / 720.0f);
    c = rb_fma(c, r2, 1.0f / 24.0f);
    c = rb_fma(c, r2, -0.5f);
    c = rb_fma(c, r2, 1.0f);
    const int q = ((int)kf) & 3;
    float v = (q & 1) ? c : s;
    return (q & 2) ? -v : v;
}

#endif
