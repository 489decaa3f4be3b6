/*
 * illume_pmath.h - portable single-precision math, PCG32 and the per-pass seed hash.
 *
 * Every function here is built only from IEEE-754 operations that are correctly rounded on
 * both x86-64 and gfx950 (+, -, *, /, sqrt, fma, floor, integer/bit operations). Compiled with
 * -ffp-contract=off on both sides, the GPU kernels and a host build therefore return the same
 * bits, which is what lets a path traced on the device be compared bit-for-bit with the same
 * path traced on the host.
 *
 * Methods (all written from the mathematics, no library code):
 *   sin/cos : Cody-Waite reduction by pi/2 in three fma steps, Taylor polynomials on [-pi/4, pi/4]
 *   exp     : reduction by ln2 (two-part constant), Taylor to r^7, scaling by two exact powers of 2
 *   log     : x = 2^e m, m in (sqrt2/2, sqrt2], log m = 2 atanh(s), s = (m-1)/(m+1), series to s^13
 *   atan    : 1/x for |x|>1, pi/4 shift for |x|>tan(pi/8), Taylor to t^19
 *   atan2, asin (via atan2 with sqrt((1-x)(1+x))), pow = exp(y log x)
 * Accuracy is a few ulp over the ranges the renderer uses; tests/test_oracle.py states the bounds
 * measured against double-precision libm.
 *
 * PCG32 is O'Neill's permuted congruential generator (pcg-random.org, "pcg32" XSH-RR, 64-bit
 * state): the multiplier and output permutation are the published ones.
 */
#ifndef ILLUME_PMATH_H
#define ILLUME_PMATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#define ILM_HD __host__ __device__ inline
#else
#define ILM_HD static inline
#endif

ILM_HD uint32_t pm_f2u(float f) { uint32_t u; __builtin_memcpy(&u, &f, 4); return u; }
ILM_HD float pm_u2f(uint32_t u) { float f; __builtin_memcpy(&f, &u, 4); return f; }
ILM_HD float pm_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
ILM_HD float pm_fabs(float x) { return pm_u2f(pm_f2u(x) & 0x7fffffffu); }
ILM_HD float pm_copysign(float x, float s) { return pm_u2f((pm_f2u(x) & 0x7fffffffu) | (pm_f2u(s) & 0x80000000u)); }
ILM_HD int pm_signbit(float x) { return (int)(pm_f2u(x) >> 31); }
ILM_HD int pm_isnan(float x) { return (pm_f2u(x) & 0x7fffffffu) > 0x7f800000u; }

#define PM_INF pm_u2f(0x7f800000u)
#define PM_NAN pm_u2f(0x7fc00000u)
#define PM_PI 0x1.921fb6p+1f
#define PM_PI_2 0x1.921fb6p+0f
#define PM_PI_4 0x1.921fb6p-1f

/* ---------------- sin / cos ---------------- */
ILM_HD float pm_sin_poly(float r, float r2) {
    float p = pm_fma(r2, -0x1.ae6456p-26f, 0x1.71de3ap-19f);
    p = pm_fma(r2, p, -0x1.a01a02p-13f);
    p = pm_fma(r2, p, 0x1.111112p-7f);
    p = pm_fma(r2, p, -0x1.555556p-3f);
    return pm_fma(r * r2, p, r);
}
ILM_HD float pm_cos_poly(float r2) {
    float p = pm_fma(r2, 0x1.1eed8ep-29f, -0x1.27e4fcp-22f);
    p = pm_fma(r2, p, 0x1.a01a02p-16f);
    p = pm_fma(r2, p, -0x1.6c16c2p-10f);
    p = pm_fma(r2, p, 0x1.555556p-5f);
    p = pm_fma(r2, p, -0.5f);
    return pm_fma(r2, p, 1.0f);
}
/* reduced argument r in [-pi/4, pi/4] and quadrant q */
ILM_HD float pm_reduce_pio2(float x, int* q) {
    float k = __builtin_floorf(x * 0x1.45f306p-1f + 0.5f);
    float r = pm_fma(-k, 0x1.921fb6p+0f, x);
    r = pm_fma(-k, -0x1.777a5cp-25f, r);
    r = pm_fma(-k, -0x1.ee59dap-50f, r);
    *q = ((int)k) & 3;
    return r;
}
ILM_HD void pm_sincosf(float x, float* s, float* c) {
    int q;
    float r = pm_reduce_pio2(x, &q);
    float r2 = r * r;
    float sr = pm_sin_poly(r, r2), cr = pm_cos_poly(r2);
    switch (q) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
    }
}
ILM_HD float pm_sinf(float x) { float s, c; pm_sincosf(x, &s, &c); return s; }
ILM_HD float pm_cosf(float x) { float s, c; pm_sincosf(x, &s, &c); return c; }

/* ---------------- exp ---------------- */
ILM_HD float pm_expf(float x) {
    if (pm_isnan(x)) return x;
    if (x > 0x1.62e42ep+6f) return PM_INF;      /* > 88.7228 */
    if (x < -0x1.9fe368p+6f) return 0.0f;       /* < -103.972 */
    float k = __builtin_floorf(x * 0x1.715476p+0f + 0.5f);
    float r = pm_fma(-k, 0x1.62e400p-1f, x);
    r = pm_fma(-k, 0x1.7f7d1cp-20f, r);
    float p = pm_fma(r, 0x1.a01a02p-13f, 0x1.6c16c2p-10f);
    p = pm_fma(r, p, 0x1.111112p-7f);
    p = pm_fma(r, p, 0x1.555556p-5f);
    p = pm_fma(r, p, 0x1.555556p-3f);
    p = pm_fma(r, p, 0.5f);
    p = pm_fma(r, p, 1.0f);
    p = pm_fma(r, p, 1.0f);
    int ki = (int)k;                  /* in [-150, 128] */
    int k1 = ki / 2, k2 = ki - k1;    /* both in [-75, 64]: normal powers of two */
    float s1 = pm_u2f((uint32_t)(k1 + 127) << 23);
    float s2 = pm_u2f((uint32_t)(k2 + 127) << 23);
    return (p * s1) * s2;
}

/* ---------------- log ---------------- */
ILM_HD float pm_logf(float x) {
    if (pm_isnan(x)) return x;
    if (x == 0.0f) return -PM_INF;
    if (x < 0.0f) return PM_NAN;
    if (x == PM_INF) return x;
    int e = 0;
    uint32_t u = pm_f2u(x);
    if (u < 0x00800000u) { x = x * 0x1p+23f; u = pm_f2u(x); e = -23; }
    e += (int)(u >> 23) - 127;
    float m = pm_u2f((u & 0x007fffffu) | 0x3f800000u);
    if (m > 0x1.6a09e6p+0f) { m = m * 0.5f; e += 1; }
    float f = m - 1.0f;
    float s = f / (2.0f + f);
    float s2 = s * s;
    float p = pm_fma(s2, 0x1.3b13b2p-3f, 0x1.745d18p-3f);
    p = pm_fma(s2, p, 0x1.c71c72p-3f);
    p = pm_fma(s2, p, 0x1.24924ap-2f);
    p = pm_fma(s2, p, 0x1.99999ap-2f);
    p = pm_fma(s2, p, 0x1.555556p-1f);
    float lm = pm_fma(s * s2, p, 2.0f * s);
    float ef = (float)e;
    return pm_fma(ef, 0x1.62e400p-1f, pm_fma(ef, 0x1.7f7d1cp-20f, lm));
}

/* ---------------- atan family ---------------- */
/* atan(a) for a in [0, 1] */
ILM_HD float pm_atan01(float a) {
    float off = 0.0f;
    if (a > 0x1.a8279ap-2f) { a = (a - 1.0f) / (a + 1.0f); off = PM_PI_4; }   /* tan(pi/8) */
    float t2 = a * a;
    float p = pm_fma(t2, 0x1.861862p-5f, -0x1.af286cp-5f);
    p = pm_fma(t2, p, 0x1.e1e1e2p-5f);
    p = pm_fma(t2, p, -0x1.111112p-4f);
    p = pm_fma(t2, p, 0x1.3b13b2p-4f);
    p = pm_fma(t2, p, -0x1.745d18p-4f);
    p = pm_fma(t2, p, 0x1.c71c72p-4f);
    p = pm_fma(t2, p, -0x1.24924ap-3f);
    p = pm_fma(t2, p, 0x1.99999ap-3f);
    p = pm_fma(t2, p, -0x1.555556p-2f);
    float r = pm_fma(a * t2, p, a);
    return off + r;
}
ILM_HD float pm_atanf(float x) {
    if (pm_isnan(x)) return x;
    float a = pm_fabs(x);
    float r = (a > 1.0f) ? PM_PI_2 - pm_atan01(1.0f / a) : pm_atan01(a);
    return pm_copysign(r, x);
}
ILM_HD float pm_atan2f(float y, float x) {
    if (pm_isnan(x) || pm_isnan(y)) return x + y;
    float ax = pm_fabs(x), ay = pm_fabs(y);
    float r;
    if (ay == 0.0f) {
        r = pm_signbit(x) ? PM_PI : 0.0f;
    } else if (ax == PM_INF && ay == PM_INF) {
        r = pm_signbit(x) ? 3.0f * PM_PI_4 : PM_PI_4;
    } else {
        r = (ay <= ax) ? pm_atan01(ay / ax) : PM_PI_2 - pm_atan01(ax / ay);
        if (pm_signbit(x)) r = PM_PI - r;
    }
    return pm_copysign(r, y);
}
ILM_HD float pm_asinf(float x) {
    float a = pm_fabs(x);
    if (!(a <= 1.0f)) return PM_NAN;
    float c = __builtin_sqrtf((1.0f - a) * (1.0f + a));
    return pm_atan2f(x, c);
}

/* ---------------- pow ---------------- */
ILM_HD float pm_powf(float x, float y) {
    if (y == 0.0f) return 1.0f;
    if (pm_isnan(x) || pm_isnan(y)) return PM_NAN;
    if (x == 0.0f) return y > 0.0f ? 0.0f : PM_INF;
    if (x < 0.0f) return PM_NAN;
    if (x == PM_INF) return y > 0.0f ? PM_INF : 0.0f;
    return pm_expf(y * pm_logf(x));
}

/* ---------------- per-pass seed hash and PCG32 ---------------- */
/* Thomas Wang's 32-bit integer hash, the per-pass seed of src/renderer.cu:488-496 */
ILM_HD uint32_t ilm_wang_hash(uint32_t a) {
    a = (a ^ 61u) ^ (a >> 16);
    a = a + (a << 3);
    a = a ^ (a >> 4);
    a = a * 0x27d4eb2du;
    a = a ^ (a >> 15);
    return a;
}

#define ILM_PCG_MULT 6364136223846793005ULL

ILM_HD uint32_t ilm_pcg32_next(uint64_t* state, uint64_t inc) {
    uint64_t old = *state;
    *state = old * ILM_PCG_MULT + inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((0u - rot) & 31u));
}
ILM_HD uint64_t ilm_pcg32_inc(uint64_t initseq) { return (initseq << 1u) | 1u; }
ILM_HD uint64_t ilm_pcg32_seed(uint64_t initstate, uint64_t initseq) {
    uint64_t inc = ilm_pcg32_inc(initseq);
    uint64_t s = 0u;
    ilm_pcg32_next(&s, inc);
    s += initstate;
    ilm_pcg32_next(&s, inc);
    return s;
}
/* uniform in (0, 1]: k * 2^-24 for k in [1, 2^24]; both steps are exact */
ILM_HD float ilm_pcg32_uniform(uint64_t* state, uint64_t inc) {
    uint32_t r = ilm_pcg32_next(state, inc);
    return (float)((r >> 8) + 1u) * 0x1p-24f;
}
/* Path seed: pass i, pixel p (the seeding shape of src/renderer.cu:61-69,541) */
ILM_HD uint64_t ilm_path_seed(uint32_t pass, uint32_t pixel) {
    return ilm_pcg32_seed((uint64_t)ilm_wang_hash(pass), (uint64_t)pixel);
}

#endif /* ILLUME_PMATH_H */
