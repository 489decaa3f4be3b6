/*
 * shade.h - device BSDF, medium, texture and intersection routines of the hot path.
 *
 * Each function restates one reference routine (file:line given) with the same float
 * operation order. Mixed float/double expressions of the reference (ILLUME_PI and
 * ILLUME_INV_PI are double literals, src/math/mathutils.h:6-7) are evaluated in double here as
 * well, so the float result is the same.
 */
#pragma once
#include "common.h"

/* Explicit address spaces: pointers read out of scene records are generic to the compiler,
 * which would turn every access into a flat_* instruction; these casts make them global (1)
 * or LDS (3) accesses. */
#if defined(__HIP_DEVICE_COMPILE__)
#define GLOBAL __attribute__((address_space(1)))
#define LDSAS __attribute__((address_space(3)))
#else   /* host pass: these functions are never called there */
#define GLOBAL
#define LDSAS
#endif
template <class T> __device__ inline const GLOBAL T* gp(const T* p) { return (const GLOBAL T*)p; }
template <class T> __device__ inline GLOBAL T* gpw(T* p) { return (GLOBAL T*)p; }

#define ILM_PI_D 3.14159265358979323846
#define ILM_INV_PI_D 0.318309886183790671538

/* src/fresnel.h:10-26 */
__device__ inline float fresnel_dielectric(float cosI, float ior, float& cosT) {
    float eta = cosI < 0.0f ? ior : 1.0f / ior;
    float sinTSq = eta * eta * (1.0f - cosI * cosI);
    if (sinTSq >= 1.0f) { cosT = 0.0f; return 1.0f; }
    cosI = pm_fabs(cosI);
    cosT = __builtin_sqrtf(1.0f - sinTSq);
    float rperp = (eta * cosI - cosT) / (eta * cosI + cosT);
    float rpar = (cosI - eta * cosT) / (cosI + eta * cosT);
    return (rperp * rperp + rpar * rpar) * 0.5f;
}
/* src/fresnel.h:28-49 */
__device__ inline float fresnel_conductor(float eta, float k, float cosI) {
    float cosISq = cosI * cosI;
    float sinISq = 1 - cosISq;
    float etaSq = eta * eta;
    float kSq = k * k;
    float t0 = etaSq - kSq - sinISq;
    float a2pb2 = __builtin_sqrtf(t0 * t0 + 4.0f * etaSq * kSq);
    float t1 = a2pb2 + cosISq;
    float a = __builtin_sqrtf(0.5f * (a2pb2 + t0));
    float t2 = 2.0f * a * cosI;
    float Rs = (t1 - t2) / (t1 + t2);
    float t3 = cosISq * a2pb2 + sinISq * sinISq;
    float t4 = t2 * sinISq;
    float Rp = Rs * (t3 - t4) / (t3 + t4);
    return 0.5f * (Rp + Rs);
}

/* src/microfacet.cu:23-37 */
__device__ inline float G1_beckmann(V3 v, V3 n, float alpha) {
    float cosV = dot3(v, n);
    float tanV = __builtin_sqrtf(1.0f - cosV * cosV) / cosV;
    float a = 1.0f / (alpha * pm_fabs(tanV));
    if ((double)a < 1.6) return (3.535f * a + 2.181f * a * a) / (1.0f + 2.276f * a + 2.577f * a * a);
    return 1.0f;
}
/* src/microfacet.cu:46-52 */
__device__ inline V3 sample_beckmann(float a, float u1, float u2) {
    float theta = pm_atanf(__builtin_sqrtf(-a * a * pm_logf(1.0f - u1)));
    float phi = (float)(2.0f * ILM_PI_D * (double)u2);
    float st, ct, sp, cp;
    pm_sincosf(theta, &st, &ct);
    pm_sincosf(phi, &sp, &cp);
    return mk3(st * cp, ct, st * sp);
}
/* GGX extension (Walter et al. 2007): D(m)cos(m) sampling, Smith G1 against the macro normal */
__device__ inline V3 sample_ggx(float a, float u1, float u2) {
    float theta = pm_atanf(a * __builtin_sqrtf(u1 / (1.0f - u1)));
    float phi = (float)(2.0f * ILM_PI_D * (double)u2);
    float st, ct, sp, cp;
    pm_sincosf(theta, &st, &ct);
    pm_sincosf(phi, &sp, &cp);
    return mk3(st * cp, ct, st * sp);
}
__device__ inline float G1_ggx(V3 v, V3 m, V3 n, float a) {
    float cosV = dot3(v, n);
    if (dot3(v, m) * cosV <= 0.0f) return 0.0f;
    float c2 = cosV * cosV;
    float tan2 = (1.0f - c2) / c2;
    return 2.0f / (1.0f + __builtin_sqrtf(1.0f + a * a * tan2));
}
__device__ inline V3 mf_sample(int dist, float a, float u1, float u2) {
    return dist == ILM_DIST_GGX ? sample_ggx(a, u1, u2) : sample_beckmann(a, u1, u2);
}
/* src/microfacet.cu:40-43 receives the microfacet normal as "n" (src/renderer.cu:259) */
__device__ inline float mf_G(int dist, V3 wi, V3 wo, V3 m, V3 n, float a) {
    if (dist == ILM_DIST_GGX) return G1_ggx(wi, m, n, a) * G1_ggx(wo, m, n, a);
    return G1_beckmann(wi, m, a) * G1_beckmann(wo, m, a);
}
/* src/renderer.cu:240,267,329 */
__device__ inline float mf_alpha(int dist, float roughness, float wiDotN) {
    if (dist == ILM_DIST_GGX) return roughness;
    return roughness * (1.2f - 0.2f * __builtin_sqrtf(pm_fabs(wiDotN)));
}

/* src/math/sample.cu:7-52 */
__device__ inline V3 sample_hemisphere_cosine(float u1, float u2) {
    float r = __builtin_sqrtf(u1);
    float t = (float)(2 * ILM_PI_D * (double)u2);
    float s, c;
    pm_sincosf(t, &s, &c);
    return mk3(r * c, __builtin_sqrtf(1 - u1), r * s);
}
__device__ inline V3 sample_circle(float u1, float u2) {
    float r = __builtin_sqrtf(u1);
    float d = (float)(2 * ILM_PI_D * (double)u2);
    float s, c;
    pm_sincosf(d, &s, &c);
    return mk3(r * c, r * s, 0.0f);
}
__device__ inline V3 sample_sphere(float u1, float u2) {
    float up = u1 * 2 - 1;
    float theta = (float)(2 * ILM_PI_D * (double)u2);
    float o = __builtin_sqrtf(1.0f - up * up);
    float s, c;
    pm_sincosf(theta, &s, &c);
    return mk3(c * o, up, s * o);
}
__device__ inline V3 sample_hg(float g, float u1, float u2) {
    if (g == 0.0f) return sample_sphere(u1, u2);
    float phi = (float)(2.0f * ILM_PI_D * (double)u1);
    float frac = (1.0f - g * g) / (1.0f + g * (2.0f * u2 - 1));
    float cosT = (1.0f + g * g - frac * frac) / (2.0f * g);
    float sinT = __builtin_sqrtf(__builtin_fmaxf(1.0f - cosT, 0.0f));
    float s, c;
    pm_sincosf(phi, &s, &c);
    return mk3(c * sinT, cosT, s * sinT);
}

/* src/texture.cu:129-162. Bitmap filtering: wrap, normalised coordinates, texel centres at
 * (i + 0.5) / w, bilinear for HDR and point for LDR, computed in software. */
__device__ inline V3 tex_eval(const GLOBAL DTex* tp, float u, float v) {
    const int type = tp->type;
    if (type == ILM_TEX_CONSTANT) return mk3(tp->a[0], tp->a[1], tp->a[2]);
    if (type == ILM_TEX_CHECKERBOARD) {
        int x = (int)((float)(int)tp->scale[0] * (u + 0.25f));
        int y = (int)((float)(int)tp->scale[1] * v);
        bool on = (x + y) % 2 == 0;
        return on ? mk3(tp->a[0], tp->a[1], tp->a[2]) : mk3(tp->b[0], tp->b[1], tp->b[2]);
    }
    const DTex t = *tp;
    u = u * t.scale[0];
    v = v * t.scale[1];
    float uw = u - __builtin_floorf(u), vw = v - __builtin_floorf(v);
    if (t.hdr) {
        float x = uw * (float)t.w - 0.5f, y = vw * (float)t.h - 0.5f;
        float fx = __builtin_floorf(x), fy = __builtin_floorf(y);
        float ax = x - fx, ay = y - fy;
        int x0 = (int)fx, y0 = (int)fy;
        x0 = ((x0 % t.w) + t.w) % t.w;
        y0 = ((y0 % t.h) + t.h) % t.h;
        int x1 = (x0 + 1) % t.w, y1 = (y0 + 1) % t.h;
        const GLOBAL float4* tf = gp(t.f);
        float4 p00 = tf[(size_t)y0 * t.w + x0];
        float4 p10 = tf[(size_t)y0 * t.w + x1];
        float4 p01 = tf[(size_t)y1 * t.w + x0];
        float4 p11 = tf[(size_t)y1 * t.w + x1];
        float wx = 1.0f - ax, wy = 1.0f - ay;
        float r0 = (p00.x * wx + p10.x * ax) * wy + (p01.x * wx + p11.x * ax) * ay;
        float r1 = (p00.y * wx + p10.y * ax) * wy + (p01.y * wx + p11.y * ax) * ay;
        float r2 = (p00.z * wx + p10.z * ax) * wy + (p01.z * wx + p11.z * ax) * ay;
        return mk3(r0, r1, r2);
    }
    int x = (int)(uw * (float)t.w), y = (int)(vw * (float)t.h);
    if (x >= t.w) x = t.w - 1;
    if (y >= t.h) y = t.h - 1;
    uchar4 p = gp(t.u8)[(size_t)y * t.w + x];
    return mk3((float)p.x / 255.0f, (float)p.y / 255.0f, (float)p.z / 255.0f);
}

/* Moller-Trumbore, src/primitives/mesh.cu:446-476 (e1, e2 formed from the vertices as there) */
__device__ inline float tri_mt(V3 v0, V3 v1, V3 v2, V3 o, V3 d, float& u, float& v) {
    V3 e1 = sub3(v1, v0), e2 = sub3(v2, v0);
    V3 P = cross3(d, e2);
    float det = dot3(e1, P);
    if (pm_fabs(det) < ILM_FLT_EPSILON) return -1.0f;
    det = 1.f / det;
    V3 T = sub3(o, v0);
    u = dot3(T, P) * det;
    if (u < 0.f || u > 1.f) return -1.0f;
    V3 Q = cross3(T, e1);
    v = dot3(d, Q) * det;
    if (v < 0.f || u + v > 1.f) return -1.0f;
    return dot3(e2, Q) * det;
}

/* Watertight ray/triangle test (Woop, Benthin, Wald, "Watertight Ray/Triangle Intersection",
 * JCGT 2(1), 2013): shear to ray space, 2D edge functions, double-precision fallback on zero
 * edge values. Barycentric u weights v1 and v weights v2 as in Moller-Trumbore. */
struct WTRay { int kx, ky, kz; float Sx, Sy, Sz; };
__device__ inline float comp3(V3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
__device__ inline WTRay wt_setup(V3 d) {
    WTRay r;
    float ax = pm_fabs(d.x), ay = pm_fabs(d.y), az = pm_fabs(d.z);
    r.kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    r.kx = (r.kz + 1) % 3;
    r.ky = (r.kx + 1) % 3;
    float dz = comp3(d, r.kz);
    if (dz < 0.0f) { int t = r.kx; r.kx = r.ky; r.ky = t; }
    r.Sx = comp3(d, r.kx) / dz;
    r.Sy = comp3(d, r.ky) / dz;
    r.Sz = 1.0f / dz;
    return r;
}
__device__ inline float tri_wt(const WTRay& w, V3 v0, V3 v1, V3 v2, V3 o, float& u, float& v) {
    V3 A = sub3(v0, o), B = sub3(v1, o), C = sub3(v2, o);
    float Az = comp3(A, w.kz), Bz = comp3(B, w.kz), Cz = comp3(C, w.kz);
    float Ax = comp3(A, w.kx) - w.Sx * Az, Ay = comp3(A, w.ky) - w.Sy * Az;
    float Bx = comp3(B, w.kx) - w.Sx * Bz, By = comp3(B, w.ky) - w.Sy * Bz;
    float Cx = comp3(C, w.kx) - w.Sx * Cz, Cy = comp3(C, w.ky) - w.Sy * Cz;
    float U = Cx * By - Cy * Bx;
    float Vv = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || Vv == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        Vv = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || Vv < 0.0f || W < 0.0f) && (U > 0.0f || Vv > 0.0f || W > 0.0f)) return -1.0f;
    float det = U + Vv + W;
    if (det == 0.0f) return -1.0f;
    float T = U * (w.Sz * Az) + Vv * (w.Sz * Bz) + W * (w.Sz * Cz);
    if ((T < 0.0f) != (det < 0.0f) && T != 0.0f) return -1.0f;
    float inv = 1.0f / det;
    u = Vv * inv;
    v = W * inv;
    return T * inv;
}

/* src/primitives/sphere.cu:43-77 (hit whenever the two rejection tests pass, no t > 0 test) */
__device__ inline bool sphere_hit(const DSphere& s, V3 o, V3 d, float& t) {
    V3 l = sub3(mk3(s.c[0], s.c[1], s.c[2]), o);
    float sd = dot3(l, d);
    float ls = dot3(l, l);
    float rs = s.r * s.r;
    if (sd < 0 && ls > rs) return false;
    float ms = ls - sd * sd;
    if (ms > rs) return false;
    float q = __builtin_sqrtf(rs - ms);
    t = sd;
    if (ls > rs) t -= q; else t += q;
    return true;
}
