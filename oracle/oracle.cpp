/*
 * oracle.cpp - CPU oracle for the path-tracer hot path.  TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library,
 * and only as the checker (or the timed CPU baseline), never as the product path.
 *
 * A fresh restatement of the reference's per-path routine (matthewshang/illume), templated on
 * the real type:
 *   R = float  : faithful FP32 restatement. Transcendentals come from include/illume_pmath.h so
 *                the stream of decisions is identical to the device kernels'.
 *   R = double : FP64 twin with libm, the precision oracle for image RMSE.
 * Algorithmic choices kept faithful to the reference (the "reference" CPU baseline):
 *   - BVH: the reference's sweep builder (src/accel/bvh.cpp:65-249, flattening :309-331)
 *   - traversal: exhaustive DFS with no t-pruning and a 32-entry stack whose overflow drops
 *     the instance (src/primitives/mesh.cu:479-562); instances in a linear loop
 *     (src/renderer.cu:112-137); Moller-Trumbore (src/primitives/mesh.cu:446-476)
 *   - AoS data, one path at a time.
 * The RNG is PCG32 seeded (wang_hash(pass), pixel) with the reference's draw order; the
 * reference's cuRAND stream is not reproducible (no CUDA here and its constants are not ours
 * to copy), so parity is against this stream.
 *
 * Reference behaviour restated here deliberately (so the oracle agrees with the reference):
 *   - the instance AABB test result is ignored (src/primitives/mesh_instance.cu:91-94)
 *   - leaf boxes are never tested, only inner-node boxes (src/primitives/mesh.cu:496-529)
 *   - G1 of the Beckmann model is evaluated against the microfacet normal m
 *     (src/renderer.cu:259 passes m as "n" to src/microfacet.cu:40-43)
 *   - Henyey-Greenstein sinT = sqrt(max(1 - cosT, 0)) (src/math/sample.cu:49)
 *   - plastic diffuse_fresnel has its arguments swapped (src/fresnel.h:68)
 *   - matrix determinant expansion as written (src/math/matrix4.cu:174-201); one term has a
 *     typo that vanishes for affine matrices
 * One deliberate deviation: get_bounds (src/primitives/mesh.cu:342-361) starts from an
 * uninitialised AABB; here it starts empty, the evident intent.
 *
 * How each part is pinned (tests/test_oracle.py, tests/test_bvh.py): PCG32 by its published
 * known-answer sequence; wang_hash by its integer recipe; portable math against double libm;
 * Fresnel, slab, sphere, Moller-Trumbore, tonemap, transforms and camera by closed forms and
 * constructed cases; samplers by their textbook distributions; the medium stage by Beer-Lambert
 * transmittance; traversal by brute force over every triangle.
 * Parity unpinned (nothing but the reference's source text fixes them, and the reference cannot
 * run here): the Henyey-Greenstein sinT = sqrt(1 - cosT) quirk (only cosT is checked), the
 * swapped arguments of the plastic diffuse_fresnel beyond its closed-form value, texture
 * bilinear weights between texel centres (CUDA tex2D's are not reproducible), and the
 * reference's cuRAND stream (replaced by PCG32 with the reference's draw order).
 */
#include "../include/illume_c.h"
#include "../include/illume_pmath.h"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

static const double ILLUME_PI = 3.14159265358979323846;     /* src/math/mathutils.h:6-7 */
static const double ILLUME_INV_PI = 0.318309886183790671538;

/* ---------------- math policy ---------------- */
template <class R> struct MathP;
template <> struct MathP<float> {
    static float sin(float x) { return pm_sinf(x); }
    static float cos(float x) { return pm_cosf(x); }
    static float exp(float x) { return pm_expf(x); }
    static float log(float x) { return pm_logf(x); }
    static float atan(float x) { return pm_atanf(x); }
    static float atan2(float y, float x) { return pm_atan2f(y, x); }
    static float asin(float x) { return pm_asinf(x); }
    static float pow(float x, float y) { return pm_powf(x, y); }
    static float sqrt(float x) { return std::sqrt(x); }
};
template <> struct MathP<double> {
    static double sin(double x) { return std::sin(x); }
    static double cos(double x) { return std::cos(x); }
    static double exp(double x) { return std::exp(x); }
    static double log(double x) { return std::log(x); }
    static double atan(double x) { return std::atan(x); }
    static double atan2(double y, double x) { return std::atan2(y, x); }
    static double asin(double x) { return std::asin(x); }
    static double pow(double x, double y) { return std::pow(x, y); }
    static double sqrt(double x) { return std::sqrt(x); }
};

template <class R> static inline R rmax() { return (R)FLT_MAX; }

/* ---------------- Vector3 (src/math/vector3.cu) ---------------- */
template <class R> struct V3 { R x, y, z; };
template <class R> static inline V3<R> v3(R x, R y, R z) { return V3<R>{x, y, z}; }
template <class R> static inline V3<R> vadd(V3<R> a, V3<R> b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class R> static inline V3<R> vadds(V3<R> a, R s) { return v3(s + a.x, s + a.y, s + a.z); }
template <class R> static inline V3<R> vsub(V3<R> a, V3<R> b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class R> static inline V3<R> vmul(V3<R> a, R s) { return v3(a.x * s, a.y * s, a.z * s); }
template <class R> static inline V3<R> vmulv(V3<R> a, V3<R> b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
template <class R> static inline V3<R> vdivs(V3<R> a, R s) { return v3(a.x / s, a.y / s, a.z / s); }
template <class R> static inline V3<R> vdivv(V3<R> a, V3<R> b) { return v3(a.x / b.x, a.y / b.y, a.z / b.z); }
template <class R> static inline R vdot(V3<R> a, V3<R> b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
template <class R> static inline V3<R> vcross(V3<R> a, V3<R> b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class R> static inline R vlen2(V3<R> v) { return (v.x * v.x) + (v.y * v.y) + (v.z * v.z); }
template <class R> static inline R vlen(V3<R> v) { return MathP<R>::sqrt(vlen2(v)); }
template <class R> static inline void vnormalize(V3<R>* v) {
    R inv = (R)1 / vlen(*v);
    v->x *= inv; v->y *= inv; v->z *= inv;
}
template <class R> static inline V3<R> vnormalized(V3<R> v) {
    R inv = (R)1 / vlen(v);
    return v3(v.x * inv, v.y * inv, v.z * inv);
}
template <class R> static inline V3<R> vmin(V3<R> v, R m) { return v3(std::fmin(v.x, m), std::fmin(v.y, m), std::fmin(v.z, m)); }
template <class R> static inline V3<R> vmax(V3<R> v, R m) { return v3(std::fmax(v.x, m), std::fmax(v.y, m), std::fmax(v.z, m)); }
template <class R> static inline V3<R> vpow(V3<R> v, R p) { return v3(MathP<R>::pow(v.x, p), MathP<R>::pow(v.y, p), MathP<R>::pow(v.z, p)); }
/* src/math/vector3.cu:129-133 */
template <class R> static inline V3<R> vreflect(V3<R> v, V3<R> n) { return vsub(v, vmul(vmul(n, vdot(v, n)), (R)2)); }
/* src/math/vector3.cu:85-104: sample y maps to the normal */
template <class R> static inline V3<R> vto_basis(V3<R> v, V3<R> n) {
    V3<R> t;
    if (std::fabs(n.x) > std::fabs(n.y)) t = v3(n.z, (R)0, -n.x);
    else t = v3((R)0, n.z, -n.y);
    V3<R> b = vcross(n, t);
    t = vcross(b, n);
    V3<R> r = vadd(vadd(vmul(t, v.x), vmul(n, v.y)), vmul(b, v.z));
    vnormalize(&r);
    return r;
}

template <class R> struct Ray { V3<R> o, d; };
/* src/math/ray.cu:3-22 */
template <class R> static inline void ray_set(Ray<R>* r, V3<R> o, V3<R> d) { r->o = o; r->d = d; vnormalize(&r->d); }
template <class R> static inline V3<R> ray_at(const Ray<R>& r, R t) { return vadd(r.o, vmul(r.d, t)); }

/* host-side sin/cos of the transform builder: the reference calls libm sinf/cosf on the host */
static inline float host_sin(float x) { return sinf(x); }
static inline double host_sin(double x) { return std::sin(x); }
static inline float host_cos(float x) { return cosf(x); }
static inline double host_cos(double x) { return std::cos(x); }

/* ---------------- Matrix4 (src/math/matrix4.cu) ---------------- */
template <class R> struct M4 { R m[4][4]; };
template <class R> static M4<R> m4_identity() {
    M4<R> a; for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) a.m[i][j] = (i == j) ? (R)1 : (R)0; return a;
}
template <class R> static M4<R> m4_axis_angle(V3<R> axis, R angle) {      /* :16-34 */
    vnormalize(&axis);
    V3<R> v = vmul(axis, host_sin(angle / 2));
    R w = host_cos(angle / 2);
    M4<R> m = m4_identity<R>();
    m.m[0][0] = 1 - 2 * (v.y * v.y + v.z * v.z);
    m.m[1][0] = 2 * (v.x * v.y + w * v.z);
    m.m[2][0] = 2 * (v.x * v.z - w * v.y);
    m.m[0][1] = 2 * (v.x * v.y - w * v.z);
    m.m[1][1] = 1 - 2 * (v.x * v.x + v.z * v.z);
    m.m[2][1] = 2 * (v.y * v.z + w * v.x);
    m.m[0][2] = 2 * (v.x * v.z + w * v.y);
    m.m[1][2] = 2 * (v.y * v.z - w * v.x);
    m.m[2][2] = 1 - 2 * (v.x * v.x + v.y * v.y);
    return m;
}
template <class R> static M4<R> m4_mul(const M4<R>& a, const M4<R>& b) {      /* :82-96 */
    M4<R> m;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            m.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j] + a.m[i][3] * b.m[3][j];
    return m;
}
template <class R> static M4<R> m4_transpose(const M4<R>& a) {
    M4<R> t; for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) t.m[j][i] = a.m[i][j]; return t;
}
static const int SUBIDX[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
template <class R> static R m4_subdet(const M4<R>& m, int i, int j) {          /* :122-138 */
    R d[3][3];
    for (int y = 0; y < 3; y++) for (int x = 0; x < 3; x++) d[y][x] = m.m[SUBIDX[i][y]][SUBIDX[j][x]];
    return (d[0][0] * d[1][1] * d[2][2]) + (d[0][1] * d[1][2] * d[2][0]) + (d[0][2] * d[1][0] * d[2][1])
         - (d[0][2] * d[1][1] * d[2][0]) - (d[0][1] * d[1][0] * d[2][2]) - (d[0][0] * d[1][2] * d[2][1]);
}
template <class R> static R m4_det(const M4<R>& M) {                           /* :174-201, as written */
    const R (*m)[4] = M.m;
    return (m[0][0] * m[1][1] * m[2][2] * m[3][3]) + (m[0][0] * m[1][2] * m[2][3] * m[3][1])
         + (m[0][0] * m[1][3] * m[2][1] * m[3][2]) + (m[0][1] * m[1][0] * m[2][3] * m[3][2])
         + (m[0][1] * m[1][2] * m[2][0] * m[3][3]) + (m[0][1] * m[1][3] * m[2][2] * m[3][0])
         + (m[0][2] * m[1][0] * m[2][1] * m[3][3]) + (m[0][2] * m[1][1] * m[2][3] * m[3][0])
         + (m[0][2] * m[1][3] * m[2][0] * m[3][1]) + (m[0][3] * m[2][0] * m[2][2] * m[3][1])
         + (m[0][3] * m[1][1] * m[2][0] * m[3][2]) + (m[0][3] * m[1][2] * m[2][1] * m[3][0])
         - (m[0][0] * m[1][1] * m[2][3] * m[3][2]) - (m[0][0] * m[1][2] * m[2][1] * m[3][3])
         - (m[0][0] * m[1][3] * m[2][2] * m[3][1]) - (m[0][1] * m[1][0] * m[2][2] * m[3][3])
         - (m[0][1] * m[1][2] * m[2][3] * m[3][0]) - (m[0][1] * m[1][3] * m[2][0] * m[3][2])
         - (m[0][2] * m[1][0] * m[2][3] * m[3][1]) - (m[0][2] * m[1][1] * m[2][0] * m[3][3])
         - (m[0][2] * m[1][3] * m[2][1] * m[3][0]) - (m[0][3] * m[1][0] * m[2][1] * m[3][2])
         - (m[0][3] * m[1][1] * m[2][2] * m[3][0]) - (m[0][3] * m[1][2] * m[2][0] * m[3][1]);
}
template <class R> static M4<R> m4_inverse(const M4<R>& m) {                   /* :115-120,140-172 */
    M4<R> ad;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            ad.m[i][j] = ((i + j) % 2 == 0) ? m4_subdet(m, j, i) : (R)-1 * m4_subdet(m, j, i);
    R s = (R)1 / m4_det(m);
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) ad.m[i][j] *= s;
    return ad;
}
template <class R> static M4<R> m4_lookat(V3<R> pos, V3<R> fwd, V3<R> up) {   /* :55-67 */
    V3<R> f = vnormalized(fwd);
    V3<R> r = vnormalized(vcross(up, f));
    V3<R> u = vnormalized(vcross(f, r));
    M4<R> m;
    m.m[0][0] = r.x; m.m[0][1] = u.x; m.m[0][2] = f.x; m.m[0][3] = pos.x;
    m.m[1][0] = r.y; m.m[1][1] = u.y; m.m[1][2] = f.y; m.m[1][3] = pos.y;
    m.m[2][0] = r.z; m.m[2][1] = u.z; m.m[2][2] = f.z; m.m[2][3] = pos.z;
    m.m[3][0] = 0; m.m[3][1] = 0; m.m[3][2] = 0; m.m[3][3] = 1;
    return m;
}
template <class R> static inline V3<R> m4_mul_v3(const M4<R>& m, V3<R> v, R w) {  /* :203-209 */
    return v3(m.m[0][0] * v.x + m.m[0][1] * v.y + m.m[0][2] * v.z + m.m[0][3] * w,
              m.m[1][0] * v.x + m.m[1][1] * v.y + m.m[1][2] * v.z + m.m[1][3] * w,
              m.m[2][0] * v.x + m.m[2][1] * v.y + m.m[2][2] * v.z + m.m[2][3] * w);
}

template <class R> struct Xform { M4<R> mat, inv, trans_inv; };

/* src/math/transform.cu:6-44 */
template <class R> static Xform<R> xform_from_desc(const ilm_transform_desc& t) {
    Xform<R> x;
    if (t.has_matrix) {
        for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) x.mat.m[i][j] = (R)t.matrix[i * 4 + j];
    } else if (t.has_lookat) {
        V3<R> tr = v3((R)t.translation[0], (R)t.translation[1], (R)t.translation[2]);
        V3<R> la = v3((R)t.lookat[0], (R)t.lookat[1], (R)t.lookat[2]);
        x.mat = m4_lookat(tr, vsub(la, tr), v3((R)0, (R)1, (R)0));
    } else {
        /* degtorad: deg * ILLUME_PI / 180.0f evaluated in double, returned as float */
        R rx = (R)((double)t.rotation[0] * ILLUME_PI / 180.0f);
        R ry = (R)((double)t.rotation[1] * ILLUME_PI / 180.0f);
        R rz = (R)((double)t.rotation[2] * ILLUME_PI / 180.0f);
        M4<R> mx = m4_axis_angle(v3((R)1, (R)0, (R)0), rx);
        M4<R> my = m4_axis_angle(v3((R)0, (R)1, (R)0), ry);
        M4<R> mz = m4_axis_angle(v3((R)0, (R)0, (R)1), rz);
        M4<R> rot = m4_mul(m4_mul(mx, my), mz);
        M4<R> tr = m4_identity<R>();
        tr.m[0][3] = (R)t.translation[0]; tr.m[1][3] = (R)t.translation[1]; tr.m[2][3] = (R)t.translation[2];
        M4<R> sc = m4_identity<R>();
        sc.m[0][0] = (R)t.scale[0]; sc.m[1][1] = (R)t.scale[1]; sc.m[2][2] = (R)t.scale[2];
        x.mat = m4_mul(m4_mul(tr, rot), sc);
    }
    x.inv = m4_inverse(x.mat);
    x.trans_inv = m4_transpose(x.inv);
    return x;
}

/* ---------------- AABB (src/math/aabb.cu) ---------------- */
template <class R> struct AABB { V3<R> lo, hi; };
template <class R> static inline AABB<R> aabb_empty() {
    return AABB<R>{v3((R)FLT_MAX, (R)FLT_MAX, (R)FLT_MAX), v3((R)-FLT_MAX, (R)-FLT_MAX, (R)-FLT_MAX)};
}
template <class R> static inline void aabb_update(AABB<R>* a, V3<R> p) {
    a->lo.x = std::fmin(a->lo.x, p.x); a->lo.y = std::fmin(a->lo.y, p.y); a->lo.z = std::fmin(a->lo.z, p.z);
    a->hi.x = std::fmax(a->hi.x, p.x); a->hi.y = std::fmax(a->hi.y, p.y); a->hi.z = std::fmax(a->hi.z, p.z);
}
/* :118-151 */
template <class R> static inline R aabb_ray_intersect(const AABB<R>& b, const Ray<R>& r) {
    R tmin = (R)-FLT_MAX, tmax = (R)FLT_MAX;
    if (r.d.x != 0) {
        R t1 = (b.lo.x - r.o.x) / r.d.x, t2 = (b.hi.x - r.o.x) / r.d.x;
        tmin = std::fmax(tmin, std::fmin(t1, t2)); tmax = std::fmin(tmax, std::fmax(t1, t2));
    }
    if (r.d.y != 0) {
        R t1 = (b.lo.y - r.o.y) / r.d.y, t2 = (b.hi.y - r.o.y) / r.d.y;
        tmin = std::fmax(tmin, std::fmin(t1, t2)); tmax = std::fmin(tmax, std::fmax(t1, t2));
    }
    if (r.d.z != 0) {
        R t1 = (b.lo.z - r.o.z) / r.d.z, t2 = (b.hi.z - r.o.z) / r.d.z;
        tmin = std::fmax(tmin, std::fmin(t1, t2)); tmax = std::fmin(tmax, std::fmax(t1, t2));
    }
    if (tmax > tmin && tmax >= 0) return tmin;
    return (R)-FLT_MAX;
}
/* src/primitives/mesh.cu:93-107 */
static inline void fix_aabb(AABB<float>* a) {
    if (a->hi.x - a->lo.x < FLT_EPSILON) a->hi.x += FLT_EPSILON;
    if (a->hi.y - a->lo.y < FLT_EPSILON) a->hi.y += FLT_EPSILON;
    if (a->hi.z - a->lo.z < FLT_EPSILON) a->hi.z += FLT_EPSILON;
}

/* ---------------- reference BVH (src/accel/bvh.cpp) ---------------- */
struct RefNode { AABB<float> bounds; uint32_t a, b; };   /* GPUNode: {left,right} or {count|0x80000000, start} */
struct RefBVH { std::vector<RefNode> nodes; std::vector<int> tri_indices; };

struct BWork { AABB<float> bounds; V3<float> center; int tri; };
struct BTree {
    AABB<float> bounds;
    int left = -1, right = -1;
    std::vector<int> tris;
    bool leaf = false;
};

static float axis_of(const V3<float>& v, int axis) { return axis == 0 ? v.x : (axis == 1 ? v.y : v.z); }

/* src/accel/bvh.cpp:65-249. The candidate planes of one node are costed in parallel; the
 * argmin keeps the reference's sequential order (axis-major, increasing split, strict <). */
static int ref_build(std::vector<BTree>& tree, std::vector<BWork*>& work, int depth, int tris_per_node) {
    const int BVH_STACK_SIZE = 32;
    int me = (int)tree.size();
    tree.emplace_back();
    if ((int)work.size() < tris_per_node || depth > BVH_STACK_SIZE) {
        tree[me].leaf = true;
        for (BWork* w : work) tree[me].tris.push_back(w->tri);
        return me;
    }
    AABB<float> bounds = aabb_empty<float>();
    for (BWork* w : work) { aabb_update(&bounds, w->bounds.lo); aabb_update(&bounds, w->bounds.hi); }
    float xdim = bounds.hi.x - bounds.lo.x, ydim = bounds.hi.y - bounds.lo.y, zdim = bounds.hi.z - bounds.lo.z;
    float best_cost = (float)work.size() * ((xdim * ydim) + (ydim * zdim) + (zdim * xdim));
    float best_split = FLT_MAX;
    int best_axis = -1;

    struct Cand { int axis; float split; };
    std::vector<Cand> cands;
    for (int axis = 0; axis < 3; axis++) {
        float start = axis_of(bounds.lo, axis), end = axis_of(bounds.hi, axis);
        if ((double)std::fabs(end - start) < 1e-4) continue;
        float step = (end - start) / (1024.f / ((float)depth + 1.0f));
        for (float split = start + step; split < end - step; split += step) cands.push_back({axis, split});
    }
    std::vector<float> costs(cands.size());
    const long n = (long)work.size();
#pragma omp parallel for schedule(dynamic, 4) if (n * (long)cands.size() > 200000)
    for (long c = 0; c < (long)cands.size(); c++) {
        int axis = cands[c].axis;
        float split = cands[c].split;
        AABB<float> lb = aabb_empty<float>(), rb = aabb_empty<float>();
        int ln = 0, rn = 0;
        for (long i = 0; i < n; i++) {
            BWork* w = work[i];
            if (axis_of(w->center, axis) < split) { aabb_update(&lb, w->bounds.lo); aabb_update(&lb, w->bounds.hi); ln++; }
            else { aabb_update(&rb, w->bounds.lo); aabb_update(&rb, w->bounds.hi); rn++; }
        }
        if (ln <= 1 || rn <= 1) { costs[c] = NAN; continue; }
        float lx = lb.hi.x - lb.lo.x, ly = lb.hi.y - lb.lo.y, lz = lb.hi.z - lb.lo.z;
        float ls = lx * ly + ly * lz + lz * lx;
        float rx = rb.hi.x - rb.lo.x, ry = rb.hi.y - rb.lo.y, rz = rb.hi.z - rb.lo.z;
        float rs = rx * ry + ry * rz + rz * rx;
        costs[c] = ls * (float)ln + rs * (float)rn;
    }
    for (size_t c = 0; c < cands.size(); c++) {
        if (std::isnan(costs[c])) continue;
        if (costs[c] < best_cost) { best_cost = costs[c]; best_split = cands[c].split; best_axis = cands[c].axis; }
    }
    if (best_axis == -1) {
        tree[me].leaf = true;
        for (BWork* w : work) tree[me].tris.push_back(w->tri);
        return me;
    }
    std::vector<BWork*> left, right;
    AABB<float> lb = aabb_empty<float>(), rb = aabb_empty<float>();
    for (BWork* w : work) {
        if (axis_of(w->center, best_axis) < best_split) { left.push_back(w); aabb_update(&lb, w->bounds.lo); aabb_update(&lb, w->bounds.hi); }
        else { right.push_back(w); aabb_update(&rb, w->bounds.lo); aabb_update(&rb, w->bounds.hi); }
    }
    std::vector<BWork*>().swap(work);
    int l = ref_build(tree, left, depth + 1, tris_per_node);
    int r = ref_build(tree, right, depth + 1, tris_per_node);
    tree[l].bounds = lb;
    tree[r].bounds = rb;
    tree[me].left = l;
    tree[me].right = r;
    return me;
}

/* src/accel/bvh.cpp:309-331: depth-first preorder */
static void ref_flatten(const std::vector<BTree>& tree, int node, RefBVH& out, int* node_index, int* tri_index) {
    int cur = *node_index;
    out.nodes[cur].bounds = tree[node].bounds;
    if (tree[node].leaf) {
        out.nodes[cur].a = 0x80000000u | (uint32_t)tree[node].tris.size();
        out.nodes[cur].b = (uint32_t)*tri_index;
        for (int t : tree[node].tris) out.tri_indices[(*tri_index)++] = t;
    } else {
        int li = ++(*node_index);
        ref_flatten(tree, tree[node].left, out, node_index, tri_index);
        int ri = ++(*node_index);
        ref_flatten(tree, tree[node].right, out, node_index, tri_index);
        out.nodes[cur].a = (uint32_t)li;
        out.nodes[cur].b = (uint32_t)ri;
    }
}

/* Median-split builder (not the reference's): longest centroid axis, split at the median,
 * leaves of at most tris_per_node triangles. O(N log N), depth about log2(N / leaf) so the
 * 32-entry traversal stack never overflows. Closest hits do not depend on the tree except for
 * ties, so the FP64 precision oracle may use it on large meshes where the reference sweep
 * builder takes minutes; the CPU baseline and the FP32 restatement keep the reference builder. */
static int median_build(std::vector<BTree>& tree, std::vector<BWork*>& work, size_t b, size_t e, int tris_per_node) {
    int me = (int)tree.size();
    tree.emplace_back();
    AABB<float> box = aabb_empty<float>(), cbox = aabb_empty<float>();
    for (size_t i = b; i < e; i++) {
        aabb_update(&box, work[i]->bounds.lo);
        aabb_update(&box, work[i]->bounds.hi);
        aabb_update(&cbox, work[i]->center);
    }
    tree[me].bounds = box;
    if (e - b <= (size_t)std::max(1, tris_per_node)) {
        tree[me].leaf = true;
        for (size_t i = b; i < e; i++) tree[me].tris.push_back(work[i]->tri);
        return me;
    }
    V3<float> ext = vsub(cbox.hi, cbox.lo);
    int axis = ext.x >= ext.y && ext.x >= ext.z ? 0 : (ext.y >= ext.z ? 1 : 2);
    size_t mid = b + (e - b) / 2;
    std::nth_element(work.begin() + b, work.begin() + mid, work.begin() + e, [axis](const BWork* x, const BWork* y) {
        float cx = axis_of(x->center, axis), cy = axis_of(y->center, axis);
        return cx < cy || (cx == cy && x->tri < y->tri);
    });
    int l = median_build(tree, work, b, mid, tris_per_node);
    int r = median_build(tree, work, mid, e, tris_per_node);
    tree[me].left = l;
    tree[me].right = r;
    return me;
}

static bool g_median_bvh = false;

static RefBVH ref_bvh_build(const std::vector<AABB<float>>& aabbs, AABB<float> bounds, int tris_per_node) {
    std::vector<BWork> work(aabbs.size());
    std::vector<BWork*> list(aabbs.size());
    for (size_t i = 0; i < aabbs.size(); i++) {
        work[i].bounds = aabbs[i];
        work[i].center = vmul(vadd(aabbs[i].lo, aabbs[i].hi), 0.5f);
        work[i].tri = (int)i;
        list[i] = &work[i];
    }
    std::vector<BTree> tree;
    tree.reserve(aabbs.size() / 2 + 16);
    int root = g_median_bvh ? median_build(tree, list, 0, list.size(), tris_per_node)
                            : ref_build(tree, list, 0, tris_per_node);
    tree[root].bounds = bounds;
    RefBVH out;
    out.nodes.resize(tree.size());
    size_t ntri = 0;
    for (auto& t : tree) if (t.leaf) ntri += t.tris.size();
    out.tri_indices.resize(ntri);
    int ni = 0, ti = 0;
    ref_flatten(tree, root, out, &ni, &ti);
    return out;
}

/* ---------------- scene ---------------- */
template <class R> struct Tex {
    int type = ILM_TEX_CONSTANT;
    V3<R> a{0, 0, 0}, b{0, 0, 0};
    float scale[2] = {1, 1};
    int w = 0, h = 0, hdr = 0;
    std::vector<R> f;          /* RGBA */
    std::vector<uint8_t> u8;   /* RGBA */
};
template <class R> struct Medium { V3<R> absorption{0, 0, 0}; R scattering = 0, g = 0; bool active = false; };
template <class R> struct Material {
    int type = ILM_MAT_DIFFUSE;
    Tex<R> albedo, roughness;
    R ior = 0, diffuse_fresnel = 0;
    int medium = -1;
    V3<R> eta{0, 0, 0}, k{0, 0, 0};
    int distribution = ILM_DIST_BECKMANN;
};
template <class R> struct Mesh {
    std::vector<V3<R>> pos, nrm;
    std::vector<float> uv;        /* 2 per vertex */
    std::vector<int> idx;         /* 3 per triangle */
    bool face_normals = true, has_texcoords = false;
    RefBVH bvh;
};
template <class R> struct Instance { int mesh, material; Xform<R> t; };
template <class R> struct Sphere { R r; V3<R> c; int material; };
template <class R> struct Scene {
    std::vector<Mesh<R>> meshes;
    std::vector<Instance<R>> inst;
    std::vector<Sphere<R>> spheres;
    std::vector<Material<R>> mats;
    std::vector<Medium<R>> mediums;
    Tex<R> env;
    bool tpruned = false;   /* mesh_intersect_pruned instead of the reference walk */
    R fov, dof, aperture;
    M4<R> cam;
};

/* src/texture.cu:129-162; bitmap filtering is this project's own (wrap, normalised
 * coordinates, texel centres at (i+0.5)/w, bilinear for HDR, point for LDR) */
template <class R> static V3<R> tex_eval(const Tex<R>& t, R u, R v) {
    switch (t.type) {
    case ILM_TEX_CONSTANT: return t.a;
    case ILM_TEX_CHECKERBOARD: {
        int x = (int)((R)(int)t.scale[0] * (u + (R)0.25f));
        int y = (int)((R)(int)t.scale[1] * v);
        bool on = (x + y) % 2 == 0;
        return on ? t.a : t.b;
    }
    case ILM_TEX_BITMAP: {
        u = u * (R)t.scale[0];
        v = v * (R)t.scale[1];
        R uw = u - std::floor(u), vw = v - std::floor(v);
        if (t.hdr) {
            R x = uw * (R)t.w - (R)0.5f, y = vw * (R)t.h - (R)0.5f;
            R fx = std::floor(x), fy = std::floor(y);
            R ax = x - fx, ay = y - fy;
            int x0 = (int)fx, y0 = (int)fy;
            x0 = ((x0 % t.w) + t.w) % t.w; y0 = ((y0 % t.h) + t.h) % t.h;
            int x1 = (x0 + 1) % t.w, y1 = (y0 + 1) % t.h;
            const R* p00 = &t.f[((size_t)y0 * t.w + x0) * 4];
            const R* p10 = &t.f[((size_t)y0 * t.w + x1) * 4];
            const R* p01 = &t.f[((size_t)y1 * t.w + x0) * 4];
            const R* p11 = &t.f[((size_t)y1 * t.w + x1) * 4];
            R c[3];
            for (int i = 0; i < 3; i++) {
                R top = p00[i] * ((R)1 - ax) + p10[i] * ax;
                R bot = p01[i] * ((R)1 - ax) + p11[i] * ax;
                c[i] = top * ((R)1 - ay) + bot * ay;
            }
            return v3(c[0], c[1], c[2]);
        } else {
            int x = (int)(uw * (R)t.w), y = (int)(vw * (R)t.h);
            if (x >= t.w) x = t.w - 1;
            if (y >= t.h) y = t.h - 1;
            const uint8_t* p = &t.u8[((size_t)y * t.w + x) * 4];
            return v3((R)p[0] / (R)255.0f, (R)p[1] / (R)255.0f, (R)p[2] / (R)255.0f);
        }
    }
    }
    return v3((R)0, (R)0, (R)0);
}

template <class R> static Tex<R> tex_from_desc(const ilm_texture_desc& d) {
    Tex<R> t;
    t.type = d.type;
    if (d.type == ILM_TEX_CONSTANT) t.a = v3((R)d.color[0], (R)d.color[1], (R)d.color[2]);
    else if (d.type == ILM_TEX_CHECKERBOARD) {
        t.a = v3((R)d.on[0], (R)d.on[1], (R)d.on[2]);
        t.b = v3((R)d.off[0], (R)d.off[1], (R)d.off[2]);
    }
    t.scale[0] = d.scale[0]; t.scale[1] = d.scale[1];
    if (d.type == ILM_TEX_BITMAP) {
        t.w = d.width; t.h = d.height; t.hdr = d.is_hdr;
        size_t n = (size_t)d.width * d.height * 4;
        if (d.is_hdr) { t.f.resize(n); for (size_t i = 0; i < n; i++) t.f[i] = (R)d.texels_f32[i]; }
        else t.u8.assign(d.texels_u8, d.texels_u8 + n);
    }
    return t;
}

/* ---------------- Fresnel (src/fresnel.h) ---------------- */
template <class R> static inline R fresnel_dielectric(R cosI, R ior, R& cosT) {    /* :10-26 */
    R eta = cosI < (R)0 ? ior : (R)1 / ior;
    R sinTSq = eta * eta * ((R)1 - cosI * cosI);
    if (sinTSq >= (R)1) { cosT = 0; return (R)1; }
    cosI = std::fabs(cosI);
    cosT = MathP<R>::sqrt((R)1 - sinTSq);
    R rperp = (eta * cosI - cosT) / (eta * cosI + cosT);
    R rpar = (cosI - eta * cosT) / (cosI + eta * cosT);
    return (rperp * rperp + rpar * rpar) * (R)0.5f;
}
template <class R> static inline R fresnel_conductor(R eta, R k, R cosI) {          /* :28-49 */
    R cosISq = cosI * cosI;
    R sinISq = 1 - cosISq;
    R etaSq = eta * eta;
    R kSq = k * k;
    R t0 = etaSq - kSq - sinISq;
    R a2pb2 = MathP<R>::sqrt(t0 * t0 + (R)4.0f * etaSq * kSq);
    R t1 = a2pb2 + cosISq;
    R a = MathP<R>::sqrt((R)0.5f * (a2pb2 + t0));
    R t2 = (R)2.0f * a * cosI;
    R Rs = (t1 - t2) / (t1 + t2);
    R t3 = cosISq * a2pb2 + sinISq * sinISq;
    R t4 = t2 * sinISq;
    R Rp = Rs * (t3 - t4) / (t3 + t4);
    return (R)0.5f * (Rp + Rs);
}
template <class R> static inline V3<R> fresnel_conductor3(V3<R> eta, V3<R> k, R cosI) {
    return v3(fresnel_conductor(eta.x, k.x, cosI), fresnel_conductor(eta.y, k.y, cosI), fresnel_conductor(eta.z, k.z, cosI));
}
/* :59-73, arguments swapped as in the reference */
template <class R> static R diffuse_fresnel(R ior, int samples) {
    R cosT;
    double ret = 0.0;
    R fb = fresnel_dielectric((R)0.0, ior, cosT);
    for (int i = 1; i <= samples; i++) {
        R cosTSq = (R)i / (R)samples;
        R fa = fresnel_dielectric(ior, MathP<R>::sqrt(cosTSq), cosT);
        ret += (double)(fa + fb) * (0.5 / samples);
        fb = fa;
    }
    return (R)ret;
}

/* ---------------- microfacet (src/microfacet.cu) ---------------- */
template <class R> static inline R G1_beckmann(V3<R> v, V3<R> n, R alpha) {        /* :23-37 */
    R cosV = vdot(v, n);
    R tanV = MathP<R>::sqrt((R)1 - cosV * cosV) / cosV;
    R a = (R)1 / (alpha * std::fabs(tanV));
    if ((double)a < 1.6) return ((R)3.535f * a + (R)2.181f * a * a) / ((R)1 + (R)2.276f * a + (R)2.577f * a * a);
    return (R)1;
}
template <class R> static inline V3<R> sample_beckmann(R a, R u1, R u2) {           /* :46-52 */
    R theta = MathP<R>::atan(MathP<R>::sqrt(-a * a * MathP<R>::log((R)1 - u1)));
    R phi = (R)(2.0f * ILLUME_PI * (double)u2);
    R st = MathP<R>::sin(theta), ct = MathP<R>::cos(theta);
    return v3(st * MathP<R>::cos(phi), ct, st * MathP<R>::sin(phi));
}
/* GGX extension (Walter et al. 2007): D(m)cos sampling, Smith G1 against the macro normal */
template <class R> static inline V3<R> sample_ggx(R a, R u1, R u2) {
    R theta = MathP<R>::atan(a * MathP<R>::sqrt(u1 / ((R)1 - u1)));
    R phi = (R)(2.0f * ILLUME_PI * (double)u2);
    R st = MathP<R>::sin(theta), ct = MathP<R>::cos(theta);
    return v3(st * MathP<R>::cos(phi), ct, st * MathP<R>::sin(phi));
}
template <class R> static inline R G1_ggx(V3<R> v, V3<R> m, V3<R> n, R a) {
    R cosV = vdot(v, n);
    if (vdot(v, m) * cosV <= (R)0) return (R)0;
    R c2 = cosV * cosV;
    R tan2 = ((R)1 - c2) / c2;
    return (R)2 / ((R)1 + MathP<R>::sqrt((R)1 + a * a * tan2));
}
template <class R> static inline V3<R> mf_sample(int dist, R a, R u1, R u2) {
    return dist == ILM_DIST_GGX ? sample_ggx(a, u1, u2) : sample_beckmann(a, u1, u2);
}
template <class R> static inline R mf_G(int dist, V3<R> wi, V3<R> wo, V3<R> m, V3<R> n, R a) {
    if (dist == ILM_DIST_GGX) return G1_ggx(wi, m, n, a) * G1_ggx(wo, m, n, a);
    return G1_beckmann(wi, m, a) * G1_beckmann(wo, m, a);   /* src/microfacet.cu:40-43 with n := m */
}
/* roughness remap of src/renderer.cu:240,267,329 (Beckmann); GGX uses alpha = roughness */
template <class R> static inline R mf_alpha(int dist, R roughness, R wiDotN) {
    if (dist == ILM_DIST_GGX) return roughness;
    return roughness * ((R)1.2f - (R)0.2f * MathP<R>::sqrt(std::fabs(wiDotN)));
}

/* ---------------- samplers (src/math/sample.cu) ---------------- */
template <class R> static inline V3<R> sample_hemisphere_cosine(R u1, R u2) {
    R r = MathP<R>::sqrt(u1);
    R t = (R)(2 * ILLUME_PI * (double)u2);
    return v3(r * MathP<R>::cos(t), MathP<R>::sqrt((R)1 - u1), r * MathP<R>::sin(t));
}
template <class R> static inline V3<R> sample_circle(R u1, R u2) {
    R r = MathP<R>::sqrt(u1);
    R d = (R)(2 * ILLUME_PI * (double)u2);
    return v3(r * MathP<R>::cos(d), r * MathP<R>::sin(d), (R)0);
}
template <class R> static inline V3<R> sample_sphere(R u1, R u2) {
    R up = u1 * 2 - 1;
    R theta = (R)(2 * ILLUME_PI * (double)u2);
    R o = MathP<R>::sqrt((R)1.0f - up * up);
    return v3(MathP<R>::cos(theta) * o, up, MathP<R>::sin(theta) * o);
}
template <class R> static inline V3<R> sample_hg(R g, R u1, R u2) {
    if (g == (R)0.0f) return sample_sphere(u1, u2);
    R phi = (R)(2.0f * ILLUME_PI * (double)u1);
    R frac = ((R)1.0f - g * g) / ((R)1.0f + g * ((R)2.0f * u2 - 1));
    R cosT = ((R)1.0f + g * g - frac * frac) / ((R)2.0f * g);
    R sinT = MathP<R>::sqrt(std::fmax((R)1.0f - cosT, (R)0.0f));
    return v3(MathP<R>::cos(phi) * sinT, cosT, MathP<R>::sin(phi) * sinT);
}

/* ---------------- intersection ---------------- */
template <class R> struct Hit {
    bool is = false;
    R d = (R)FLT_MAX;
    V3<R> normal{0, 0, 0};
    int mat = -1;
    R u = 0, v = 0;
    int inst = -1, prim = -1;
};

/* src/primitives/mesh.cu:446-476 */
template <class R> static inline R tri_intersect(V3<R> v0, V3<R> v1, V3<R> v2, const Ray<R>& ray, R& u, R& v) {
    V3<R> e1 = vsub(v1, v0), e2 = vsub(v2, v0);
    V3<R> P = vcross(ray.d, e2);
    R det = vdot(e1, P);
    if (std::fabs(det) < (R)FLT_EPSILON) return (R)-1;
    det = (R)1.f / det;
    V3<R> T = vsub(ray.o, v0);
    u = vdot(T, P) * det;
    if (u < (R)0.f || u > (R)1.f) return (R)-1;
    V3<R> Q = vcross(T, e1);
    v = vdot(ray.d, Q) * det;
    if (v < (R)0.f || u + v > (R)1.f) return (R)-1;
    return vdot(e2, Q) * det;
}

template <class R> static void mesh_hit_attributes(const Mesh<R>& mesh, int tri_index, Hit<R>* hit);

/* The build's traversal order over the same tree, for the second CPU baseline figure (not the
 * reference's walk): a node's two children are box-tested together, the nearer is entered
 * first, and a subtree is skipped once its entry distance lies beyond the closest hit so far
 * (tmax: the closest hit of the objects tested before this instance). The hits equal
 * mesh_intersect's except at exact distance ties, and no subtree is dropped for stack depth. */
template <class R> static void mesh_intersect_pruned(const Mesh<R>& mesh, const Ray<R>& ray, R tmax, Hit<R>* hit,
                                                     int* box_tests, int* tri_tests) {
    struct Entry { uint32_t node; R t; };
    Entry stack[128];
    int sp = 0;
    hit->is = false;
    hit->d = (R)FLT_MAX;
    R best = tmax;
    int tri_index = -1;
    /* boxes of inner nodes only: the reference walk never tests a leaf's box, and a leaf of
     * a flat mesh can have a box of zero thickness; a leaf inherits its parent's distance */
    R parent_t = 0;
    auto enter = [&](uint32_t n, R* t) {
        const RefNode& c = mesh.bvh.nodes[n];
        if (c.a & 0x80000000u) { *t = parent_t; return true; }
        AABB<R> b{v3((R)c.bounds.lo.x, (R)c.bounds.lo.y, (R)c.bounds.lo.z), v3((R)c.bounds.hi.x, (R)c.bounds.hi.y, (R)c.bounds.hi.z)};
        if (box_tests) (*box_tests)++;
        R tb = aabb_ray_intersect(b, ray);
        if (tb == (R)-FLT_MAX) return false;
        *t = tb > (R)0 ? tb : (R)0;
        return *t <= best;
    };
    R t0;
    if (mesh.bvh.nodes.empty() || !enter(0, &t0)) return;
    stack[sp++] = {0u, t0};
    while (sp > 0) {
        const Entry e = stack[--sp];
        if (e.t > best) continue;
        const RefNode& node = mesh.bvh.nodes[e.node];
        if (node.a & 0x80000000u) {
            for (uint32_t i = node.b; i < node.b + (node.a & 0x7fffffffu); i++) {
                int t = mesh.bvh.tri_indices[i];
                R u = 0, v = 0;
                if (tri_tests) (*tri_tests)++;
                R d = tri_intersect(mesh.pos[mesh.idx[3 * t]], mesh.pos[mesh.idx[3 * t + 1]], mesh.pos[mesh.idx[3 * t + 2]], ray, u, v);
                if (d > 0 && d < best) { tri_index = t; hit->u = u; hit->v = v; best = d; }
            }
        } else {
            R ta = 0, tb = 0;
            parent_t = e.t;
            const bool ha = enter(node.a, &ta), hb = enter(node.b, &tb);
            if (ha && hb) {
                const bool a_first = ta <= tb;
                stack[sp++] = a_first ? Entry{node.b, tb} : Entry{node.a, ta};
                stack[sp++] = a_first ? Entry{node.a, ta} : Entry{node.b, tb};
            } else if (ha) {
                stack[sp++] = {node.a, ta};
            } else if (hb) {
                stack[sp++] = {node.b, tb};
            }
        }
    }
    if (tri_index != -1) {
        hit->is = true;
        hit->d = best;
        mesh_hit_attributes(mesh, tri_index, hit);
    }
}

/* src/primitives/mesh.cu:479-562; returns is_intersect, d, uv, normal (object space), prim */
template <class R> static void mesh_intersect(const Mesh<R>& mesh, const Ray<R>& ray, Hit<R>* hit, int* box_tests, int* tri_tests) {
    const int BVH_STACK_SIZE = 32;
    uint32_t stack[BVH_STACK_SIZE + 2];
    int sp = 0;
    stack[sp++] = 0;
    hit->is = false;
    hit->d = (R)FLT_MAX;
    int tri_index = -1;
    while (sp > 0) {
        const RefNode& node = mesh.bvh.nodes[stack[sp - 1]];
        sp--;
        if (node.a & 0x80000000u) {
            for (uint32_t i = node.b; i < node.b + (node.a & 0x7fffffffu); i++) {
                int t = mesh.bvh.tri_indices[i];
                R u = 0, v = 0;
                if (tri_tests) (*tri_tests)++;
                R d = tri_intersect(mesh.pos[mesh.idx[3 * t]], mesh.pos[mesh.idx[3 * t + 1]], mesh.pos[mesh.idx[3 * t + 2]], ray, u, v);
                if (d > 0 && d < hit->d) { tri_index = t; hit->u = u; hit->v = v; hit->d = d; hit->is = true; }
            }
        } else {
            AABB<R> b{v3((R)node.bounds.lo.x, (R)node.bounds.lo.y, (R)node.bounds.lo.z),
                      v3((R)node.bounds.hi.x, (R)node.bounds.hi.y, (R)node.bounds.hi.z)};
            if (box_tests) (*box_tests)++;
            if (aabb_ray_intersect(b, ray) != (R)-FLT_MAX) {
                stack[sp++] = node.b;
                stack[sp++] = node.a;
                if (sp > BVH_STACK_SIZE) { hit->is = false; return; }
            }
        }
    }
    if (tri_index != -1) mesh_hit_attributes(mesh, tri_index, hit);
}

/* normal and texture coordinates of a triangle hit, src/primitives/mesh.cu:532-561 */
template <class R> static void mesh_hit_attributes(const Mesh<R>& mesh, int tri_index, Hit<R>* hit) {
    {
        const int* ix = &mesh.idx[3 * tri_index];
        if (mesh.face_normals) {
            V3<R> v0 = mesh.pos[ix[0]], v1 = mesh.pos[ix[1]], v2 = mesh.pos[ix[2]];
            hit->normal = vcross(vsub(v1, v0), vsub(v2, v0));
            vnormalize(&hit->normal);
        } else {
            V3<R> n0 = mesh.nrm[ix[0]], n1 = mesh.nrm[ix[1]], n2 = mesh.nrm[ix[2]];
            hit->normal = vadd(vmul(n0, (R)1.0f - hit->u - hit->v), vadd(vmul(n1, hit->u), vmul(n2, hit->v)));
            vnormalize(&hit->normal);
        }
        if (mesh.has_texcoords) {
            R w = (R)1.0f - hit->u - hit->v;
            const float* t0 = &mesh.uv[2 * ix[0]];
            const float* t1 = &mesh.uv[2 * ix[1]];
            const float* t2 = &mesh.uv[2 * ix[2]];
            R nu = ((R)t0[0] * w + (R)t1[0] * hit->u) + (R)t2[0] * hit->v;
            R nv = ((R)t0[1] * w + (R)t1[1] * hit->u) + (R)t2[1] * hit->v;
            hit->u = nu; hit->v = nv;
        }
        hit->prim = tri_index;
    }
}

/* src/primitives/sphere.cu:43-77 */
template <class R> static void sphere_intersect(const Sphere<R>& s, const Ray<R>& ray, Hit<R>* hit) {
    V3<R> l = vsub(s.c, ray.o);
    R sd = vdot(l, ray.d);
    R ls = vdot(l, l);
    R rs = s.r * s.r;
    if (sd < 0 && ls > rs) { hit->is = false; return; }
    R ms = ls - sd * sd;
    if (ms > rs) { hit->is = false; return; }
    R q = MathP<R>::sqrt(rs - ms);
    R t = sd;
    if (ls > rs) t -= q; else t += q;
    V3<R> pos = ray_at(ray, t);
    V3<R> normal = vsub(pos, s.c);
    vnormalize(&normal);
    hit->is = true;
    hit->d = t;
    hit->normal = normal;
    hit->mat = s.material;
    hit->u = (R)((double)(MathP<R>::atan2(normal.z, normal.x) * (R)0.5f) * ILLUME_INV_PI + 0.5f);
    hit->v = (R)(0.5f - (double)MathP<R>::asin(normal.y) * ILLUME_INV_PI);
}

/* src/renderer.cu:112-137 with src/primitives/mesh_instance.cu:88-108 */
template <class R> static Hit<R> get_min_hit(const Scene<R>& sc, const Ray<R>& ray, int* box_tests, int* tri_tests) {
    Hit<R> min;
    min.d = (R)FLT_MAX;
    for (size_t i = 0; i < sc.spheres.size(); i++) {
        Hit<R> h;
        sphere_intersect(sc.spheres[i], ray, &h);
        if (h.is && h.d < min.d) { min = h; min.inst = -2 - (int)i; min.prim = -1; }
    }
    for (size_t i = 0; i < sc.inst.size(); i++) {
        const Instance<R>& in = sc.inst[i];
        Ray<R> r2;
        r2.o = m4_mul_v3(in.t.inv, ray.o, (R)1);
        r2.d = m4_mul_v3(in.t.inv, ray.d, (R)0);
        Hit<R> h;
        if (sc.tpruned) mesh_intersect_pruned(sc.meshes[in.mesh], r2, min.d, &h, box_tests, tri_tests);
        else mesh_intersect(sc.meshes[in.mesh], r2, &h, box_tests, tri_tests);
        if (h.is) {
            h.normal = m4_mul_v3(in.t.trans_inv, h.normal, (R)0);
            vnormalize(&h.normal);
            h.mat = in.material;
            if (h.d < min.d) { min = h; min.inst = (int)i; }
        }
    }
    return min;
}

/* ---------------- camera (src/renderer.cu:71-79,443-458) ---------------- */
template <class R> struct RenderInfo { R image_width, pixel_size, left, top; };
template <class R> static RenderInfo<R> make_render_info(const Scene<R>& sc, int width, int height) {
    RenderInfo<R> i;
    i.image_width = (R)width;
    R dim_ratio = (R)height / (R)width;
    R tan_half_fov = sizeof(R) == 4 ? (R)tanf((float)(ILLUME_PI * (double)sc.fov / 360))
                                    : (R)std::tan(ILLUME_PI * (double)sc.fov / 360);
    R dofmfov = sc.dof * tan_half_fov;
    i.pixel_size = dofmfov * 2 / (R)width;
    i.left = -1 * dofmfov;
    i.top = dim_ratio * dofmfov;
    return i;
}

/* ---------------- one path (src/renderer.cu:81-110 and 139-388) ---------------- */
struct PathStats {
    int segments = 0;
    float* rec = nullptr;   /* nullable: 9 floats per segment (origin, direction, inst, prim, t) */
    int rec_max = 0;
};

template <class R> static V3<R> trace_path(const Scene<R>& sc, const RenderInfo<R>& ri, const ilm_render_params& p,
                                            uint32_t pass, uint32_t pixel, PathStats* st) {
    typedef MathP<R> Mth;
    uint64_t rs = ilm_path_seed(pass, pixel);
    const uint64_t inc = ilm_pcg32_inc(pixel);
    auto rnd = [&]() -> R { return (R)ilm_pcg32_uniform(&rs, inc); };

    /* init_rays */
    int W = (int)ri.image_width;
    int x = (int)pixel % W, y = (int)pixel / W;
    R left_edge = ri.left + ri.pixel_size * (R)x;
    R top_edge = ri.top - ri.pixel_size * (R)y;
    R rx = left_edge + ri.pixel_size * rnd();
    R ry = top_edge - ri.pixel_size * rnd();
    V3<R> pos = v3((R)0, (R)0, (R)0);
    if (sc.aperture > (R)FLT_EPSILON) {
        R u1 = rnd(), u2 = rnd();
        pos = vmul(sample_circle(u1, u2), sc.aperture);
    }
    V3<R> origin = m4_mul_v3(sc.cam, pos, (R)1);
    V3<R> image_pos = m4_mul_v3(sc.cam, v3(rx, ry, sc.dof), (R)1);
    Ray<R> ray;
    ray_set(&ray, origin, vsub(image_pos, origin));
    V3<R> color = v3((R)1, (R)1, (R)1);
    int medium = -1;
    const R bias = (R)p.ray_bias;

    for (int depth = 0; depth < p.max_depth; depth++) {
        st->segments++;
        Hit<R> min = get_min_hit(sc, ray, nullptr, nullptr);
        if (st->rec && depth < st->rec_max) {
            float* q = st->rec + 9 * depth;
            q[0] = (float)ray.o.x; q[1] = (float)ray.o.y; q[2] = (float)ray.o.z;
            q[3] = (float)ray.d.x; q[4] = (float)ray.d.y; q[5] = (float)ray.d.z;
            q[6] = min.is ? (float)min.inst : -1.0f; q[7] = min.is ? (float)min.prim : -1.0f;
            q[8] = min.is ? (float)min.d : FLT_MAX;
        }
        if (medium >= 0 && sc.mediums[medium].active) {
            const Medium<R>& md = sc.mediums[medium];
            R u = rnd();
            R scatter_dist = -Mth::log(u) / md.scattering;
            if (scatter_dist < min.d) {
                R u1 = rnd(), u2 = rnd();
                V3<R> sdir = vto_basis(sample_hg(md.g, u1, u2), ray.d);
                ray_set(&ray, ray_at(ray, scatter_dist), sdir);
                color = vmulv(color, v3(Mth::exp((R)-1.0f * scatter_dist * md.absorption.x),
                                         Mth::exp((R)-1.0f * scatter_dist * md.absorption.y),
                                         Mth::exp((R)-1.0f * scatter_dist * md.absorption.z)));
                if ((double)vlen2(color) < 1e-4) return v3((R)0, (R)0, (R)0);
                continue;
            } else {
                color = vmulv(color, v3(Mth::exp((R)-1.0f * min.d * md.absorption.x),
                                         Mth::exp((R)-1.0f * min.d * md.absorption.y),
                                         Mth::exp((R)-1.0f * min.d * md.absorption.z)));
            }
        }
        if (min.is) {
            const Material<R>& m = sc.mats[min.mat];
            Ray<R> r = ray;
            V3<R> n = min.normal;
            V3<R> new_dir;
            V3<R> norm_o = vmul(n, vdot(n, r.d) > 0 ? (R)-1.0f : (R)1.0f);
            V3<R> new_origin = ray_at(r, min.d);
            V3<R> albedo = tex_eval(m.albedo, min.u, min.v);
            R roughness = tex_eval(m.roughness, min.u, min.v).x;
            switch (m.type) {
            case ILM_MAT_EMISSIVE:
                color = vmulv(color, albedo);
                return color;
            case ILM_MAT_DIFFUSE: {
                color = vmulv(color, albedo);
                R u1 = rnd(), u2 = rnd();
                new_dir = vto_basis(sample_hemisphere_cosine(u1, u2), norm_o);
                new_origin = vadd(new_origin, vmul(norm_o, bias));
                break;
            }
            case ILM_MAT_SPECULAR:
                color = vmulv(color, v3((R)0.99f, (R)0.99f, (R)0.99f));
                new_dir = vreflect(r.d, norm_o);
                new_origin = vadd(new_origin, vmul(norm_o, bias));
                break;
            case ILM_MAT_REFRACTIVE: {
                R cosI = -vdot(r.d, n);
                R cosT = 0;
                R F = fresnel_dielectric(cosI, m.ior, cosT);
                if (F == (R)1.0f || rnd() < F) {
                    new_dir = vreflect(r.d, n);
                    new_origin = vadd(new_origin, vmul(norm_o, bias));
                } else {
                    R eta = cosI < (R)0 ? m.ior : (R)1.0f / m.ior;
                    medium = cosI > (R)0 ? m.medium : -1;
                    new_dir = vadd(vmul(r.d, eta), vmul(norm_o, eta * std::fabs(cosI) - cosT));
                    new_origin = vadd(new_origin, vmul(norm_o, -bias));
                }
                color = vmulv(color, albedo);
                break;
            }
            case ILM_MAT_ROUGHREFLECTIVE: {
                V3<R> wi = vmul(r.d, (R)-1.0f);
                R wiDotN = vdot(wi, n);
                R a = mf_alpha(m.distribution, roughness, wiDotN);
                R u1 = rnd(), u2 = rnd();
                V3<R> mm = vto_basis(mf_sample(m.distribution, a, u1, u2), n);
                R wiDotT = 0;
                R wiDotM = vdot(wi, mm);
                R F = fresnel_dielectric(wiDotM, m.ior, wiDotT);
                new_dir = vreflect(r.d, mm);
                if (wiDotN * vdot(new_dir, n) <= (R)0) return v3((R)0, (R)0, (R)0);
                new_origin = vadd(new_origin, vmul(norm_o, bias));
                R G = mf_G(m.distribution, wi, new_dir, mm, n, a);
                R weight = (F * G * std::fabs(wiDotM)) / (std::fabs(wiDotN) * std::fabs(vdot(mm, n)));
                color = vmulv(color, vmul(albedo, weight));
                break;
            }
            case ILM_MAT_ROUGHREFRACTIVE: {
                V3<R> wi = vmul(r.d, (R)-1.0f);
                R wiDotN = vdot(wi, n);
                R a = mf_alpha(m.distribution, roughness, wiDotN);
                R u1 = rnd(), u2 = rnd();
                V3<R> mm = vto_basis(mf_sample(m.distribution, a, u1, u2), n);
                R wiDotT = 0;
                R wiDotM = vdot(wi, mm);
                R F = fresnel_dielectric(wiDotM, m.ior, wiDotT);
                if (F == (R)1.0f || rnd() < F) {
                    new_dir = vreflect(r.d, mm);
                    if (wiDotN * vdot(new_dir, n) <= (R)0) return v3((R)0, (R)0, (R)0);
                    new_origin = vadd(new_origin, vmul(norm_o, bias));
                } else {
                    R eta = wiDotM < (R)0 ? m.ior : (R)1.0f / m.ior;
                    medium = wiDotM > (R)0 ? m.medium : -1;
                    new_dir = vsub(vmul(mm, wiDotM * eta - (wiDotM > (R)0 ? (R)1.0f : (R)-1.0f) * wiDotT), vmul(wi, eta));
                    if (wiDotN * vdot(new_dir, n) >= (R)0) return v3((R)0, (R)0, (R)0);
                    new_origin = vadd(new_origin, vmul(norm_o, -bias));
                }
                R G = mf_G(m.distribution, wi, new_dir, mm, n, a);
                R weight = (G * std::fabs(wiDotM)) / (std::fabs(wiDotN) * std::fabs(vdot(mm, n)));
                color = vmulv(color, vmul(albedo, weight));
                break;
            }
            case ILM_MAT_CONDUCTOR: {
                R cosI = -vdot(r.d, n);
                if (cosI <= 0) return v3((R)0, (R)0, (R)0);
                new_dir = vreflect(r.d, norm_o);
                new_origin = vadd(new_origin, vmul(norm_o, bias));
                color = vmulv(color, fresnel_conductor3(m.eta, m.k, cosI));
                break;
            }
            case ILM_MAT_ROUGHCONDUCTOR: {
                V3<R> wi = vmul(r.d, (R)-1.0f);
                R wiDotN = vdot(wi, n);
                if (wiDotN <= 0) return v3((R)0, (R)0, (R)0);
                R a = mf_alpha(m.distribution, roughness, wiDotN);
                R u1 = rnd(), u2 = rnd();
                V3<R> mm = vto_basis(mf_sample(m.distribution, a, u1, u2), n);
                V3<R> F = fresnel_conductor3(m.eta, m.k, wiDotN);
                new_dir = vreflect(r.d, mm);
                if (wiDotN * vdot(new_dir, n) <= (R)0) return v3((R)0, (R)0, (R)0);
                R wiDotM = vdot(wi, mm);
                R G = mf_G(m.distribution, wi, new_dir, mm, n, a);
                R weight = (G * std::fabs(wiDotM)) / (std::fabs(wiDotN) * std::fabs(vdot(mm, n)));
                color = vmulv(color, vmul(F, weight));
                new_origin = vadd(new_origin, vmul(mm, bias));
                break;
            }
            case ILM_MAT_PLASTIC: {
                R cosI = -vdot(r.d, n);
                R cosT = 0;
                R F = fresnel_dielectric(cosI, m.ior, cosT);
                R u1 = rnd();
                if (F == (R)1.0f || u1 < F) {
                    new_dir = vreflect(r.d, n);
                    color = vmulv(color, v3((R)0.99f, (R)0.99f, (R)0.99f));
                } else {
                    R u2 = rnd();
                    R uu = (R)((double)(u1 - F) / (1.0 - (double)F));
                    new_dir = vto_basis(sample_hemisphere_cosine(uu, u2), norm_o);
                    R cosO = vdot(new_dir, n);
                    R Fo = fresnel_dielectric(cosO, m.ior, cosT);
                    V3<R> diff = vdivs(albedo, (R)1.0f - m.diffuse_fresnel);
                    R inv_eta = (R)1.0f / m.ior;
                    color = vmulv(color, vmul(diff, inv_eta * inv_eta * ((R)1 - Fo)));
                }
                new_origin = vadd(new_origin, vmul(norm_o, bias));
                break;
            }
            default:
                return v3((R)0, (R)0, (R)0);
            }
            ray_set(&ray, new_origin, new_dir);
        } else {
            R u = (R)((double)(Mth::atan2(ray.d.z, ray.d.x) * (R)0.5f) * ILLUME_INV_PI + 0.5f);
            R v = (R)(0.5f - (double)Mth::asin(ray.d.y) * ILLUME_INV_PI);
            color = vmulv(color, tex_eval(sc.env, u, v));
            return color;
        }
    }
    return v3((R)0, (R)0, (R)0);   /* alive after max_depth: contributes nothing */
}

/* src/tonemapper.cu:33-76 */
template <class R> static V3<R> uncharted2(V3<R> x) {
    const R A = (R)0.15f, B = (R)0.50f, C = (R)0.10f, D = (R)0.20f, E = (R)0.02f, F = (R)0.30f;
    V3<R> num = vadds(vmulv(x, vadds(vmul(x, A), C * B)), D * E);
    V3<R> den = vadds(vmulv(x, vadds(vmul(x, A), B)), D * F);
    return vadds(vdivv(num, den), -E / F);
}
template <class R> static V3<R> tonemap_eval(int op, V3<R> in) {
    const R gamma = (R)(1.0f / 2.2f);
    switch (op) {
    case ILM_TONEMAP_REINHARD: return vpow(vdivv(in, vadds(in, (R)1.0f)), gamma);
    case ILM_TONEMAP_FILMIC: {
        V3<R> x = vmax(vadds(in, (R)-0.004f), (R)0.0f);
        return vdivv(vmulv(x, vadds(vmul(x, (R)6.2f), (R)0.5f)), vadds(vmulv(x, vadds(vmul(x, (R)6.2f), (R)1.7f)), (R)0.06f));
    }
    case ILM_TONEMAP_UNCHARTED: {
        const R W = (R)11.2f, exposure_bias = (R)2.0f;
        V3<R> curr = uncharted2(vmul(in, exposure_bias));
        V3<R> white = vdivv(v3((R)1.0f, (R)1.0f, (R)1.0f), uncharted2(v3(W, W, W)));
        return vpow(vmulv(curr, white), gamma);
    }
    default: return vpow(in, gamma);
    }
}
/* src/renderer.cu:426-441 */
template <class R> static ilm_pixel tonemap_pixel(int op, V3<R> sum, int spp) {
    R inv = (R)1 / (R)spp;
    V3<R> avg = vmul(sum, inv);
    avg = tonemap_eval(op, avg);
    avg = vmax(vmin(avg, (R)1), (R)0);
    ilm_pixel px;
    px.r = (uint64_t)(int)((R)255 * avg.x);
    px.g = (uint64_t)(int)((R)255 * avg.y);
    px.b = (uint64_t)(int)((R)255 * avg.z);
    return px;
}

/* ---------------- scene construction ---------------- */
template <class R> static Scene<R>* scene_build(const ilm_scene_desc* d, const std::vector<RefBVH>* shared_bvh) {
    Scene<R>* sc = new Scene<R>();
    sc->meshes.resize(d->num_meshes);
    for (int i = 0; i < d->num_meshes; i++) {
        const ilm_mesh_desc& md = d->meshes[i];
        Mesh<R>& m = sc->meshes[i];
        m.pos.resize(md.num_vertices);
        for (int v = 0; v < md.num_vertices; v++)
            m.pos[v] = v3((R)md.positions[3 * v], (R)md.positions[3 * v + 1], (R)md.positions[3 * v + 2]);
        m.face_normals = md.face_normals != 0;
        m.has_texcoords = md.has_texcoords != 0;
        if (!m.face_normals) {
            m.nrm.resize(md.num_vertices);
            for (int v = 0; v < md.num_vertices; v++)
                m.nrm[v] = v3((R)md.normals[3 * v], (R)md.normals[3 * v + 1], (R)md.normals[3 * v + 2]);
        }
        if (m.has_texcoords) m.uv.assign(md.texcoords, md.texcoords + 2 * md.num_vertices);
        m.idx.assign(md.indices, md.indices + 3 * md.num_triangles);
        if (shared_bvh) {
            m.bvh = (*shared_bvh)[i];
        } else {
            /* get_bounds (src/primitives/mesh.cu:342-361) in float, then the reference builder */
            std::vector<AABB<float>> tri_aabbs(md.num_triangles);
            AABB<float> bounds = aabb_empty<float>();
            for (int t = 0; t < md.num_triangles; t++) {
                AABB<float> ta = aabb_empty<float>();
                for (int j = 0; j < 3; j++) {
                    int vi = md.indices[3 * t + j];
                    V3<float> p = v3(md.positions[3 * vi], md.positions[3 * vi + 1], md.positions[3 * vi + 2]);
                    aabb_update(&ta, p);
                    aabb_update(&bounds, p);
                }
                fix_aabb(&ta);
                tri_aabbs[t] = ta;
            }
            fix_aabb(&bounds);
            int leaf = md.leaf_size > 0 ? md.leaf_size : 4;
            m.bvh = ref_bvh_build(tri_aabbs, bounds, leaf);
        }
    }
    sc->mediums.resize(d->num_mediums);
    for (int i = 0; i < d->num_mediums; i++) {
        Medium<R>& m = sc->mediums[i];
        m.absorption = v3((R)d->mediums[i].sigma_a[0], (R)d->mediums[i].sigma_a[1], (R)d->mediums[i].sigma_a[2]);
        m.scattering = (R)d->mediums[i].sigma_s;
        m.g = (R)d->mediums[i].g;
        m.active = ((double)vlen2(m.absorption) > 1e-5 || m.scattering > 0);   /* src/medium.cu:12 */
    }
    sc->mats.resize(d->num_materials);
    for (int i = 0; i < d->num_materials; i++) {
        const ilm_material_desc& md = d->materials[i];
        Material<R>& m = sc->mats[i];
        m.type = md.type;
        m.albedo = tex_from_desc<R>(md.albedo);
        m.roughness = tex_from_desc<R>(md.roughness);
        m.ior = (R)md.ior;
        m.medium = md.medium;
        m.eta = v3((R)md.eta[0], (R)md.eta[1], (R)md.eta[2]);
        m.k = v3((R)md.k[0], (R)md.k[1], (R)md.k[2]);
        m.distribution = md.distribution;
        if (m.type == ILM_MAT_PLASTIC) m.diffuse_fresnel = diffuse_fresnel((R)1.0f / m.ior, 1000000);  /* src/material.cu:176 */
    }
    sc->inst.resize(d->num_instances);
    for (int i = 0; i < d->num_instances; i++) {
        sc->inst[i].mesh = d->instances[i].mesh;
        sc->inst[i].material = d->instances[i].material;
        sc->inst[i].t = xform_from_desc<R>(d->instances[i].transform);
    }
    sc->spheres.resize(d->num_spheres);
    for (int i = 0; i < d->num_spheres; i++) {
        sc->spheres[i].r = (R)d->spheres[i].radius;
        sc->spheres[i].c = v3((R)d->spheres[i].center[0], (R)d->spheres[i].center[1], (R)d->spheres[i].center[2]);
        sc->spheres[i].material = d->spheres[i].material;
    }
    if (d->has_environment) sc->env = tex_from_desc<R>(d->environment);
    else sc->env.type = ILM_TEX_CONSTANT;   /* black constant (src/texture.h:39-43) */
    sc->fov = (R)d->camera.fov;
    sc->dof = (R)d->camera.dof;
    sc->aperture = (R)d->camera.aperture;
    sc->cam = xform_from_desc<R>(d->camera.transform).mat;
    return sc;
}

}  // namespace orc

using namespace orc;

struct orc_scene {
    Scene<float>* f = nullptr;
    Scene<double>* d = nullptr;
};

static void set_threads(int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
}

extern "C" {

int orc_threads_available(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* fp64: 0 = FP32 faithful only, 1 = also build the FP64 twin (same BVH);
 * median_bvh: 1 = the median-split builder instead of the reference sweep builder */
orc_scene* orc_scene_create2(const ilm_scene_desc* desc, int fp64, int threads, int median_bvh) {
    set_threads(threads);
    orc_scene* s = new orc_scene();
    g_median_bvh = median_bvh != 0;
    s->f = scene_build<float>(desc, nullptr);
    g_median_bvh = false;
    if (fp64) {
        std::vector<RefBVH> bvhs;
        for (auto& m : s->f->meshes) bvhs.push_back(m.bvh);
        s->d = scene_build<double>(desc, &bvhs);
    }
    return s;
}

orc_scene* orc_scene_create(const ilm_scene_desc* desc, int fp64, int threads) {
    return orc_scene_create2(desc, fp64, threads, 0);
}

/* 1: closest hits by the build's ordered, t-pruned traversal over the same trees (second CPU
 * baseline figure); 0: the reference's exhaustive walk (the default) */
void orc_scene_set_tpruned(orc_scene* s, int on) {
    if (!s) return;
    if (s->f) s->f->tpruned = on != 0;
    if (s->d) s->d->tpruned = on != 0;
}

void orc_scene_free(orc_scene* s) {
    if (!s) return;
    delete s->f;
    delete s->d;
    delete s;
}

/* Reference BVH of mesh i: returns node count; nodes: n*8 (lo.xyz hi.xyz a b as float bits) */
int orc_ref_bvh_size(orc_scene* s, int mesh, int* num_tri_indices) {
    if (!s || mesh < 0 || mesh >= (int)s->f->meshes.size()) return -1;
    if (num_tri_indices) *num_tri_indices = (int)s->f->meshes[mesh].bvh.tri_indices.size();
    return (int)s->f->meshes[mesh].bvh.nodes.size();
}
int orc_ref_bvh_export(orc_scene* s, int mesh, float* boxes, uint32_t* ab, int32_t* tri_indices) {
    const RefBVH& b = s->f->meshes[mesh].bvh;
    for (size_t i = 0; i < b.nodes.size(); i++) {
        const RefNode& n = b.nodes[i];
        float* o = boxes + 6 * i;
        o[0] = n.bounds.lo.x; o[1] = n.bounds.lo.y; o[2] = n.bounds.lo.z;
        o[3] = n.bounds.hi.x; o[4] = n.bounds.hi.y; o[5] = n.bounds.hi.z;
        ab[2 * i] = n.a; ab[2 * i + 1] = n.b;
    }
    std::memcpy(tri_indices, b.tri_indices.data(), b.tri_indices.size() * 4);
    return 0;
}

/* Closest hit with the reference traversal (exhaustive, FP32). brute != 0 replaces every
 * instance's BVH walk by a loop over all its triangles (pins BVH-independence).
 * counts (nullable): n*2 box and triangle tests per ray. */
int orc_closest_hit(orc_scene* s, const float* rays, int n, int32_t* inst, int32_t* prim, float* t,
                    int brute, int32_t* counts, int threads) {
    set_threads(threads);
    const Scene<float>& sc = *s->f;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; i++) {
        Ray<float> r;
        r.o = v3(rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]);
        r.d = v3(rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]);
        int bt = 0, tt = 0;
        Hit<float> h;
        if (!brute) {
            h = get_min_hit(sc, r, &bt, &tt);
        } else {
            h.d = FLT_MAX;
            for (size_t k = 0; k < sc.spheres.size(); k++) {
                Hit<float> hh;
                sphere_intersect(sc.spheres[k], r, &hh);
                if (hh.is && hh.d < h.d) { h = hh; h.inst = -2 - (int)k; h.prim = -1; }
            }
            for (size_t k = 0; k < sc.inst.size(); k++) {
                const Instance<float>& in = sc.inst[k];
                Ray<float> r2;
                r2.o = m4_mul_v3(in.t.inv, r.o, 1.0f);
                r2.d = m4_mul_v3(in.t.inv, r.d, 0.0f);
                const Mesh<float>& m = sc.meshes[in.mesh];
                float best = FLT_MAX;
                int bp = -1;
                for (size_t q = 0; q < m.idx.size() / 3; q++) {
                    float u, v;
                    float d = tri_intersect(m.pos[m.idx[3 * q]], m.pos[m.idx[3 * q + 1]], m.pos[m.idx[3 * q + 2]], r2, u, v);
                    if (d > 0 && d < best) { best = d; bp = (int)q; }
                }
                if (bp >= 0 && best < h.d) { h.is = true; h.d = best; h.inst = (int)k; h.prim = bp; }
            }
        }
        if (h.is) { inst[i] = h.inst; prim[i] = h.prim; t[i] = h.d; }
        else { inst[i] = -1; prim[i] = -1; t[i] = FLT_MAX; }
        if (counts) { counts[2 * i] = bt; counts[2 * i + 1] = tt; }
    }
    return 0;
}

/* Every hit of one ray under the reference's own tests, nearest first (brute force over every
 * sphere and every instance's triangles; a triangle is tested on the FP32 object-space ray as
 * get_min_hit forms it): inst, prim (-1 for a sphere), FP32 distance, and for a triangle the
 * margin min(u, v, 1 - u - v) of the same Moller-Trumbore test evaluated in FP64 on the same
 * FP32 inputs (+inf for a sphere). A negative margin marks a point outside the triangle that
 * FP32 rounding accepted: whether a BVH walk ever tests that triangle depends on the tree's boxes
 * (the reference tests the triangles of every leaf whose parent box the ray meets), so a walk
 * over another tree may report the next hit instead. Fills at most cap entries; returns the count. */
int orc_ray_hits(orc_scene* s, const float* ray6, int cap, int32_t* inst, int32_t* prim, float* t, double* margin) {
    const Scene<float>& sc = *s->f;
    Ray<float> r;
    r.o = v3(ray6[0], ray6[1], ray6[2]);
    r.d = v3(ray6[3], ray6[4], ray6[5]);
    struct H { int inst, prim; float t; double m; };
    std::vector<H> hits;
    for (size_t k = 0; k < sc.spheres.size(); k++) {
        Hit<float> hh;
        sphere_intersect(sc.spheres[k], r, &hh);
        if (hh.is) hits.push_back({-2 - (int)k, -1, hh.d, HUGE_VAL});
    }
    for (size_t k = 0; k < sc.inst.size(); k++) {
        const Instance<float>& in = sc.inst[k];
        Ray<float> r2;
        r2.o = m4_mul_v3(in.t.inv, r.o, 1.0f);
        r2.d = m4_mul_v3(in.t.inv, r.d, 0.0f);
        const Mesh<float>& m = sc.meshes[in.mesh];
        for (size_t q = 0; q < m.idx.size() / 3; q++) {
            const V3<float> a = m.pos[m.idx[3 * q]], b = m.pos[m.idx[3 * q + 1]], c = m.pos[m.idx[3 * q + 2]];
            float u, v;
            const float d = tri_intersect(a, b, c, r2, u, v);
            if (!(d > 0)) continue;
            const V3<double> A = v3((double)a.x, (double)a.y, (double)a.z);
            const V3<double> e1 = vsub(v3((double)b.x, (double)b.y, (double)b.z), A);
            const V3<double> e2 = vsub(v3((double)c.x, (double)c.y, (double)c.z), A);
            const V3<double> D = v3((double)r2.d.x, (double)r2.d.y, (double)r2.d.z);
            const V3<double> P = vcross(D, e2);
            const double det = vdot(e1, P);
            const V3<double> T = vsub(v3((double)r2.o.x, (double)r2.o.y, (double)r2.o.z), A);
            const double uu = vdot(T, P) / det;
            const double vv = vdot(D, vcross(T, e1)) / det;
            hits.push_back({(int)k, (int)q, d, std::min(std::min(uu, vv), 1.0 - uu - vv)});
        }
    }
    std::stable_sort(hits.begin(), hits.end(), [](const H& x, const H& y) { return x.t < y.t; });
    for (int i = 0; i < (int)hits.size() && i < cap; i++) {
        inst[i] = hits[i].inst; prim[i] = hits[i].prim; t[i] = hits[i].t; margin[i] = hits[i].m;
    }
    return (int)hits.size();
}

/* Radiance of paths (pixel, pass). fp64 selects the twin; rgb is n*3 doubles when fp64 else floats. */
int orc_path_radiance(orc_scene* s, const ilm_render_params* p, const uint32_t* pixel_pass, int n,
                      void* rgb, int32_t* segments, int fp64, int threads) {
    set_threads(threads);
    if (fp64 && !s->d) return ILM_E_STATE;
    if (fp64) {
        RenderInfo<double> ri = make_render_info(*s->d, p->width, p->height);
        double* out = (double*)rgb;
#pragma omp parallel for schedule(dynamic, 16)
        for (int i = 0; i < n; i++) {
            PathStats st;
            V3<double> c = trace_path(*s->d, ri, *p, pixel_pass[2 * i + 1], pixel_pass[2 * i], &st);
            out[3 * i] = c.x; out[3 * i + 1] = c.y; out[3 * i + 2] = c.z;
            if (segments) segments[i] = st.segments;
        }
    } else {
        RenderInfo<float> ri = make_render_info(*s->f, p->width, p->height);
        float* out = (float*)rgb;
#pragma omp parallel for schedule(dynamic, 16)
        for (int i = 0; i < n; i++) {
            PathStats st;
            V3<float> c = trace_path(*s->f, ri, *p, pixel_pass[2 * i + 1], pixel_pass[2 * i], &st);
            out[3 * i] = c.x; out[3 * i + 1] = c.y; out[3 * i + 2] = c.z;
            if (segments) segments[i] = st.segments;
        }
    }
    return 0;
}

/* The segments of one FP32 path: rec holds 9 floats per segment (ray origin, ray direction,
 * closest-hit instance, primitive, distance; -1, -1, FLT_MAX for a miss), at most rec_max of them.
 * Returns the path's segment count. */
int orc_path_segments(orc_scene* s, const ilm_render_params* p, uint32_t pixel, uint32_t pass,
                      float* rec, int rec_max) {
    RenderInfo<float> ri = make_render_info(*s->f, p->width, p->height);
    PathStats st;
    st.rec = rec;
    st.rec_max = rec_max;
    trace_path(*s->f, ri, *p, pass, pixel, &st);
    return st.segments;
}

/* orc_path_segments over n (pixel, pass) pairs on all threads: rec holds rec_max records of 9
 * floats per path (unused records untouched), counts[i] = min(segments of path i, rec_max). */
int orc_path_segments_batch(orc_scene* s, const ilm_render_params* p, const uint32_t* pixel_pass, int n,
                            float* rec, int rec_max, int32_t* counts, int threads) {
    set_threads(threads);
    RenderInfo<float> ri = make_render_info(*s->f, p->width, p->height);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; i++) {
        PathStats st;
        st.rec = rec + (size_t)9 * rec_max * i;
        st.rec_max = rec_max;
        trace_path(*s->f, ri, *p, pixel_pass[2 * i + 1], pixel_pass[2 * i], &st);
        counts[i] = st.segments < rec_max ? st.segments : rec_max;
    }
    return 0;
}

/* Render pixels [pix_begin, pix_end) of the image for passes first_sample..first_sample+spp-1.
 * Per pixel the samples are summed in pass order: plain float sum (kahan=0, the reference's
 * final_colors += c) or Kahan-compensated (kahan=1). accum: (pix_end-pix_begin)*3 floats or
 * doubles (fp64). out (nullable): tonemapped pixels. Returns the elapsed seconds of the loop. */
double orc_render(orc_scene* s, const ilm_render_params* p, int64_t pix_begin, int64_t pix_end, int kahan,
                  int fp64, void* accum, ilm_pixel* out, uint64_t* segments_out, int threads) {
    set_threads(threads);
    if (fp64 && !s->d) return -1.0;
    int64_t npx = pix_end - pix_begin;
    uint64_t segs = 0;
    auto t0 = std::chrono::steady_clock::now();
    if (fp64) {
        RenderInfo<double> ri = make_render_info(*s->d, p->width, p->height);
        double* acc = (double*)accum;
#pragma omp parallel for schedule(dynamic, 8) reduction(+ : segs)
        for (int64_t q = 0; q < npx; q++) {
            uint32_t px = (uint32_t)(pix_begin + q);
            V3<double> sum = v3(0.0, 0.0, 0.0);
            PathStats st;
            for (int i = 0; i < p->spp; i++) {
                V3<double> c = trace_path(*s->d, ri, *p, (uint32_t)(p->first_sample + i), px, &st);
                sum = vadd(sum, c);
            }
            segs += (uint64_t)st.segments;
            if (acc) { acc[3 * q] = sum.x; acc[3 * q + 1] = sum.y; acc[3 * q + 2] = sum.z; }
            if (out) out[q] = tonemap_pixel(p->tonemap, sum, p->spp);
        }
    } else {
        RenderInfo<float> ri = make_render_info(*s->f, p->width, p->height);
        float* acc = (float*)accum;
#pragma omp parallel for schedule(dynamic, 8) reduction(+ : segs)
        for (int64_t q = 0; q < npx; q++) {
            uint32_t px = (uint32_t)(pix_begin + q);
            V3<float> sum = v3(0.f, 0.f, 0.f), comp = v3(0.f, 0.f, 0.f);
            PathStats st;
            for (int i = 0; i < p->spp; i++) {
                V3<float> c = trace_path(*s->f, ri, *p, (uint32_t)(p->first_sample + i), px, &st);
                if (kahan) {
                    float* sp = &sum.x; float* cp = &comp.x; float* xp = &c.x;
                    for (int k = 0; k < 3; k++) {
                        float y = xp[k] - cp[k];
                        float tt = sp[k] + y;
                        cp[k] = (tt - sp[k]) - y;
                        sp[k] = tt;
                    }
                } else {
                    sum = vadd(sum, c);
                }
            }
            segs += (uint64_t)st.segments;
            if (acc) { acc[3 * q] = sum.x; acc[3 * q + 1] = sum.y; acc[3 * q + 2] = sum.z; }
            if (out) out[q] = tonemap_pixel(p->tonemap, sum, p->spp);
        }
    }
    auto t1 = std::chrono::steady_clock::now();
    if (segments_out) *segments_out = segs;
    return std::chrono::duration<double>(t1 - t0).count();
}

/* Tonemap an accumulated image (sum over spp) exactly as src/renderer.cu:426-441 */
int orc_tonemap_image(const float* accum, int64_t npx, int spp, int op, ilm_pixel* out) {
    for (int64_t i = 0; i < npx; i++)
        out[i] = tonemap_pixel(op, v3(accum[3 * i], accum[3 * i + 1], accum[3 * i + 2]), spp);
    return 0;
}

/* ---- unit exports for closed-form tests ---- */
float orc_fresnel_dielectric(float cosI, float ior, float* cosT) { float c; float f = fresnel_dielectric(cosI, ior, c); if (cosT) *cosT = c; return f; }
double orc_fresnel_dielectric64(double cosI, double ior) { double c; return fresnel_dielectric(cosI, ior, c); }
float orc_fresnel_conductor(float eta, float k, float cosI) { return fresnel_conductor(eta, k, cosI); }
float orc_diffuse_fresnel(float ior, int samples) { return diffuse_fresnel(ior, samples); }
float orc_triangle_intersect(const float* a, const float* b, const float* c, const float* o, const float* d, float* u, float* v) {
    Ray<float> r{v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2])};
    float uu = 0, vv = 0;
    float t = tri_intersect(v3(a[0], a[1], a[2]), v3(b[0], b[1], b[2]), v3(c[0], c[1], c[2]), r, uu, vv);
    if (u) *u = uu;
    if (v) *v = vv;
    return t;
}
float orc_aabb_intersect(const float* lo, const float* hi, const float* o, const float* d) {
    AABB<float> b{v3(lo[0], lo[1], lo[2]), v3(hi[0], hi[1], hi[2])};
    Ray<float> r{v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2])};
    return aabb_ray_intersect(b, r);
}
int orc_sphere_intersect(const float* c, float radius, const float* o, const float* d, float* t, float* normal, float* uv) {
    Sphere<float> s{radius, v3(c[0], c[1], c[2]), 0};
    Ray<float> r{v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2])};
    Hit<float> h;
    sphere_intersect(s, r, &h);
    if (!h.is) return 0;
    *t = h.d;
    normal[0] = h.normal.x; normal[1] = h.normal.y; normal[2] = h.normal.z;
    uv[0] = h.u; uv[1] = h.v;
    return 1;
}
void orc_tonemap_eval(int op, const float* in, float* out) {
    V3<float> r = tonemap_eval(op, v3(in[0], in[1], in[2]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
uint32_t orc_wang_hash(uint32_t a) { return ilm_wang_hash(a); }
/* which: 0 cosine hemisphere, 1 circle, 2 sphere, 3 HG(g), 4 Beckmann(a), 5 GGX(a) */
void orc_sample(int which, float param, const float* u, int n, float* out) {
    for (int i = 0; i < n; i++) {
        float u1 = u[2 * i], u2 = u[2 * i + 1];
        V3<float> r;
        switch (which) {
        case 0: r = sample_hemisphere_cosine(u1, u2); break;
        case 1: r = sample_circle(u1, u2); break;
        case 2: r = sample_sphere(u1, u2); break;
        case 3: r = sample_hg(param, u1, u2); break;
        case 4: r = sample_beckmann(param, u1, u2); break;
        default: r = sample_ggx(param, u1, u2); break;
        }
        out[3 * i] = r.x; out[3 * i + 1] = r.y; out[3 * i + 2] = r.z;
    }
}
float orc_g1_beckmann(const float* v, const float* n, float alpha) { return G1_beckmann(v3(v[0], v[1], v[2]), v3(n[0], n[1], n[2]), alpha); }
float orc_g1_ggx(const float* v, const float* m, const float* n, float alpha) {
    return G1_ggx(v3(v[0], v[1], v[2]), v3(m[0], m[1], m[2]), v3(n[0], n[1], n[2]), alpha);
}
void orc_to_basis(const float* v, const float* n, float* out) {
    V3<float> r = vto_basis(v3(v[0], v[1], v[2]), v3(n[0], n[1], n[2]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void orc_transform_matrices(const ilm_transform_desc* t, float* out48, double* out48_64) {
    Xform<float> x = xform_from_desc<float>(*t);
    for (int i = 0; i < 16; i++) {
        out48[i] = x.mat.m[i / 4][i % 4];
        out48[16 + i] = x.inv.m[i / 4][i % 4];
        out48[32 + i] = x.trans_inv.m[i / 4][i % 4];
    }
    if (out48_64) {
        Xform<double> y = xform_from_desc<double>(*t);
        for (int i = 0; i < 16; i++) {
            out48_64[i] = y.mat.m[i / 4][i % 4];
            out48_64[16 + i] = y.inv.m[i / 4][i % 4];
            out48_64[32 + i] = y.trans_inv.m[i / 4][i % 4];
        }
    }
}
/* pixel_size, left, top of src/renderer.cu:443-458 */
void orc_render_info(const ilm_camera_desc* c, int width, int height, float* out3) {
    Scene<float> sc;
    sc.fov = c->fov; sc.dof = c->dof; sc.aperture = c->aperture;
    RenderInfo<float> ri = make_render_info(sc, width, height);
    out3[0] = ri.pixel_size; out3[1] = ri.left; out3[2] = ri.top;
}
float orc_texture_eval(const ilm_texture_desc* d, float u, float v, float* out3) {
    Tex<float> t = tex_from_desc<float>(*d);
    V3<float> r = tex_eval(t, u, v);
    out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
    return r.x;
}
/* pmath and PCG exports, so the tests can pin them against double libm / published sequences */
void orc_pmath(int fn, const float* x, const float* y, int n, float* out) {
    for (int i = 0; i < n; i++) {
        switch (fn) {
        case 0: out[i] = pm_sinf(x[i]); break;
        case 1: out[i] = pm_cosf(x[i]); break;
        case 2: out[i] = pm_expf(x[i]); break;
        case 3: out[i] = pm_logf(x[i]); break;
        case 4: out[i] = pm_atanf(x[i]); break;
        case 5: out[i] = pm_atan2f(y[i], x[i]); break;
        case 6: out[i] = pm_asinf(x[i]); break;
        default: out[i] = pm_powf(x[i], y[i]); break;
        }
    }
}
void orc_pcg32(uint64_t initstate, uint64_t initseq, int n, uint32_t* out) {
    uint64_t s = ilm_pcg32_seed(initstate, initseq);
    uint64_t inc = ilm_pcg32_inc(initseq);
    for (int i = 0; i < n; i++) out[i] = ilm_pcg32_next(&s, inc);
}
void orc_path_uniforms(uint32_t pass, uint32_t pixel, int n, float* out) {
    uint64_t s = ilm_path_seed(pass, pixel);
    uint64_t inc = ilm_pcg32_inc(pixel);
    for (int i = 0; i < n; i++) out[i] = ilm_pcg32_uniform(&s, inc);
}

}  /* extern "C" */
