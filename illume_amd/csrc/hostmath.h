/*
 * hostmath.h - host-side scene math of the reference, restated for the product's scene upload:
 * the transform builder (src/math/transform.cu:6-44 with src/math/matrix4.cu), the camera's
 * RenderInfo (src/renderer.cu:443-458) and the plastic diffuse-Fresnel integral
 * (src/fresnel.h:59-73). Same float operation order as the reference, so the matrices and
 * constants the device uses equal the reference's bit for bit.
 */
#pragma once
#include <cmath>
#include <cstdint>

#include "../../include/illume_c.h"

namespace hm {

static const double PI_D = 3.14159265358979323846;   /* src/math/mathutils.h:6 */

struct M4 { float m[4][4]; };
struct F3 { float x, y, z; };

inline F3 f3(float x, float y, float z) { F3 r; r.x = x; r.y = y; r.z = z; return r; }
inline F3 sub(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline F3 mul(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
inline F3 cross(F3 a, F3 b) { return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline float len(F3 v) { return std::sqrt((v.x * v.x) + (v.y * v.y) + (v.z * v.z)); }
inline F3 normalized(F3 v) { float inv = 1.0f / len(v); return f3(v.x * inv, v.y * inv, v.z * inv); }

inline M4 identity() {
    M4 a;
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) a.m[i][j] = (i == j) ? 1.0f : 0.0f;
    return a;
}
inline M4 mul(const M4& a, const M4& b) {
    M4 r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j] + a.m[i][3] * b.m[3][j];
    return r;
}
inline M4 transpose(const M4& a) {
    M4 t;
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) t.m[j][i] = a.m[i][j];
    return t;
}
/* src/math/matrix4.cu:16-34 */
inline M4 axis_angle(F3 axis, float angle) {
    axis = [](F3 v) { float inv = 1.0f / len(v); return f3(v.x * inv, v.y * inv, v.z * inv); }(axis);
    F3 v = mul(axis, sinf(angle / 2));
    float w = cosf(angle / 2);
    M4 m = identity();
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
/* src/math/matrix4.cu:55-67 */
inline M4 lookat(F3 pos, F3 fwd, F3 up) {
    F3 f = normalized(fwd);
    F3 r = normalized(cross(up, f));
    F3 u = normalized(cross(f, r));
    M4 m;
    m.m[0][0] = r.x; m.m[0][1] = u.x; m.m[0][2] = f.x; m.m[0][3] = pos.x;
    m.m[1][0] = r.y; m.m[1][1] = u.y; m.m[1][2] = f.y; m.m[1][3] = pos.y;
    m.m[2][0] = r.z; m.m[2][1] = u.z; m.m[2][2] = f.z; m.m[2][3] = pos.z;
    m.m[3][0] = 0; m.m[3][1] = 0; m.m[3][2] = 0; m.m[3][3] = 1;
    return m;
}
/* src/math/matrix4.cu:122-138 */
inline float subdet(const M4& m, int i, int j) {
    static const int S[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    float d[3][3];
    for (int y = 0; y < 3; y++) for (int x = 0; x < 3; x++) d[y][x] = m.m[S[i][y]][S[j][x]];
    return (d[0][0] * d[1][1] * d[2][2]) + (d[0][1] * d[1][2] * d[2][0]) + (d[0][2] * d[1][0] * d[2][1])
         - (d[0][2] * d[1][1] * d[2][0]) - (d[0][1] * d[1][0] * d[2][2]) - (d[0][0] * d[1][2] * d[2][1]);
}
/* src/math/matrix4.cu:174-201, including its m[0][3]*m[2][0]*m[2][2]*m[3][1] term (zero for
 * affine matrices) */
inline float det(const M4& M) {
    const float (*m)[4] = M.m;
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
/* src/math/matrix4.cu:115-120,140-172: adjoint times 1/det */
inline M4 inverse(const M4& m) {
    M4 ad;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            ad.m[i][j] = ((i + j) % 2 == 0) ? subdet(m, j, i) : -1 * subdet(m, j, i);
    float s = 1 / det(m);
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) ad.m[i][j] *= s;
    return ad;
}

/* src/math/transform.cu:6-44 */
inline void build_transform(const ilm_transform_desc& t, M4& mat, M4& inv, M4& tinv) {
    if (t.has_matrix) {
        for (int i = 0; i < 16; i++) mat.m[i / 4][i % 4] = t.matrix[i];
    } else if (t.has_lookat) {
        F3 tr = f3(t.translation[0], t.translation[1], t.translation[2]);
        F3 la = f3(t.lookat[0], t.lookat[1], t.lookat[2]);
        mat = lookat(tr, sub(la, tr), f3(0, 1, 0));
    } else {
        /* degtorad (src/math/mathutils.h:9-12): deg * ILLUME_PI / 180.0f in double, as float */
        float rx = (float)((double)t.rotation[0] * PI_D / 180.0f);
        float ry = (float)((double)t.rotation[1] * PI_D / 180.0f);
        float rz = (float)((double)t.rotation[2] * PI_D / 180.0f);
        M4 x = axis_angle(f3(1, 0, 0), rx);
        M4 y = axis_angle(f3(0, 1, 0), ry);
        M4 z = axis_angle(f3(0, 0, 1), rz);
        M4 rot = mul(mul(x, y), z);
        M4 tr = identity();
        tr.m[0][3] = t.translation[0]; tr.m[1][3] = t.translation[1]; tr.m[2][3] = t.translation[2];
        M4 sc = identity();
        sc.m[0][0] = t.scale[0]; sc.m[1][1] = t.scale[1]; sc.m[2][2] = t.scale[2];
        mat = mul(mul(tr, rot), sc);
    }
    inv = inverse(mat);
    tinv = transpose(inv);
}

/* src/renderer.cu:443-458 */
inline void render_info(const ilm_camera_desc& c, int width, int height, float& image_width, float& pixel_size,
                        float& left, float& top) {
    image_width = (float)width;
    float dim_ratio = (float)height / (float)width;
    float tan_half_fov = tanf((float)(PI_D * (double)c.fov / 360));
    float dofmfov = c.dof * tan_half_fov;
    pixel_size = dofmfov * 2 / (float)width;
    left = -1 * dofmfov;
    top = dim_ratio * dofmfov;
}

/* src/fresnel.h:10-26 */
inline float fresnel_dielectric(float cosI, float ior, float& cosT) {
    float eta = cosI < 0.0f ? ior : 1.0f / ior;
    float sinTSq = eta * eta * (1.0f - cosI * cosI);
    if (sinTSq >= 1.0f) { cosT = 0.0f; return 1.0f; }
    cosI = std::fabs(cosI);
    cosT = std::sqrt(1.0f - sinTSq);
    float rperp = (eta * cosI - cosT) / (eta * cosI + cosT);
    float rpar = (cosI - eta * cosT) / (cosI + eta * cosT);
    return (rperp * rperp + rpar * rpar) * 0.5f;
}
/* src/fresnel.h:59-73 with the reference's argument order */
inline float diffuse_fresnel(float ior, int samples) {
    float cosT;
    double ret = 0.0;
    float fb = fresnel_dielectric(0.0f, ior, cosT);
    for (int i = 1; i <= samples; i++) {
        float cosTSq = (float)i / samples;
        float fa = fresnel_dielectric(ior, std::sqrt(cosTSq), cosT);
        ret += (double)(fa + fb) * (0.5 / samples);
        fb = fa;
    }
    return (float)ret;
}

}  // namespace hm
