/*
 * traverse.h - closest-hit traversal and the extend kernels.
 *
 * Two-level walk: spheres in index order (src/renderer.cu:117-125), then a TLAS over instance
 * world boxes whose leaves enter a mesh BLAS in object space (ray transformed by the
 * instance's inverse, direction not renormalised, src/primitives/mesh_instance.cu:96-101).
 * Triangles use Moller-Trumbore exactly as src/primitives/mesh.cu:446-476 (or the watertight
 * test). Unlike the reference's exhaustive walk (src/primitives/mesh.cu:492-530) the walk is
 * ordered (near child first) and t-pruned; box tests are conservative by a relative margin, so
 * the closest hit equals the reference's except for distance ties.
 *
 * The stack lives in LDS, 64 lanes interleaved (entry s of lane l at s*64 + l, one bank per
 * lane); scenes whose proven stack bound exceeds the LDS depth use the spilling variant
 * (OVF = 1) whose extra entries go to a per-lane global area.
 */
#pragma once
#include "shade.h"

#define EXT_BLOCK 64            /* one wave per workgroup: the LDS stack belongs to one wave */
#ifndef LDS_STACK
#define LDS_STACK 20            /* stack entries per lane kept in LDS (5 KB per wave) */
#endif
#define OVF_STACK (128 - LDS_STACK)  /* spill entries per lane (spilling variant) */
#define STACK_SENTINEL ((int)0x7ffffffe)
/* Minimum resident one-wave workgroups per CU the register allocation must allow (0: no
 * constraint); with LDS_STACK it sets the extend kernels' occupancy. */
#ifndef EXT_MIN_BLOCKS
#define EXT_MIN_BLOCKS 28       /* 7 waves per SIMD: registers capped at 72 */
#endif
#if EXT_MIN_BLOCKS > 0
#define EXT_BOUNDS __launch_bounds__(EXT_BLOCK) __attribute__((amdgpu_waves_per_eu(EXT_MIN_BLOCKS / 4, 8)))
#else
#define EXT_BOUNDS __launch_bounds__(EXT_BLOCK)
#endif

/* Conservative slab-test margins (relative); they cover the rcp and fma rounding of the box
 * test. A child is hit when tn * BOX_LO <= tf * BOX_HI and tn * BOX_LO <= tmax; the tests fold
 * the two factors into one: tn <= tf * BOX_K and tn <= tmax * BOX_TMAX, with BOX_K and BOX_TMAX
 * the floats just above BOX_HI / BOX_LO and 1 / BOX_LO (0x3f8000a9 and 0x3f800055), so they
 * accept every box the unfolded form accepts. */
#define BOX_LO 0.99999f
#define BOX_HI 1.00001f
#define BOX_K 1.0000201f
#define BOX_TMAX 1.0000101f

__device__ inline uint32_t f2u(float f) { return __float_as_uint(f); }
__device__ inline float u2f(uint32_t u) { return __uint_as_float(u); }
/* values the caller knows to be equal in every lane, made uniform (SGPRs) for the compiler */
__device__ inline int uni_i(int v) { return __builtin_amdgcn_readfirstlane(v); }
template <class T>
__device__ inline const GLOBAL T* uni_p(const GLOBAL T* p) {
    const uint64_t b = (uint64_t)p;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)b);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(b >> 32));
    return (const GLOBAL T*)(((uint64_t)hi << 32) | lo);
}

__device__ inline uint32_t udiv(uint32_t n, UDiv m) {
    const uint32_t t = __umulhi(n, m.lo);
    const uint32_t q = (uint32_t)(((uint64_t)n * m.hi + t) >> 32);
    return (m.lo | m.hi) ? q : n;
}
/* slice_decode for a work item, with its precomputed divisor (sgroup is a power of two) */
__device__ inline void slice_decode_w(uint32_t l, const GLOBAL DWork* W, uint32_t& k, uint32_t& q) {
    const uint32_t sh = W->sg_log2, span = W->npx << sh;
    const uint32_t kb = udiv(l, W->span_div), r = l - kb * span;
    q = r >> sh;
    k = (kb << sh) + (r - (q << sh));
}
/* tile_pixel for a request of the chunk, with its precomputed divisors */
__device__ inline uint32_t tile_pixel_r(uint32_t i, const GLOBAL DReq* R) {
    const uint32_t th = R->tile_th;
    if (th <= 1 || i >= R->tile_full) return i;
    const uint32_t w = (uint32_t)R->width, band_px = th * w;
    const uint32_t band = udiv(i, R->band_div), r = i - band * band_px;
    const uint32_t x = udiv(r, R->th_div), y = band * th + (r - x * th);
    return y * w + x;
}

/* The work item of the chunk that holds path id i (work items sorted by pid_base). */
__device__ inline int find_work_pid(const GLOBAL DWork* work, int n, uint32_t i) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (work[mid].pid_base <= i) lo = mid; else hi = mid - 1;
    }
    return lo;
}


struct Best {
    float t, u, v;
    int inst, prim;
};

__device__ inline void best_init(Best& B) {
    B.t = ILM_FLT_MAX; B.u = 0.0f; B.v = 0.0f; B.inst = -1; B.prim = -1;
}
/* record a closer triangle hit (A.w carries the original triangle index); shading forms the
 * face normal from the mesh's own vertex arrays */
__device__ inline void best_tri(Best& B, float t, float u, float v, int inst, float4 A) {
    B.t = t; B.u = u; B.v = v; B.inst = inst; B.prim = __float_as_int(A.w);
}
__device__ inline void write_hit(const KParams& P, uint32_t i, const Best& B) {
    gpw(P.hits.h)[i] = make_float4(B.t, B.u, B.v, __int_as_float(B.prim));
    gpw(P.hits.inst)[i] = B.inst;
}

/* Reciprocal direction for the slab test, clamped to [-1e30, 1e30] (one med3 after the rcp):
 * components below about 1e-30 in magnitude (axis-parallel rays, +-0 included, the sign kept)
 * get +-1e30. With a true 1/0 a plane through the origin gives 0 * inf = NaN and a box the ray
 * crosses could be rejected; with 1e30 the slab distances (p - o) / d are +-huge (or exactly 0),
 * which keeps parallel slabs exact and the products finite for |p - o| < 3e8. Only the box test
 * uses this; triangles and spheres see the true direction. */
__device__ inline float safe_rcp(float x) { return __builtin_amdgcn_fmed3f(__builtin_amdgcn_rcpf(x), -1e30f, 1e30f); }
__device__ inline V3 rcp3(V3 d) { return mk3(safe_rcp(d.x), safe_rcp(d.y), safe_rcp(d.z)); }

/* Two child boxes of an FP32 BVH2 node against the ray (o, 1/d). A slab distance is formed as
 * (p - o) * (1/d): each of the two roundings and v_rcp_f32's error is relative to the distance
 * itself, so the BOX_ margins cover them wherever the ray starts (the fma form p/d - o/d carries
 * the rounding of o/d, an absolute error of |o/d| * 2^-24 that a ray far from the world origin
 * meeting a box at small distance would not be covered for). t0, t1: entry distances, used to
 * order the children. */
__device__ inline void box2(const float4 a, const float4 b, const float4 c, V3 id, V3 o, float tmax,
                            float& t0, float& t1, bool& h0, bool& h1) {
    const float tm = tmax * BOX_TMAX;
    float x0 = (a.x - o.x) * id.x, x1 = (a.y - o.x) * id.x;
    float y0 = (a.z - o.y) * id.y, y1 = (a.w - o.y) * id.y;
    float z0 = (b.x - o.z) * id.z, z1 = (b.y - o.z) * id.z;
    float tn = __builtin_fmaxf(__builtin_fmaxf(__builtin_fminf(x0, x1), __builtin_fminf(y0, y1)),
                               __builtin_fmaxf(__builtin_fminf(z0, z1), 0.0f));
    float tf = __builtin_fminf(__builtin_fminf(__builtin_fmaxf(x0, x1), __builtin_fmaxf(y0, y1)),
                               __builtin_fmaxf(z0, z1));
    t0 = tn;
    h0 = (tn <= tf * BOX_K) && (tn <= tm);

    float X0 = (b.z - o.x) * id.x, X1 = (b.w - o.x) * id.x;
    float Y0 = (c.x - o.y) * id.y, Y1 = (c.y - o.y) * id.y;
    float Z0 = (c.z - o.z) * id.z, Z1 = (c.w - o.z) * id.z;
    float Tn = __builtin_fmaxf(__builtin_fmaxf(__builtin_fminf(X0, X1), __builtin_fminf(Y0, Y1)),
                               __builtin_fmaxf(__builtin_fminf(Z0, Z1), 0.0f));
    float Tf = __builtin_fminf(__builtin_fminf(__builtin_fmaxf(X0, X1), __builtin_fmaxf(Y0, Y1)),
                               __builtin_fmaxf(Z0, Z1));
    t1 = Tn;
    h1 = (Tn <= Tf * BOX_K) && (Tn <= tm);
}

/* Up to four dequantised child boxes of a QNode (dequantisation as bvh.cpp's qnode_child_box):
 * bit k of hitmask is set when child k is hit, and tn_out[k] is its entry distance, or +inf for a
 * child missed (or absent), so the distances sort the hit children first. */
/* w0 = ox oy oz | ex ey ez nchild, w1 = qlo[x][0..3] qlo[y] qlo[z] qhi[x], w2 = qhi[y] qhi[z] */
__device__ inline void box4q(const float4 w0, const float4 w1, const float4 w2, V3 id, V3 o, float tmax,
                             float* tn_out, uint32_t& hitmask, int& nchild) {
    const uint32_t e = f2u(w0.w);
    const float sx = u2f((e & 0xffu) << 23), sy = u2f(((e >> 8) & 0xffu) << 23), sz = u2f(((e >> 16) & 0xffu) << 23);
    nchild = (int)(e >> 24);
    /* slab distance of plane q: (o_n + q*s - o) / d = q * (s/d) + (o_n - o)/d. s is a power of
     * two, so A = s * (1/d) is exact while it is a normal float, and a plane costs one fma on top
     * of the byte conversion. The node origin and the ray origin are differenced first, so no
     * error grows with the distance of either from the world origin: with B = (o_n - o) * (1/d)
     * rounded twice, a plane's error is relative to its distance (covered by the BOX_ margins)
     * plus at most 2^-23 * 255 * |A|, which the builder's outward pad of every child box
     * (bvh.cpp QPAD_REL, 2^-20 of the node extent) covers four times over. |1/d| <= 1e30
     * (safe_rcp) and s <= 2^ILM_QUANT_MAX_EXP (scenes that need a larger quantum do not get the
     * 8-bit layout) keep A below 2^124; a subnormal A only adds an absolute error below 2^-140. */
    const float ax = sx * id.x, ay = sy * id.y, az = sz * id.z;
    const float bx = (w0.x - o.x) * id.x, by = (w0.y - o.y) * id.y, bz = (w0.z - o.z) * id.z;
    /* q -> fma(q, A, B) is non-decreasing for A >= 0 and non-increasing for A < 0 (rounding is
     * monotonic), so of a child's two planes on an axis the lo plane is the near one exactly when
     * A >= 0: picking each axis's near and far byte words once per visit gives the same tn and tf
     * as the min and max of the two planes, for two fewer operations per axis and child. */
    const uint32_t lx = f2u(w1.x), ly = f2u(w1.y), lz = f2u(w1.z);
    const uint32_t hx = f2u(w1.w), hy = f2u(w2.x), hz = f2u(w2.y);
    const bool px = f2u(ax) >> 31 == 0u, py = f2u(ay) >> 31 == 0u, pz = f2u(az) >> 31 == 0u;
    const uint32_t nx = px ? lx : hx, fx = px ? hx : lx;
    const uint32_t ny = py ? ly : hy, fy = py ? hy : ly;
    const uint32_t nz = pz ? lz : hz, fz = pz ? hz : lz;
    const float tm = tmax * BOX_TMAX;
    hitmask = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int sh = 8 * k;
        const float xn = __builtin_fmaf((float)((nx >> sh) & 0xffu), ax, bx);
        const float xf = __builtin_fmaf((float)((fx >> sh) & 0xffu), ax, bx);
        const float yn = __builtin_fmaf((float)((ny >> sh) & 0xffu), ay, by);
        const float yf = __builtin_fmaf((float)((fy >> sh) & 0xffu), ay, by);
        const float zn = __builtin_fmaf((float)((nz >> sh) & 0xffu), az, bz);
        const float zf = __builtin_fmaf((float)((fz >> sh) & 0xffu), az, bz);
        const float tn = __builtin_fmaxf(__builtin_fmaxf(xn, yn), __builtin_fmaxf(zn, 0.0f));
        const float tf = __builtin_fminf(__builtin_fminf(xf, yf), zf);
        /* evaluated without short-circuit: a branch per child would split the walk into
         * exec-masked blocks (and make the compiler re-canonicalise the slab values in each).
         * Every node has at least two children (the builder collapses BVH2 nodes; cached trees
         * are checked on load), so only slots 2 and 3 test nchild. */
        const bool hit = (k < 2 || k < nchild) & (tn <= __builtin_fminf(tf * BOX_K, tm));
        tn_out[k] = hit ? tn : __builtin_inff();
        hitmask |= (uint32_t)hit << k;
    }
}

template <int OVF>
struct Stack {
    LDSAS int* lds;     /* &stk[lane]; entry s at lds[s * EXT_BLOCK] */
    GLOBAL int* ovf;    /* the workgroup's spill area (OVF = 1), OVF_STACK entries per lane; uniform,
                           so the lane's own slot is formed where it is used (the spill path is rare) */
    int sp;
    __device__ inline GLOBAL int* spill(int s) const { return ovf + __lane_id() * OVF_STACK + (s - LDS_STACK); }
    __device__ inline void push(int v) {
        if (!OVF) {
            if (sp < LDS_STACK) { lds[sp * EXT_BLOCK] = v; sp++; }
        } else {
            if (sp < LDS_STACK) lds[sp * EXT_BLOCK] = v;
            else if (sp < LDS_STACK + OVF_STACK) *spill(sp) = v;
            sp = sp < LDS_STACK + OVF_STACK ? sp + 1 : sp;
        }
    }
    /* Push the sorted far children of a node visit that hit nhit (1..4) children: c3 if nhit > 3,
     * then c2 if nhit > 2, then c1 if nhit > 1. While three more entries fit in LDS the lane
     * writes rows sp..sp+2 unconditionally (rows past the new top are not live) and moves sp
     * once, instead of three masked pushes. */
    __device__ inline void push_far(int nhit, int c1, int c2, int c3) {
        if (sp <= LDS_STACK - 3) {
            const int e0 = nhit > 3 ? c3 : nhit > 2 ? c2 : c1;
            const int e1 = nhit > 3 ? c2 : c1;
            LDSAS int* p = lds + sp * EXT_BLOCK;
            p[0] = e0;
            p[EXT_BLOCK] = e1;
            p[2 * EXT_BLOCK] = c1;
            sp += nhit - 1;
        } else {
            if (nhit > 3) push(c3);
            if (nhit > 2) push(c2);
            if (nhit > 1) push(c1);
        }
    }
    __device__ inline int pop() {
        --sp;
        if (!OVF || sp < LDS_STACK) return lds[sp * EXT_BLOCK];
        return *spill(sp);
    }
};

struct TravCount {
    uint32_t boxes, tris;
};

__device__ inline void load_rows12(const GLOBAL float* src, float* dst) {
    const GLOBAL float4* s4 = (const GLOBAL float4*)src;
    float4 a = s4[0], b = s4[1], c = s4[2];
    dst[0] = a.x; dst[1] = a.y; dst[2] = a.z; dst[3] = a.w;
    dst[4] = b.x; dst[5] = b.y; dst[6] = b.z; dst[7] = b.w;
    dst[8] = c.x; dst[9] = c.y; dst[10] = c.z; dst[11] = c.w;
}

/* Entering instance k (a TLAS leaf's `first`): its inverse rows, BLAS pointers and root come
 * from its one DInst record (bytes 0..47 and 112..143). All five 16-byte words are requested
 * before the root is tested: the empty asm takes one value of each, so the compiler cannot sink
 * the matrix and pointer loads behind the test and make an entry wait on memory twice. Returns
 * the BLAS root (ILM_NO_CHILD for an empty mesh). */
template <int QUANT>
__device__ inline int inst_load(const GLOBAL DInst* I, float* inv, const GLOBAL BNode*& nodes,
                                const GLOBAL QNode*& qnodes, const GLOBAL float4*& tris) {
    const GLOBAL int4* w = (const GLOBAL int4*)I;
    const int4 r0 = w[0], r1 = w[1], r2 = w[2], p0 = w[7], p1 = w[8];
    asm volatile("" ::"v"(r0.x), "v"(r1.x), "v"(r2.x), "v"(p0.x), "v"(p1.x));
    inv[0] = __int_as_float(r0.x); inv[1] = __int_as_float(r0.y); inv[2] = __int_as_float(r0.z); inv[3] = __int_as_float(r0.w);
    inv[4] = __int_as_float(r1.x); inv[5] = __int_as_float(r1.y); inv[6] = __int_as_float(r1.z); inv[7] = __int_as_float(r1.w);
    inv[8] = __int_as_float(r2.x); inv[9] = __int_as_float(r2.y); inv[10] = __int_as_float(r2.z); inv[11] = __int_as_float(r2.w);
    nodes = (const GLOBAL BNode*)(((uint64_t)(uint32_t)p0.y << 32) | (uint32_t)p0.x);
    qnodes = (const GLOBAL QNode*)(((uint64_t)(uint32_t)p0.w << 32) | (uint32_t)p0.z);
    tris = (const GLOBAL float4*)(((uint64_t)(uint32_t)p1.y << 32) | (uint32_t)p1.x);
    return QUANT ? p1.w : p1.z;
}

/* Camera ray of path id i, src/renderer.cu:81-110 with the per-pass seeding of :61-69: origin
 * (w = bits of req << 16: no medium), unit direction (w = bits of the pixel) and the PCG32
 * state after the ray's draws. A function of the chunk descriptor and i alone, so the first
 * depth's extend and shade launches form it where they need it instead of reading it back. */
__device__ inline void camera_ray(const KParams& P, uint32_t i, float4& o4, float4& d4, uint2& rg) {
    const GLOBAL DChunk* C = gp(P.chunk);
    int req;
    uint32_t pixel, pass;
    if (P.list) {
        req = 0;
        pixel = gp(P.list)[2 * i];
        pass = gp(P.list)[2 * i + 1];
    } else {
        const GLOBAL DWork* W = C->work + find_work_pid(C->work, C->n_work, i);
        uint32_t k, q;
        slice_decode_w(i - W->pid_base, W, k, q);
        req = W->req;
        pass = (uint32_t)W->pass0 + k;
        pixel = W->px0 + q;
    }
    const GLOBAL DReq* R = C->reqs + req;
    if (!P.list) pixel = tile_pixel_r(pixel, R);
    const GLOBAL DScene* S = gp(R->scene);
    uint64_t rs = ilm_path_seed(pass, pixel);
    const uint64_t inc = ilm_pcg32_inc(pixel);
    const float pixel_size = R->pixel_size;
    const uint32_t yq = udiv(pixel, R->w_div);
    int x = (int)(pixel - yq * (uint32_t)R->width), y = (int)yq;
    float left_edge = R->left + pixel_size * (float)x;
    float top_edge = R->top - pixel_size * (float)y;
    float rx = left_edge + pixel_size * ilm_pcg32_uniform(&rs, inc);
    float ry = top_edge - pixel_size * ilm_pcg32_uniform(&rs, inc);
    V3 pos = mk3(0.0f, 0.0f, 0.0f);
    const float aperture = S->aperture;
    if (aperture > ILM_FLT_EPSILON) {
        float u1 = ilm_pcg32_uniform(&rs, inc);
        float u2 = ilm_pcg32_uniform(&rs, inc);
        pos = mul3(sample_circle(u1, u2), aperture);
    }
    float cam[12];
    load_rows12(S->cam, cam);
    V3 origin = xform3(cam, pos, 1.0f);
    V3 image_pos = xform3(cam, mk3(rx, ry, S->dof), 1.0f);
    V3 dir = normalize3(sub3(image_pos, origin));
    o4 = make_float4(origin.x, origin.y, origin.z, u2f((uint32_t)req << 16));
    d4 = make_float4(dir.x, dir.y, dir.z, u2f(pixel));
    rg = make_uint2((uint32_t)rs, (uint32_t)(rs >> 32));
}

/* camera_ray for a chunk of one work item (a whole plain render is one): the item, its request's
 * pixel walk and image plane, and the scene's camera are the same for every path. Their
 * addresses are kept once per wave (CamSrc, three uniform pointers) and the values are fetched
 * with scalar loads where a ray is formed (camu_fetch), so they do not hold some thirty SGPRs
 * across the kernel's loop, which made the compiler spill SGPRs into VGPR lanes (a vector
 * instruction per spill and per reload); same operations, same results as camera_ray. */
struct CamU {
    uint32_t pid_base, px0, npx, sg_log2;
    int pass0;
    UDiv span_div;
    int req, width;
    float pixel_size, left, top;
    uint32_t tile_th, tile_full;
    UDiv w_div, band_div, th_div;
    float cam[12];
    float dof, aperture;
};
struct CamSrc {
    uint64_t w, r, s;   /* the work item, its request and the request's scene */
};
#define CONSTAS __attribute__((address_space(4)))
/* false (src unset) unless the chunk is one work item of generated camera rays */
__device__ inline bool camsrc_load(const KParams& P, CamSrc& src) {
    if (P.list) return false;
    const GLOBAL DChunk* C = gp(P.chunk);
    if (uni_i(C->n_work) != 1) return false;
    const GLOBAL DWork* W = C->work;
    const GLOBAL DReq* R = C->reqs + uni_i(W->req);
    src.w = (uint64_t)uni_p(W);
    src.r = (uint64_t)uni_p(R);
    src.s = (uint64_t)uni_p(gp(R->scene));
    return true;
}
/* The camera values, by scalar loads. The empty asm makes the addresses opaque, so the loads are
 * not hoisted out of the caller's loop (which would bring the register pressure back). */
__device__ inline CamU camu_fetch(CamSrc src) {
    asm volatile("" : "+s"(src.w), "+s"(src.r), "+s"(src.s));
    const CONSTAS DWork* W = (const CONSTAS DWork*)src.w;
    const CONSTAS DReq* R = (const CONSTAS DReq*)src.r;
    const CONSTAS DScene* S = (const CONSTAS DScene*)src.s;
    CamU U;
    U.pid_base = W->pid_base;
    U.px0 = W->px0;
    U.npx = W->npx;
    U.sg_log2 = W->sg_log2;
    U.pass0 = W->pass0;
    U.span_div.lo = W->span_div.lo;
    U.span_div.hi = W->span_div.hi;
    U.req = W->req;
    U.width = R->width;
    U.pixel_size = R->pixel_size;
    U.left = R->left;
    U.top = R->top;
    U.tile_th = R->tile_th;
    U.tile_full = R->tile_full;
    U.w_div.lo = R->w_div.lo;
    U.w_div.hi = R->w_div.hi;
    U.band_div.lo = R->band_div.lo;
    U.band_div.hi = R->band_div.hi;
    U.th_div.lo = R->th_div.lo;
    U.th_div.hi = R->th_div.hi;
    for (int k = 0; k < 12; k++) U.cam[k] = S->cam[k];
    U.dof = S->dof;
    U.aperture = S->aperture;
    return U;
}
__device__ inline void camera_ray_u(const CamU& U, uint32_t i, float4& o4, float4& d4, uint2& rg) {
    const uint32_t l = i - U.pid_base;
    const uint32_t sh = U.sg_log2, span = U.npx << sh;
    const uint32_t kb = udiv(l, U.span_div), r = l - kb * span;
    const uint32_t q = r >> sh;
    const uint32_t k = (kb << sh) + (r - (q << sh));
    const uint32_t pass = (uint32_t)U.pass0 + k;
    uint32_t pixel = U.px0 + q;
    /* tile_pixel_r */
    if (U.tile_th > 1 && pixel < U.tile_full) {
        const uint32_t th = U.tile_th, w = (uint32_t)U.width, band_px = th * w;
        const uint32_t band = udiv(pixel, U.band_div), rr = pixel - band * band_px;
        const uint32_t x = udiv(rr, U.th_div), y = band * th + (rr - x * th);
        pixel = y * w + x;
    }
    uint64_t rs = ilm_path_seed(pass, pixel);
    const uint64_t inc = ilm_pcg32_inc(pixel);
    const float pixel_size = U.pixel_size;
    const uint32_t yq = udiv(pixel, U.w_div);
    int x = (int)(pixel - yq * (uint32_t)U.width), y = (int)yq;
    float left_edge = U.left + pixel_size * (float)x;
    float top_edge = U.top - pixel_size * (float)y;
    float rx = left_edge + pixel_size * ilm_pcg32_uniform(&rs, inc);
    float ry = top_edge - pixel_size * ilm_pcg32_uniform(&rs, inc);
    V3 pos = mk3(0.0f, 0.0f, 0.0f);
    if (U.aperture > ILM_FLT_EPSILON) {
        float u1 = ilm_pcg32_uniform(&rs, inc);
        float u2 = ilm_pcg32_uniform(&rs, inc);
        pos = mul3(sample_circle(u1, u2), U.aperture);
    }
    V3 origin = xform3(U.cam, pos, 1.0f);
    V3 image_pos = xform3(U.cam, mk3(rx, ry, U.dof), 1.0f);
    V3 dir = normalize3(sub3(image_pos, origin));
    o4 = make_float4(origin.x, origin.y, origin.z, u2f((uint32_t)U.req << 16));
    d4 = make_float4(dir.x, dir.y, dir.z, u2f(pixel));
    rg = make_uint2((uint32_t)rs, (uint32_t)(rs >> 32));
}

__device__ inline void spheres_first(const GLOBAL DScene* S, V3 O, V3 D, Best& B) {
    const int ns = S->n_spheres;
    const GLOBAL DSphere* sph = gp(S->spheres);
    for (int s = 0; s < ns; s++) {
        DSphere sp = sph[s];
        float t;
        if (sphere_hit(sp, O, D, t) && t < B.t) { B.t = t; B.inst = -2 - s; }
    }
}

/* Free-flight cap for a ray inside an active medium. The shading stage's first draw of the
 * segment is the scatter distance s = -log(u) / sigma_s (src/renderer.cu:152-160), and the path
 * scatters when s < t_hit, so hits beyond s never matter and the closest-hit query may stop
 * there. The draw is made here from a copy of the RNG state (shade makes it again, unchanged);
 * the cap is the next float above s so that a hit at exactly s is still found and shade's
 * comparison decides as it would have. FLT_MAX when there is no cap (no medium, inactive
 * medium, or s not finite). */
__device__ inline float medium_tmax_m(const GLOBAL DMedium* mediums, uint32_t tag, uint2 rg, uint32_t pixel) {
    const int medium = (int)(tag & 0xffffu) - 1;
    if (medium < 0) return ILM_FLT_MAX;
    /* both fields in one request: the test of `active` does not wait for a second load */
    const GLOBAL DMedium* md = mediums + medium;
    const float scattering = md->scattering;
    const int active = md->active;
    asm volatile("" ::"v"(scattering), "v"(active));
    if (!active) return ILM_FLT_MAX;
    uint64_t rs = ((uint64_t)rg.y << 32) | (uint64_t)rg.x;
    const float u = ilm_pcg32_uniform(&rs, ilm_pcg32_inc(pixel));
    const float sdist = -pm_logf(u) / scattering;
    if (!(sdist < ILM_FLT_MAX)) return ILM_FLT_MAX;
    /* u = 1 gives sdist = -0.0: the path scatters whatever it hits, and the cap must stay
     * positive (the next float above +0) so that shade's "sdist < t" still holds */
    const float s0 = sdist > 0.0f ? sdist : 0.0f;
    return __uint_as_float(__float_as_uint(s0) + 1u);
}
__device__ inline float medium_tmax(const GLOBAL DScene* S, uint32_t tag, uint2 rg, uint32_t pixel) {
    return medium_tmax_m(gp(S->mediums), tag, rg, pixel);
}

/* Intersect the triangles of BLAS leaf `node`. A triangle at exactly the distance of the best
 * triangle hit so far replaces it when its (instance, original index) is smaller, so the closest
 * hit of a ray does not depend on the order its walk visits triangles in (the walk order differs
 * between node layouts, and a packet's between mixes of requests); a sphere keeps a tie, as it
 * would under the strict < of src/renderer.cu:117-136 with spheres tested first. */
template <int TRI, int COUNT>
__device__ inline void leaf_tris(const GLOBAL float4* tris, int node, V3 o, V3 d, const WTRay& wt, int cur, Best& B,
                                 TravCount* tc) {
    const uint32_t f = ilm_leaf_first(node), cnt = ilm_leaf_count(node);
    if (COUNT) tc->tris += cnt;
    for (uint32_t q = 0; q < cnt; q++) {
        const GLOBAL float4* tp = tris + 3 * (f + q);
        float4 A = tp[0], Bq = tp[1], C = tp[2];
        float u = 0.0f, v = 0.0f, t;
        if (TRI) t = tri_wt(wt, mk3(A.x, A.y, A.z), mk3(Bq.x, Bq.y, Bq.z), mk3(C.x, C.y, C.z), o, u, v);
        else t = tri_mt(mk3(A.x, A.y, A.z), mk3(Bq.x, Bq.y, Bq.z), mk3(C.x, C.y, C.z), o, d, u, v);
        const int prim = __float_as_int(A.w);
        const bool tie = t == B.t && B.inst >= 0 && (cur < B.inst || (cur == B.inst && prim < B.prim));
        if (t > 0.0f && (t < B.t || tie)) best_tri(B, t, u, v, cur, A);
    }
}

/* Push the hit children of an inner node nearest-last and continue with the nearest; returns
 * false when no child is hit. n0..n2 are the node's box words (FP32 BNode a, b, c or the QNode's
 * first three), ch its child codes. */
template <int QUANT, int OVF, int COUNT>
__device__ inline bool inner_step_words(const float4 n0, const float4 n1, const float4 n2, const int4 ch, int& node,
                                        V3 id, V3 o, float tmax, Stack<OVF>& st, TravCount* tc) {
    if (QUANT) {
        float tn[4];
        uint32_t hm;
        int nc;
        box4q(n0, n1, n2, id, o, tmax, tn, hm, nc);
        if (COUNT) tc->boxes += nc;
        if (hm == 0) return false;
        /* order the hit children near to far with a 4-input sorting network (missed children
         * sort last as +inf), descend into the nearest and push the rest far-to-near: no
         * data-dependent loops, so the lanes of a wave stay together */
        float k0 = tn[0], k1 = tn[1], k2 = tn[2], k3 = tn[3];
        int c0 = ch.x, c1 = ch.y, c2 = ch.z, c3 = ch.w;
#define ILM_CE(ka, ca, kb, cb)                                        \
        {                                                             \
            const bool sw = kb < ka;                                  \
            const float tk = sw ? kb : ka; kb = sw ? ka : kb; ka = tk; \
            const int tc_ = sw ? cb : ca; cb = sw ? ca : cb; ca = tc_; \
        }
        ILM_CE(k0, c0, k1, c1)
        ILM_CE(k2, c2, k3, c3)
        ILM_CE(k0, c0, k2, c2)
        ILM_CE(k1, c1, k3, c3)
        ILM_CE(k1, c1, k2, c2)
#undef ILM_CE
        st.push_far(__builtin_popcount(hm), c1, c2, c3);
        node = c0;
        return true;
    } else {
        float t0, t1;
        bool h0, h1;
        box2(n0, n1, n2, id, o, tmax, t0, t1, h0, h1);
        if (COUNT) tc->boxes += 2;
        if (h0 && h1) {
            int nearc = ch.x, farc = ch.y;
            if (t1 < t0) { nearc = ch.y; farc = ch.x; }
            st.push(farc);
            node = nearc;
            return true;
        }
        if (h0) { node = ch.x; return true; }
        if (h1) { node = ch.y; return true; }
        return false;
    }
}

/* The four 16-byte words of a node (both layouts are 64 B: three box words, then the codes). */
template <int QUANT>
__device__ inline void load_node(const GLOBAL BNode* nodes, const GLOBAL QNode* qnodes, int node, float4& n0,
                                 float4& n1, float4& n2, int4& ch) {
    const GLOBAL float4* p = QUANT ? (const GLOBAL float4*)(qnodes + node) : (const GLOBAL float4*)(nodes + node);
    n0 = p[0];
    n1 = p[1];
    n2 = p[2];
    ch = ((const GLOBAL int4*)p)[3];
}

template <int QUANT, int OVF, int COUNT>
__device__ inline bool inner_step(const GLOBAL BNode* nodes, const GLOBAL QNode* qnodes, int& node, V3 id, V3 o,
                                  float tmax, Stack<OVF>& st, TravCount* tc) {
    float4 n0, n1, n2;
    int4 ch;
    load_node<QUANT>(nodes, qnodes, node, n0, n1, n2, ch);
    return inner_step_words<QUANT, OVF, COUNT>(n0, n1, n2, ch, node, id, o, tmax, st, tc);
}

/* The same visit for a ray kept as (o, d): the node's loads are issued first and 1/d, -o/d are
 * formed while they are in flight. The empty asm takes the child codes, so all four words are
 * loaded together (the compiler would otherwise move the codes' load behind the box tests,
 * where only hit nodes need them, and a visit would wait on memory twice). */
template <int QUANT, int OVF, int COUNT>
__device__ inline bool inner_step_od(const GLOBAL BNode* nodes, const GLOBAL QNode* qnodes, int& node, V3 o, V3 d,
                                     float tmax, Stack<OVF>& st, TravCount* tc) {
    float4 n0, n1, n2;
    int4 ch;
    load_node<QUANT>(nodes, qnodes, node, n0, n1, n2, ch);
    const V3 id = rcp3(d);
    asm volatile("" ::"v"(ch.x), "v"(ch.y), "v"(ch.z), "v"(ch.w));
    return inner_step_words<QUANT, OVF, COUNT>(n0, n1, n2, ch, node, id, o, tmax, st, tc);
}

template <int TRI, int QUANT, int COUNT, int OVF>
__device__ inline void trace_ray(const GLOBAL DScene* S, V3 O, V3 D, float tmax0, Best& B, Stack<OVF>& st,
                                 TravCount* tc) {
    best_init(B);
    B.t = tmax0;
    spheres_first(S, O, D, B);
    int node = QUANT ? S->qtlas_root : S->tlas_root;
    if (node == ILM_NO_CHILD) return;
    const GLOBAL BNode* nodes = gp(S->tlas);
    const GLOBAL QNode* qnodes = gp(S->qtlas);
    const GLOBAL float4* tris = nullptr;
    bool blas = false;
    int cur = -1;
    V3 o = O, d = D;
    V3 id = rcp3(d);
    WTRay wt;
    st.sp = 0;
    for (;;) {
        if (node >= 0) {
            if (inner_step<QUANT, OVF, COUNT>(nodes, qnodes, node, id, o, B.t, st, tc)) continue;
        } else if (!blas) {
            /* TLAS leaf: one instance; enter its BLAS in object space */
            const int k = (int)ilm_leaf_first(node);
            float inv[12];
            const GLOBAL BNode* bn;
            const GLOBAL QNode* qn;
            const GLOBAL float4* tr;
            const int broot = inst_load<QUANT>(gp(S->inst) + k, inv, bn, qn, tr);
            if (broot != ILM_NO_CHILD) {
                o = xform3(inv, O, 1.0f);
                d = xform3(inv, D, 0.0f);
                id = rcp3(d);
                if (TRI) wt = wt_setup(d);
                nodes = bn;
                qnodes = qn;
                tris = tr;
                st.push(STACK_SENTINEL);
                blas = true;
                cur = k;
                node = broot;
                continue;
            }
        } else {
            leaf_tris<TRI, COUNT>(tris, node, o, d, wt, cur, B, tc);
        }
        /* pop */
        for (;;) {
            if (st.sp == 0) return;
            node = st.pop();
            if (node != STACK_SENTINEL) break;
            blas = false;
            nodes = gp(S->tlas);
            qnodes = gp(S->qtlas);
            o = O;
            d = D;
            id = rcp3(d);
        }
    }
}

/* ---------------- camera rays as packets ----------------
 * At the first depth a wave's 64 consecutive path ids are samples of one pixel (or of a few
 * neighbouring pixels), so their camera rays are nearly the same. The wave walks one BVH
 * traversal for all of them: the node, the stack (in LDS) and every branch are uniform, each
 * lane tests the node's children against its own ray and its own closest hit so far, and a
 * child is visited when any lane hits it, nearest first by the first hitting lane's entry
 * distance. A lane that misses a subtree only repeats tests that cannot change its result, so
 * every lane ends with the closest hit of its own ray (the order of visits can only change which
 * of two triangles at exactly the same distance is kept). */
#define PKT_STACK 128   /* = LDS_STACK + OVF_STACK: the largest stack bound a scene is accepted with */


/* The packet walk over the lanes that call it (all of them on scene S). */
template <int TRI, int QUANT, int COUNT>
__device__ inline void packet_trace(const GLOBAL DScene* S, V3 O, V3 D, Best& B, LDSAS int* stk, TravCount* tc) {
    spheres_first(S, O, D, B);
    int node = uni_i(QUANT ? S->qtlas_root : S->tlas_root);
    if (node == ILM_NO_CHILD) return;
    const GLOBAL BNode* nodes = uni_p(gp(S->tlas));
    const GLOBAL QNode* qnodes = uni_p(gp(S->qtlas));
    const GLOBAL float4* tris = nullptr;
    bool blas = false;
    int cur = -1;
    V3 o = O, d = D;
    WTRay wt;
    int sp = 0;
    for (;;) {
        bool pop = true;
        if (node >= 0) {
            float4 n0, n1, n2;
            int4 ch;
            load_node<QUANT>(nodes, qnodes, node, n0, n1, n2, ch);
            const V3 id = rcp3(d);
            float tn[4];
            uint32_t hm;
            if (QUANT) {
                int nc;
                box4q(n0, n1, n2, id, o, B.t, tn, hm, nc);
                if (COUNT) tc->boxes += nc;
            } else {
                bool h0, h1;
                box2(n0, n1, n2, id, o, B.t, tn[0], tn[1], h0, h1);
                tn[2] = tn[3] = 0.0f;
                hm = (h0 ? 1u : 0u) | (h1 ? 2u : 0u);
                if (COUNT) tc->boxes += 2;
            }
            const uint32_t any = (__ballot(hm & 1u) ? 1u : 0u) | (__ballot(hm & 2u) ? 2u : 0u) |
                                 (__ballot(hm & 4u) ? 4u : 0u) | (__ballot(hm & 8u) ? 8u : 0u);
            if (any) {
                const int lead = (int)__builtin_ctzll(__ballot(hm != 0u));
                const uint32_t lh = (uint32_t)__builtin_amdgcn_readlane((int)hm, lead);
                /* sort keys as unsigned ints: entry distances are >= 0 (the sign bit is cleared
                 * for -0), so their bit patterns order as the floats do, and the network below
                 * runs on the scalar unit (the keys are uniform) */
                uint32_t k[4];
                int c[4] = {uni_i(ch.x), uni_i(ch.y), uni_i(ch.z), uni_i(ch.w)};
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const uint32_t lt = (uint32_t)__builtin_amdgcn_readlane(__float_as_int(tn[j]), lead) & 0x7fffffffu;
                    k[j] = !((any >> j) & 1u) ? 0x7f800000u : (((lh >> j) & 1u) ? lt : 0x7f7fffffu);   /* +inf, FLT_MAX */
                }
#define ILM_CE(a, b)                                                                  \
                {                                                                     \
                    const bool sw = k[b] < k[a];                                      \
                    const uint32_t tk = sw ? k[b] : k[a]; k[b] = sw ? k[a] : k[b]; k[a] = tk; \
                    const int tc_ = sw ? c[b] : c[a]; c[b] = sw ? c[a] : c[b]; c[a] = tc_; \
                }
                ILM_CE(0, 1)
                ILM_CE(2, 3)
                ILM_CE(0, 2)
                ILM_CE(1, 3)
                ILM_CE(1, 2)
#undef ILM_CE
                const int n = __builtin_popcount(any);
                for (int j = n - 1; j >= 1; j--) {
                    if (sp < PKT_STACK) stk[sp] = uni_i(c[j]);
                    sp++;
                }
                node = uni_i(c[0]);
                pop = false;
            }
        } else if (!blas) {
            const int k = uni_i((int)ilm_leaf_first(node));
            const GLOBAL DInst* I = gp(S->inst) + k;
            const int broot = uni_i(QUANT ? I->qroot : I->root);
            if (broot != ILM_NO_CHILD) {
                float inv[12];
                load_rows12(I->inv, inv);
                o = xform3(inv, O, 1.0f);
                d = xform3(inv, D, 0.0f);
                if (TRI) wt = wt_setup(d);
                nodes = uni_p(gp(I->nodes));
                qnodes = uni_p(gp(I->qnodes));
                tris = uni_p(gp(I->tris));
                if (sp < PKT_STACK) stk[sp] = STACK_SENTINEL;
                sp++;
                blas = true;
                cur = k;
                node = broot;
                pop = false;
            }
        } else {
            leaf_tris<TRI, COUNT>(tris, node, o, d, wt, cur, B, tc);
        }
        if (pop) {
            for (;;) {
                if (sp == 0) return;
                sp--;
                if (sp >= PKT_STACK) continue;   /* a push past the capacity was dropped (not within the bound) */
                node = uni_i(stk[sp]);
                if (node != STACK_SENTINEL) break;
                blas = false;
                nodes = uni_p(gp(S->tlas));
                qnodes = uni_p(gp(S->qtlas));
                o = O;
                d = D;
            }
        }
    }
}

/* First depth with per-step refill: each wave reserves consecutive path ids (P.ext_fetch per
 * queue atomic, shrinking to 64 as the queue drains, as in k_extend), and traces them 64 at a
 * time: forms their camera rays (camera_ray) and walks one packet per distinct scene among them.
 * Every path id of the chunk passes through here once, so the kernel also clears the per-path
 * results (the chunk needs no separate clear). */
template <int TRI, int QUANT, int COUNT>
__global__ __launch_bounds__(EXT_BLOCK) __attribute__((amdgpu_waves_per_eu(7, 8))) void k_extend_packet(KParams P, int depth) {
    __shared__ int stk[PKT_STACK];
    const int lane = threadIdx.x;
    GLOBAL DCtl* ctl = gpw(P.ctl);
    const uint32_t count = ctl->count[depth];
    const uint32_t fetch = (uint32_t)P.ext_fetch;
    const GLOBAL DReq* reqs = gp(P.chunk)->reqs;
    TravCount tc = {0u, 0u};
    const uint32_t guide = gridDim.x * 2u;
    CamSrc cs;
    const bool cu = camsrc_load(P, cs);
    const uint64_t su = cu ? cs.s : 0ull;
    /* the first reservation is fixed by the wave's index; the cursor hands out the rest (one
     * address's atomics serialise, so a wave with nothing to do should not touch it) */
    const uint32_t first = gridDim.x * fetch;
    uint32_t want = fetch;
    uint32_t b = 0, end = 0;
    bool first_trip = true;
    for (;;) {
        if (b >= end) {
            uint32_t base = 0;
            if (first_trip) {
                base = blockIdx.x * fetch;
                first_trip = false;
            } else {
                if (first >= count) break;
                if (lane == 0) base = atomicAdd(&ctl->fetch[depth], want);
                base = first + (uint32_t)uni_i((int)__shfl(base, 0));
            }
            if (base >= count) break;
            b = base;
            end = base + want < count ? base + want : count;
            const uint32_t left = count - end;
            const uint32_t g = (left / guide) & ~(uint32_t)(EXT_BLOCK - 1);
            want = g < (uint32_t)EXT_BLOCK ? (uint32_t)EXT_BLOCK : (g > fetch ? fetch : g);
        }
        const uint32_t r = b + (uint32_t)lane;
        const bool valid = r < end;
        b += (uint32_t)EXT_BLOCK;
        if (valid) {
            gpw(P.contrib)[r] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (P.seg_count) gpw(P.seg_count)[r] = 0;
        }
        float4 o4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), d4 = o4;
        uint2 rg;
        uint64_t sbits = 0;
        Best B;
        best_init(B);
        if (valid) {
            if (cu) {
                camera_ray_u(camu_fetch(cs), r, o4, d4, rg);
                sbits = su;
            } else {
                camera_ray(P, r, o4, d4, rg);
                sbits = (uint64_t)gp(reqs[f2u(o4.w) >> 16].scene);
            }
        }
        unsigned long long rem = __ballot(valid);
        while (rem) {
            const int lead = (int)__builtin_ctzll(rem);
            const uint64_t sl = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(sbits >> 32), lead) << 32) |
                                (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)sbits, lead);
            const bool mine = valid && sbits == sl;
            rem &= ~__ballot(mine);
            if (mine)
                packet_trace<TRI, QUANT, COUNT>((const GLOBAL DScene*)sl, mk3(o4.x, o4.y, o4.z), mk3(d4.x, d4.y, d4.z),
                                                B, (LDSAS int*)stk, &tc);
        }
        if (valid) write_hit(P, r, B);
    }
    if (COUNT) {
        atomicAdd(&P.trav_stats[0], (unsigned long long)tc.boxes);
        atomicAdd(&P.trav_stats[1], (unsigned long long)tc.tris);
    }
}

/* One ray batch per wave at a time: a wave reserves P.ext_fetch consecutive queue entries with
 * one atomic (same-address atomics serialise, so one per 64-ray batch throttles the queue) and
 * walks them as consecutive 64-ray batches, each to completion. */
template <int TRI, int QUANT, int COUNT, int OVF>
__global__ EXT_BOUNDS void k_extend(KParams P, int depth, int* ovf_base) {
    __shared__ int stk[LDS_STACK * EXT_BLOCK];
    const int lane = threadIdx.x;
    Stack<OVF> st;
    st.lds = (LDSAS int*)stk + lane;
    st.ovf = gpw(ovf_base) + (size_t)blockIdx.x * EXT_BLOCK * OVF_STACK;
    st.sp = 0;
    GLOBAL DCtl* ctl = gpw(P.ctl);
    const uint32_t count = ctl->count[depth];
    const uint32_t fetch = (uint32_t)P.ext_fetch;
    const GLOBAL float4* ino = gp(P.in.o);
    const GLOBAL float4* ind = gp(P.in.d);
    const GLOBAL uint2* inr = gp(P.in.rng);
    const GLOBAL DReq* reqs = gp(P.chunk)->reqs;
    TravCount tc = {0u, 0u};
    /* guided self-scheduling: reservations shrink with the rays left in the queue (down to one
     * batch), so the last waves of a launch finish together; ext_fetch is the upper bound */
    const uint32_t guide = gridDim.x * 2u;
    /* the first reservation is fixed by the wave's index; the cursor hands out the rest (one
     * address's atomics serialise, so a wave with nothing to do should not touch it) */
    const uint32_t first = gridDim.x * fetch;
    uint32_t want = fetch;
    for (bool first_trip = true;; first_trip = false) {
        uint32_t base = 0;
        if (first_trip) {
            base = blockIdx.x * fetch;
        } else {
            if (first >= count) break;
            if (lane == 0) base = atomicAdd(&ctl->fetch[depth], want);
            base = first + __shfl(base, 0);
        }
        if (base >= count) break;
        const uint32_t end = base + want < count ? base + want : count;
        const uint32_t left = count - end;
        const uint32_t g = (left / guide) & ~(uint32_t)(EXT_BLOCK - 1);
        want = g < (uint32_t)EXT_BLOCK ? (uint32_t)EXT_BLOCK : (g > fetch ? fetch : g);
        for (uint32_t b = base; b < end; b += EXT_BLOCK) {
            const uint32_t i = b + lane;
            if (i < end) {
                float4 o4 = ino[i];
                float4 d4 = ind[i];
                const GLOBAL DScene* S = gp(reqs[f2u(o4.w) >> 16].scene);
                const float tmax0 = medium_tmax(S, f2u(o4.w), inr[i], f2u(d4.w));
                Best B;
                trace_ray<TRI, QUANT, COUNT, OVF>(S, mk3(o4.x, o4.y, o4.z), mk3(d4.x, d4.y, d4.z), tmax0, B, st, &tc);
                write_hit(P, i, B);
            }
        }
    }
    if (COUNT) {
        atomicAdd(&P.trav_stats[0], (unsigned long long)tc.boxes);
        atomicAdd(&P.trav_stats[1], (unsigned long long)tc.tris);
    }
}

/* ---------------- extend with lane refill ----------------
 * Each lane keeps its ray's traversal state across an outer loop; after every BLAS leaf (and
 * when its ray finishes) the wave counts its idle lanes and, once P.refill_min of them are idle,
 * hands them the next rays of a wave-local pool taken from the queue 64 at a time (persistent
 * traversal with dynamic fetch, Aila and Laine, HPG 2009). Trades ray coherence for lane
 * occupancy; selected with ilm_opts.extend_mode = 1. */
#define POP_MARK ((int)0x7ffffffd)
#define NO_RAY 0xffffffffu

struct TState {
    V3 o, d, id;           /* ray in the current (world or object) space */
    const GLOBAL BNode* nodes;
    const GLOBAL QNode* qnodes;
    const GLOBAL float4* tris;
    const GLOBAL DScene* S;
    int node;
    int cur;
    int blas;
    WTRay wt;
};

template <int QUANT, int OVF>
__device__ inline void seg_init(TState& T, Best& B, Stack<OVF>& st, const GLOBAL DScene* S, V3 O, V3 D, float tmax0) {
    best_init(B);
    B.t = tmax0;
    spheres_first(S, O, D, B);
    T.S = S;
    T.node = QUANT ? S->qtlas_root : S->tlas_root;
    if (T.node == ILM_NO_CHILD) T.node = POP_MARK;   /* no instances: nothing to walk */
    T.nodes = gp(S->tlas);
    T.qnodes = gp(S->qtlas);
    T.tris = nullptr;
    T.blas = 0;
    T.cur = -1;
    T.o = O;
    T.d = D;
    T.id = rcp3(D);
    st.sp = 0;
}

/* Walk until one BLAS leaf has been intersected (returns false) or the ray is done (true). */
template <int TRI, int QUANT, int COUNT, int OVF>
__device__ inline bool seg_run(TState& T, Best& B, Stack<OVF>& st, const GLOBAL float4* wo, const GLOBAL float4* wd,
                               uint32_t ridx, TravCount* tc) {
    int node = T.node;
    for (;;) {
        if (node == POP_MARK) goto pop;
        if (node >= 0) {
            if (inner_step<QUANT, OVF, COUNT>(T.nodes, T.qnodes, node, T.id, T.o, B.t, st, tc)) continue;
            goto pop;
        } else if (!T.blas) {
            const int k = (int)ilm_leaf_first(node);
            float inv[12];
            const GLOBAL BNode* bn;
            const GLOBAL QNode* qn;
            const GLOBAL float4* tr;
            const int broot = inst_load<QUANT>(gp(T.S->inst) + k, inv, bn, qn, tr);
            if (broot == ILM_NO_CHILD) goto pop;
            V3 O = T.o, D = T.d;
            T.o = xform3(inv, O, 1.0f);
            T.d = xform3(inv, D, 0.0f);
            T.id = rcp3(T.d);
            if (TRI) T.wt = wt_setup(T.d);
            T.nodes = bn;
            T.qnodes = qn;
            T.tris = tr;
            st.push(STACK_SENTINEL);
            T.blas = 1;
            T.cur = k;
            node = broot;
            continue;
        } else {
            leaf_tris<TRI, COUNT>(T.tris, node, T.o, T.d, T.wt, T.cur, B, tc);
            T.node = POP_MARK;          /* resume with a pop after the wave had a chance to refill */
            return st.sp == 0;
        }
    pop:
        for (;;) {
            if (st.sp == 0) return true;
            node = st.pop();
            if (node != STACK_SENTINEL) break;
            /* leaving the instance: the world ray comes back from the path state */
            float4 a = wo[ridx], b = wd[ridx];
            T.o = mk3(a.x, a.y, a.z);
            T.d = mk3(b.x, b.y, b.z);
            T.id = rcp3(T.d);
            T.blas = 0;
            T.nodes = gp(T.S->tlas);
            T.qnodes = gp(T.S->qtlas);
        }
    }
}

/* Rays per reservation: `fetch`, or for a queue shorter than the grid's first reservations,
 * about an even share of it in whole batches of 64, so that the rays spread over the waves
 * instead of queueing behind one another in a few. */
__device__ inline uint32_t pool_fetch(uint32_t count, uint32_t fetch) {
    const uint32_t per = (count + gridDim.x - 1u) / gridDim.x;
    const uint32_t share = per <= (uint32_t)EXT_BLOCK ? (uint32_t)EXT_BLOCK
                                                      : (per + (uint32_t)(EXT_BLOCK - 1)) & ~(uint32_t)(EXT_BLOCK - 1);
    return share < fetch ? share : fetch;
}
/* Work queue of the per-lane extend kernels: wave w first takes rays [w*fetch, (w+1)*fetch)
 * without an atomic, and the shared cursor then hands out what lies past the grid's first
 * reservations, `fetch` rays per atomic. A wave with no rays leaves without touching the cursor:
 * atomics on one address serialise, and a cursor that every one of the grid's ~7k waves had to
 * touch cost a launch ~120 us however short its queue (the last depths of a render hold few
 * rays). Neighbouring waves keep neighbouring rays (coherent in the queue): first spans scaled to
 * an even split of a long queue measured slower. */
__device__ inline void pool_first(uint32_t count, uint32_t fetch, uint32_t& pool, uint32_t& pool_end) {
    const uint32_t b = blockIdx.x * fetch;
    pool = b < count ? b : count;
    pool_end = b + fetch < count ? b + fetch : count;
}
__device__ inline void pool_next(GLOBAL uint32_t* cursor, uint32_t count, uint32_t fetch, int lane, uint32_t& pool,
                                 uint32_t& pool_end) {
    const uint32_t first = gridDim.x * fetch;
    if (first >= count) {
        pool = pool_end = count;
        return;
    }
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(cursor, fetch);
    base = first + __shfl(base, 0);
    pool = base < count ? base : count;
    pool_end = base + fetch < count ? base + fetch : count;
}

template <int TRI, int QUANT, int COUNT, int OVF>
__global__ EXT_BOUNDS void k_extend_refill(KParams P, int depth, int* ovf_base) {
    __shared__ int stk[LDS_STACK * EXT_BLOCK];
    const int lane = threadIdx.x;
    Stack<OVF> st;
    st.lds = (LDSAS int*)stk + lane;
    st.ovf = gpw(ovf_base) + (size_t)blockIdx.x * EXT_BLOCK * OVF_STACK;
    st.sp = 0;
    GLOBAL DCtl* ctl = gpw(P.ctl);
    const uint32_t count = ctl->count[depth];
    const uint32_t fetch = pool_fetch(count, (uint32_t)P.ext_fetch);
    const GLOBAL float4* ino = gp(P.in.o);
    const GLOBAL float4* ind = gp(P.in.d);
    const GLOBAL uint2* inr = gp(P.in.rng);
    const GLOBAL DReq* reqs = gp(P.chunk)->reqs;
    uint32_t ridx = NO_RAY;
    /* wave-local pool [pool, pool_end) of ray indices taken from the queue */
    uint32_t pool, pool_end;
    pool_first(count, fetch, pool, pool_end);
    bool drained = pool == pool_end;
    TState T;
    Best B;
    TravCount tc = {0u, 0u};
    for (;;) {
        const bool idle = ridx == NO_RAY;
        const unsigned long long idle_mask = __ballot(idle);
        const uint32_t n_idle = (uint32_t)__popcll(idle_mask);
        if (!drained && n_idle >= (uint32_t)P.refill_min) {
            if (pool == pool_end) pool_next(&ctl->fetch[depth], count, fetch, lane, pool, pool_end);
            const uint32_t take = n_idle < pool_end - pool ? n_idle : pool_end - pool;
            if (idle) {
                /* idle lanes below this one */
                const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle_mask >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((uint32_t)idle_mask, 0u));
                if (rank < take) {
                    const uint32_t r = pool + rank;
                    ridx = r;
                    float4 o4 = ino[r], d4 = ind[r];
                    const GLOBAL DScene* S = gp(reqs[f2u(o4.w) >> 16].scene);
                    seg_init<QUANT, OVF>(T, B, st, S, mk3(o4.x, o4.y, o4.z), mk3(d4.x, d4.y, d4.z),
                                         medium_tmax(S, f2u(o4.w), inr[r], f2u(d4.w)));
                }
            }
            pool += take;
            if (pool == pool_end && pool_end >= count) drained = true;
        }
        if (ridx != NO_RAY) {
            if (seg_run<TRI, QUANT, COUNT, OVF>(T, B, st, ino, ind, ridx, &tc)) {
                write_hit(P, ridx, B);
                ridx = NO_RAY;
            }
        }
        if (drained && __ballot(ridx != NO_RAY) == 0ull) break;
    }
    if (COUNT) {
        atomicAdd(&P.trav_stats[0], (unsigned long long)tc.boxes);
        atomicAdd(&P.trav_stats[1], (unsigned long long)tc.tris);
    }
}

/* ---------------- extend with per-step refill ----------------
 * Like the refill kernel, but a lane advances its ray by exactly one action per loop trip (one
 * inner-node visit, one instance entry or one BLAS leaf, plus the pops that follow), and the
 * wave re-checks for idle lanes on every trip, so a lane whose ray finished takes a new one
 * after at most one trip instead of waiting for the rest of the wave to reach a leaf. The world
 * ray stays in registers for the return from an instance. Selected with extend_mode = 3
 * (ILM_EXTEND_STEP_REFILL) and by the default, 0. */
/* PSTATE_LEAN: the reciprocal direction and -o/d are recomputed at each node visit from the
 * current ray instead of being carried across loop trips (fewer live registers and fewer
 * register copies on the hot path, for a few extra VALU operations per visit). */
#ifndef PSTATE_LEAN
#define PSTATE_LEAN 1
#endif
struct PState {
    V3 O, D;                /* world ray */
    V3 o, d;                /* ray in the current (world or object) space */
#if !PSTATE_LEAN
    V3 id;
#endif
    const GLOBAL BNode* nodes;
    const GLOBAL QNode* qnodes;
    const GLOBAL float4* tris;
    const GLOBAL DScene* S;
    int node;
    int cur;
    int blas;
    WTRay wt;
};

/* The scene values a single-scene chunk's lanes all share, read once per wave (in SGPRs) instead
 * of at every refill and instance exit: a refill then waits on memory for its ray alone. The
 * first USCENE_SPHERES spheres (centre, radius) come along. */
#define USCENE_SPHERES 2
struct UScene {
    const GLOBAL DInst* inst;
    const GLOBAL BNode* tlas;
    const GLOBAL QNode* qtlas;
    const GLOBAL DSphere* spheres;
    const GLOBAL DMedium* mediums;
    int root;               /* the layout's TLAS root code, POP_MARK when there are no instances */
    int n_spheres;
    float4 sph[USCENE_SPHERES];
    float4 r0, r1, r2;      /* the TLAS root's box words and child codes, when the root is an */
    int4 rch;               /* inner node (root >= 0) */
};
__device__ inline float uni_f(float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); }
template <int QUANT>
__device__ inline UScene uscene_load(const GLOBAL DScene* S) {
    UScene U;
    U.inst = uni_p(gp(S->inst));
    U.tlas = uni_p(gp(S->tlas));
    U.qtlas = uni_p(gp(S->qtlas));
    U.spheres = uni_p(gp(S->spheres));
    U.mediums = uni_p(gp(S->mediums));
    const int root = uni_i(QUANT ? S->qtlas_root : S->tlas_root);
    U.root = root == ILM_NO_CHILD ? POP_MARK : root;
    U.r0 = U.r1 = U.r2 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    U.rch = make_int4(0, 0, 0, 0);
    if (QUANT && U.root >= 0 && U.root != POP_MARK) {
        float4 a, b, c;
        int4 ch;
        load_node<QUANT>(U.tlas, U.qtlas, U.root, a, b, c, ch);
        U.r0 = make_float4(uni_f(a.x), uni_f(a.y), uni_f(a.z), uni_f(a.w));
        U.r1 = make_float4(uni_f(b.x), uni_f(b.y), uni_f(b.z), uni_f(b.w));
        U.r2 = make_float4(uni_f(c.x), uni_f(c.y), uni_f(c.z), uni_f(c.w));
        U.rch = make_int4(uni_i(ch.x), uni_i(ch.y), uni_i(ch.z), uni_i(ch.w));
    }
    U.n_spheres = uni_i(S->n_spheres);
#pragma unroll
    for (int k = 0; k < USCENE_SPHERES; k++) {
        U.sph[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (k < U.n_spheres) {
            const GLOBAL DSphere* q = U.spheres + k;
            U.sph[k] = make_float4(uni_f(q->c[0]), uni_f(q->c[1]), uni_f(q->c[2]), uni_f(q->r));
        }
    }
    return U;
}
/* spheres_first over the snapshot: the same spheres in the same order and comparisons */
__device__ inline void spheres_first_u(const UScene& U, V3 O, V3 D, Best& B) {
#pragma unroll
    for (int k = 0; k < USCENE_SPHERES; k++) {
        if (k < U.n_spheres) {
            DSphere sp;
            sp.c[0] = U.sph[k].x; sp.c[1] = U.sph[k].y; sp.c[2] = U.sph[k].z; sp.r = U.sph[k].w;
            float t;
            if (sphere_hit(sp, O, D, t) && t < B.t) { B.t = t; B.inst = -2 - k; }
        }
    }
    for (int k = USCENE_SPHERES; k < U.n_spheres; k++) {
        DSphere sp = U.spheres[k];
        float t;
        if (sphere_hit(sp, O, D, t) && t < B.t) { B.t = t; B.inst = -2 - k; }
    }
}

template <int QUANT, int OVF, int UNI, int COUNT>
__device__ inline void pstep_init(PState& T, Best& B, Stack<OVF>& st, const GLOBAL DScene* S, const UScene& U, V3 O,
                                  V3 D, float tmax0, TravCount* tc) {
    best_init(B);
    B.t = tmax0;
    if (UNI) {
        spheres_first_u(U, O, D, B);
        T.node = U.root;
        T.nodes = U.tlas;
        T.qnodes = U.qtlas;
    } else {
        spheres_first(S, O, D, B);
        T.node = QUANT ? S->qtlas_root : S->tlas_root;
        if (T.node == ILM_NO_CHILD) T.node = POP_MARK;
        T.nodes = gp(S->tlas);
        T.qnodes = gp(S->qtlas);
    }
    T.S = S;
    T.tris = nullptr;
    T.blas = 0;
    T.cur = -1;
    T.O = O;
    T.D = D;
    T.o = O;
    T.d = D;
#if !PSTATE_LEAN
    T.id = rcp3(D);
#endif
    st.sp = 0;
    if (UNI && QUANT && U.root >= 0 && U.root != POP_MARK) {
        /* the TLAS root's words are in registers: visit it now, so the ray's first trip starts
         * one level down (POP_MARK when no child is hit: the next trip ends the ray). 8-bit
         * layout only: in the FP32 kernel the extra code costs a wave of occupancy. */
        int node = U.root;
        T.node = inner_step_words<QUANT, OVF, COUNT>(U.r0, U.r1, U.r2, U.rch, node, rcp3(D), O, B.t, st, tc)
                     ? node : POP_MARK;
    }
}

/* One action of the walk; true when the ray is finished. UNI: every lane's ray belongs to scene
 * Su (a value uniform over the wave), which then replaces the per-lane T.S. A leaf trip tests a
 * BLAS leaf's triangles; a node trip visits an inner node, and a lane at a TLAS leaf enters the
 * instance and visits its BLAS root in the same trip (the root's box words are then the trip's
 * second dependent fetch after the instance record), unless the root is itself a leaf. */
template <int TRI, int QUANT, int COUNT, int OVF, int UNI>
__device__ inline bool pstep(PState& T, Best& B, Stack<OVF>& st, TravCount* tc, const UScene& U) {
    const GLOBAL DScene* S = T.S;
    int node = T.node;
    bool pop;
    if (node < 0 && T.blas) {
        leaf_tris<TRI, COUNT>(T.tris, node, T.o, T.d, T.wt, T.cur, B, tc);
        pop = true;
    } else {
        pop = node == POP_MARK;
        if (node < 0) {
            const int k = (int)ilm_leaf_first(node);
            float inv[12];
            const GLOBAL BNode* bn;
            const GLOBAL QNode* qn;
            const GLOBAL float4* tr;
            const int broot = inst_load<QUANT>((UNI ? U.inst : gp(S->inst)) + k, inv, bn, qn, tr);
            if (broot != ILM_NO_CHILD) {
                T.o = xform3(inv, T.O, 1.0f);
                T.d = xform3(inv, T.D, 0.0f);
#if !PSTATE_LEAN
                T.id = rcp3(T.d);
#endif
                if (TRI) T.wt = wt_setup(T.d);
                T.nodes = bn;
                T.qnodes = qn;
                T.tris = tr;
                st.push(STACK_SENTINEL);
                T.blas = 1;
                T.cur = k;
                node = broot;
            } else {
                pop = true;
            }
        }
        if (!pop && node >= 0) {
#if PSTATE_LEAN
            pop = !inner_step_od<QUANT, OVF, COUNT>(T.nodes, T.qnodes, node, T.o, T.d, B.t, st, tc);
#else
            pop = !inner_step<QUANT, OVF, COUNT>(T.nodes, T.qnodes, node, T.id, T.o, B.t, st, tc);
#endif
        }
    }
    if (pop) {
        for (;;) {
            if (st.sp == 0) return true;
            node = st.pop();
            if (node != STACK_SENTINEL) break;
            T.o = T.O;
            T.d = T.D;
#if !PSTATE_LEAN
            T.id = rcp3(T.D);
#endif
            T.blas = 0;
            T.nodes = UNI ? U.tlas : gp(S->tlas);
            T.qnodes = UNI ? U.qtlas : gp(S->qtlas);
        }
    }
    T.node = node;
    return false;
}

/* UNI: every request of the chunk renders the same scene (the host checks), so the scene
 * pointer is read once, uniform over the wave, instead of per lane at every refill. */
template <int TRI, int QUANT, int COUNT, int OVF, int CAM, int UNI>
__global__ EXT_BOUNDS void k_extend_pers(KParams P, int depth, int* ovf_base) {
    __shared__ int stk[LDS_STACK * EXT_BLOCK];
    const int lane = threadIdx.x;
    Stack<OVF> st;
    st.lds = (LDSAS int*)stk + lane;
    st.ovf = gpw(ovf_base) + (size_t)blockIdx.x * EXT_BLOCK * OVF_STACK;
    st.sp = 0;
    GLOBAL DCtl* ctl = gpw(P.ctl);
    const uint32_t count = ctl->count[depth];
    const uint32_t fetch = pool_fetch(count, (uint32_t)P.ext_fetch);
    const GLOBAL float4* ino = gp(P.in.o);
    const GLOBAL float4* ind = gp(P.in.d);
    const GLOBAL uint2* inr = gp(P.in.rng);
    const GLOBAL DReq* reqs = gp(P.chunk)->reqs;
    const GLOBAL DScene* Su = UNI ? uni_p(gp(reqs[0].scene)) : nullptr;
    UScene U = {};
    if (UNI) U = uscene_load<QUANT>(Su);
    const uint32_t refill_min = (uint32_t)P.refill_min;
    const uint32_t leaf_pct = (uint32_t)P.leaf_pct;
    uint32_t ridx = NO_RAY;
    uint32_t pool, pool_end;
    pool_first(count, fetch, pool, pool_end);
    bool drained = pool == pool_end;
    PState T;
    Best B;
    TravCount tc = {0u, 0u};
    for (;;) {
        const bool idle = ridx == NO_RAY;
        const unsigned long long idle_mask = __ballot(idle);
        const uint32_t n_idle = (uint32_t)__popcll(idle_mask);
        if (drained) {
            if (n_idle == EXT_BLOCK) break;
        } else if (n_idle >= refill_min) {
            if (pool == pool_end) pool_next(&ctl->fetch[depth], count, fetch, lane, pool, pool_end);
            const uint32_t take = n_idle < pool_end - pool ? n_idle : pool_end - pool;
            if (idle) {
                /* idle lanes below this one */
                const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle_mask >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((uint32_t)idle_mask, 0u));
                if (rank < take) {
                    const uint32_t r = pool + rank;
                    ridx = r;
                    if (CAM) {
                        /* camera rays start outside every medium: no free-flight cap */
                        float4 o4, d4;
                        uint2 rg;
                        camera_ray(P, r, o4, d4, rg);
                        const GLOBAL DScene* S = UNI ? Su : gp(reqs[f2u(o4.w) >> 16].scene);
                        pstep_init<QUANT, OVF, UNI, COUNT>(T, B, st, S, U, mk3(o4.x, o4.y, o4.z),
                                                           mk3(d4.x, d4.y, d4.z), ILM_FLT_MAX, &tc);
                    } else {
                        /* the PCG state is read with the ray (it is needed for rays inside a medium) */
                        const float4 o4 = ino[r], d4 = ind[r];
                        const uint2 rg = inr[r];
                        asm volatile("" ::"v"(o4.w), "v"(d4.w), "v"(rg.x));
                        const GLOBAL DScene* S = UNI ? Su : gp(reqs[f2u(o4.w) >> 16].scene);
                        pstep_init<QUANT, OVF, UNI, COUNT>(T, B, st, S, U, mk3(o4.x, o4.y, o4.z),
                                                           mk3(d4.x, d4.y, d4.z),
                                                           medium_tmax_m(UNI ? U.mediums : gp(S->mediums), f2u(o4.w),
                                                                         rg, f2u(d4.w)), &tc);
                    }
                }
            }
            pool += take;
            if (pool == pool_end && pool_end >= count) drained = true;
        }
        /* leaf work and inner-node work alternate at wave level: a trip runs the BLAS-leaf
         * triangle tests only once the lanes waiting at leaves reach leaf_pct percent of the
         * other busy lanes, else only node visits (and instance entries), so a trip pays for one
         * of the two code paths instead of both. Either group is non-empty when chosen unless
         * no lane is busy, so every trip makes progress. */
        const bool has = ridx != NO_RAY;
        const bool at_leaf = has && T.node < 0 && T.blas;
        const uint32_t n_leaf = (uint32_t)__popcll(__ballot(at_leaf));
        const uint32_t n_rest = (uint32_t)__popcll(__ballot(has)) - n_leaf;
        const bool leaf_trip = n_leaf * 100u >= n_rest * leaf_pct;
        if (has && at_leaf == leaf_trip) {
            if (pstep<TRI, QUANT, COUNT, OVF, UNI>(T, B, st, &tc, U)) {
                write_hit(P, ridx, B);
                ridx = NO_RAY;
            }
        }
    }
    if (COUNT) {
        atomicAdd(&P.trav_stats[0], (unsigned long long)tc.boxes);
        atomicAdd(&P.trav_stats[1], (unsigned long long)tc.tris);
    }
}
