/*
 * common.h - device data layout of the path-tracer hot path and the small vector helpers the
 * kernels use. Host code includes it for the struct layouts.
 *
 * Vector helpers keep the operation order of the reference's Vector3 functions
 * (src/math/vector3.cu) so that, compiled with -ffp-contract=off, a path traced here produces
 * the same float bits as the same path traced by a host build of the same formulas.
 */
#pragma once

#include <stdint.h>
#include <hip/hip_runtime.h>

#include "../../include/illume_c.h"
#include "../../include/illume_pmath.h"

#define ILM_FLT_MAX 3.402823466e+38f
#define ILM_FLT_EPSILON 1.192092896e-07f

/* Largest max_depth a request may ask for (sizes the per-depth queue counters). */
#define ILM_MAX_DEPTH 256

/* ---------------- BVH node layouts ---------------- */
/* FP32 BVH2 node, 64 B: both children's boxes live in the parent so one node fetch decides
 * the whole step. child >= 0: inner node index. child < 0: leaf, see ilm_leaf_*(). */
struct __attribute__((aligned(16))) BNode {
    float4 a;   /* lo0.x hi0.x lo0.y hi0.y */
    float4 b;   /* lo0.z hi0.z lo1.x hi1.x */
    float4 c;   /* lo1.y hi1.y lo1.z hi1.z */
    int4 d;     /* child0 child1 - - */
};

/* 8-bit quantised BVH4 node, 64 B. Child boxes are stored relative to the node's origin
 * with a per-axis power-of-two scale; lo rounds down and hi rounds up, so every dequantised
 * box contains the FP32 box it encodes. */
struct __attribute__((aligned(16))) QNode {
    float ox, oy, oz;       /* origin                                  bytes  0..11 */
    uint8_t ex, ey, ez;     /* scale = 2^(e-127)                       bytes 12..14 */
    uint8_t nchild;         /*                                         byte  15     */
    uint8_t qlo[3][4];      /* [axis][child]                           bytes 16..27 */
    uint8_t qhi[3][4];      /*                                         bytes 28..39 */
    uint32_t pad[2];        /*                                         bytes 40..47 */
    int32_t child[4];       /* BNode child encoding; unused slots = ILM_NO_CHILD, bytes 48..63 */
};
static_assert(sizeof(QNode) == 64, "QNode is one 64-byte line");

#define ILM_NO_CHILD ((int32_t)0x7fffffff)
#define ILM_LEAF_MAX 16
__host__ __device__ inline int32_t ilm_leaf_code(uint32_t first, uint32_t count) {
    return (int32_t)(0x80000000u | ((count - 1u) << 27) | first);
}
__host__ __device__ inline uint32_t ilm_leaf_first(int32_t c) { return (uint32_t)c & 0x07ffffffu; }
__host__ __device__ inline uint32_t ilm_leaf_count(int32_t c) { return (((uint32_t)c >> 27) & 0xfu) + 1u; }

/* ---------------- scene ---------------- */
struct DTex {
    int type;
    int w, h, hdr;
    float a[3];
    float b[3];
    float scale[2];
    const float4* f;
    const uchar4* u8;
};

struct DMedium {
    float absorption[3];
    float scattering;
    float g;
    int active;
};

struct DMat {
    int type;
    int medium;
    int distribution;
    int pad;
    float ior;
    float diffuse_fresnel;
    float eta[3];
    float k[3];
    DTex albedo;
    DTex roughness;
};

struct DInst {
    float inv[12];      /* rows 0..2 of the inverse (world -> object) */
    float tinv[12];     /* rows 0..2 of the inverse-transpose (normals object -> world) */
    int mesh;
    int material;
    int flags;          /* ILM_INST_* copied from the mesh */
    int pad;
    /* the mesh's BLAS, copied from its DMesh so that entering the instance reads this record
     * only (bytes 112..143: one 32-byte span next to the inverse rows) */
    const BNode* nodes;
    const QNode* qnodes;
    const float4* tris;
    int root;
    int qroot;
    /* the mesh's data for shading (bytes 144..175): the shading vertices of each triangle, then
     * texture coordinates (with idx, only for meshes that have them) */
    const float4* svert;
    const float* uv;
    const int* idx;
    const float* pos;
};
static_assert(sizeof(DInst) == 176, "DInst is 11 16-byte words");
#define ILM_INST_FACE_NORMALS 1
#define ILM_INST_TEXCOORDS 2

struct DMesh {
    const BNode* nodes;
    const QNode* qnodes;
    const float4* tris;     /* 3 per leaf slot in leaf order: (v0, original index bits), (v1, 0), (v2, 0) */
    const int* idx;         /* original triangle vertex indices, 3 per triangle */
    const float* pos;       /* 3 per vertex */
    const float* nrm;       /* 3 per vertex, or null */
    const float* uv;        /* 2 per vertex, or null */
    const float4* svert;    /* 3 per original triangle: its vertex normals, or with face normals its
                               vertex positions (w = 0), so shading reads them in one request */
    int root;               /* root child code (FP32 layout) */
    int qroot;              /* root child code (quantised layout) */
    int face_normals;
    int has_texcoords;
};

struct DSphere {
    float c[3];
    float r;
    int material;
    int pad[3];
};

struct DScene {
    const BNode* tlas;
    const QNode* qtlas;     /* TLAS leaves (one instance each) hold the instance index as `first` */
    int tlas_root;
    int qtlas_root;
    int n_inst;
    int n_spheres;
    const DInst* inst;
    const DMesh* meshes;
    const DSphere* spheres;
    const DMat* mats;
    const DMedium* mediums;
    DTex env;
    float cam[12];          /* rows 0..2 of the camera matrix */
    float dof;
    float aperture;
    int pad[2];
};

/* One request inside a launch. */
/* Exact unsigned division by a divisor fixed on the host: with M = ceil(2^64 / d),
 * floor(n / d) = floor(n * M / 2^64) for every 32-bit n, because the excess n * (M - 2^64/d) / 2^64
 * is below 2^-32 and so below 1/d, the smallest gap between n/d and the next integer. d = 1 is
 * stored as M = 0. Replaces a ~30-instruction integer division by three. */
struct UDiv {
    uint32_t lo, hi;
};
inline UDiv udiv_make(uint32_t d) {
    UDiv m;
    const uint64_t M = d > 1 ? ~0ull / d + 1 : 0ull;
    m.lo = (uint32_t)M;
    m.hi = (uint32_t)(M >> 32);
    return m;
}

struct DReq {
    const DScene* scene;
    int width, height;
    int max_depth;
    float ray_bias;
    float image_width;
    float pixel_size, left, top;
    int tonemap;
    int spp;
    float4* accum;          /* width*height Kahan sums */
    float4* comp;           /* width*height Kahan compensations */
    /* pixel walk of this chunk (tile_pixel with tile_th rows) and its divisors */
    uint32_t tile_th, tile_full;
    UDiv w_div, band_div, th_div;
};

/* One slice of work in a chunk: passes [pass0, pass0+npass) of pixel slots [px0, px0+npx) of
 * request req. Sample s of slot px0+q has pid = pid_base + slice_local(s, q, npx, sgroup); the
 * slot's pixel is tile_pixel(px0+q, width, height, tile_rows). */
struct DWork {
    int req;
    int pass0;
    int npass;
    uint32_t px0;
    uint32_t npx;
    uint32_t pid_base;      /* first path id of this slice */
    uint32_t acc_base;      /* first accumulate index of this slice (prefix of npx) */
    uint32_t sgroup;        /* consecutive path ids per pixel (power of two dividing npass) */
    uint32_t sg_log2;       /* log2(sgroup) */
    UDiv span_div;          /* divides by npx * sgroup */
};

#define ILM_MAX_JOBS 256
#define ILM_MAX_WORK 512

/* Everything that changes from one chunk to the next, in one device-resident block, so the
 * kernels of a chunk can be launched with fixed arguments (and captured once as a graph). */
struct DChunk {
    uint32_t total_paths;
    uint32_t total_px;
    int n_work;
    int depth_steps;
    DReq reqs[ILM_MAX_JOBS];
    DWork work[ILM_MAX_WORK];
};

/* Per-chunk control block (device). Cleared by k_ctl_init, never by a memset. */
struct DCtl {
    uint32_t count[ILM_MAX_DEPTH + 2];   /* live paths entering depth d */
    uint32_t fetch[ILM_MAX_DEPTH + 2];   /* extend work cursor at depth d */
    unsigned long long segments;
    uint32_t total;
    uint32_t pad;
};

/* Path state, struct of arrays; two copies ping-pong across depths. */
struct DPaths {
    float4* o;          /* origin, w = bits of (req << 16 | (medium + 1)) */
    float4* d;          /* direction, w = bits of pixel */
    float4* c;          /* throughput, w = bits of pid */
    uint2* rng;         /* PCG32 state */
};

struct DHits {
    float4* h;          /* t, u, v, bits of prim (original triangle index) */
    int* inst;          /* instance of the hit: -2-s for sphere s, -1 for a miss */
};

/* Everything a kernel launch needs. Passed by value. */
struct KParams {
    const DChunk* chunk;
    DCtl* ctl;
    DPaths in;
    DPaths out;
    DHits hits;
    float4* contrib;        /* per pid radiance (zero when the path ends without emission) */
    int* seg_count;         /* per pid segment counts (debug), or null */
    const uint32_t* list;   /* explicit (pixel, pass) pairs for the debug raygen, or null */
    unsigned long long* trav_stats;   /* [0] child-box tests, [1] triangle tests, or null */
    int bvh_layout;
    int tri_test;
    int extend_mode;        /* ILM_EXTEND_BATCH, _LEAF_REFILL or _STEP_REFILL */
    int use_ovf;            /* 1 when a scene's stack bound exceeds the LDS stack depth */
    int tile_rows;          /* pixel order: bands of this many rows walked column-major (1 = rows) */
    int refill_min;         /* extend modes 1-2: idle lanes that trigger a refill (1..64) */
    int ext_fetch;          /* rays reserved per queue atomic by an extend wave (multiple of 64) */
    int leaf_pct;           /* per-step refill: leaf trips when leaf lanes >= leaf_pct% of the others */
    int shade_sort;         /* shade each block's paths in shading-key order (staged in LDS) */
    int cam0;               /* the first depth forms camera rays in extend and shade (no raygen) */
    int packet0;            /* ... and its extend traces each wave's camera rays as one packet */
    int uni_scene;          /* every request of the chunk renders the same scene */
};

/* Path id order of a slice: local index l <-> (sample k of the slice, pixel slot q). Samples
 * come in groups of g consecutive ids per pixel: l = (k / g) * npx * g + q * g + k % g. */
__host__ __device__ inline uint32_t slice_local(uint32_t k, uint32_t q, uint32_t npx, uint32_t g) {
    return (k / g) * npx * g + q * g + (k % g);
}
__host__ __device__ inline void slice_decode(uint32_t l, uint32_t npx, uint32_t g, uint32_t& k, uint32_t& q) {
    const uint32_t span = npx * g;
    const uint32_t kb = l / span, r = l - kb * span;
    q = r / g;
    k = kb * g + (r - q * g);
}
/* Pixel slot i (0 <= i < w*h) -> pixel index: bands of th rows, each walked column by column
 * (th consecutive slots cover one column of the band); rows past the last full band in order. */
__host__ __device__ inline uint32_t tile_pixel(uint32_t i, uint32_t w, uint32_t h, uint32_t th) {
    if (th <= 1) return i;
    const uint32_t band_px = th * w, full = (h / th) * band_px;
    if (i >= full) return i;
    const uint32_t band = i / band_px, r = i - band * band_px;
    const uint32_t x = r / th, y = band * th + (r - x * th);
    return y * w + x;
}

/* ---------------- vector helpers (reference operation order) ---------------- */
struct V3 { float x, y, z; };
__host__ __device__ inline V3 mk3(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
__host__ __device__ inline V3 add3(V3 a, V3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__host__ __device__ inline V3 adds3(V3 a, float s) { return mk3(s + a.x, s + a.y, s + a.z); }
__host__ __device__ inline V3 sub3(V3 a, V3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__host__ __device__ inline V3 mul3(V3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
__host__ __device__ inline V3 mulv3(V3 a, V3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
__host__ __device__ inline V3 divs3(V3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
__host__ __device__ inline V3 divv3(V3 a, V3 b) { return mk3(a.x / b.x, a.y / b.y, a.z / b.z); }
__host__ __device__ inline float dot3(V3 a, V3 b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
__host__ __device__ inline V3 cross3(V3 a, V3 b) {
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__host__ __device__ inline float len2_3(V3 v) { return (v.x * v.x) + (v.y * v.y) + (v.z * v.z); }
__host__ __device__ inline V3 normalize3(V3 v) {
    float inv = 1.0f / __builtin_sqrtf(len2_3(v));
    return mk3(v.x * inv, v.y * inv, v.z * inv);
}
__host__ __device__ inline V3 reflect3(V3 v, V3 n) { return sub3(v, mul3(mul3(n, dot3(v, n)), 2.0f)); }
__host__ __device__ inline V3 to_basis3(V3 v, V3 n) {
    V3 t;
    if (pm_fabs(n.x) > pm_fabs(n.y)) t = mk3(n.z, 0.0f, -n.x);
    else t = mk3(0.0f, n.z, -n.y);
    V3 b = cross3(n, t);
    t = cross3(b, n);
    V3 r = add3(add3(mul3(t, v.x), mul3(n, v.y)), mul3(b, v.z));
    return normalize3(r);
}
/* rows 0..2 of a 4x4 matrix times (v, w), src/math/matrix4.cu:203-209 */
__host__ __device__ inline V3 xform3(const float* m, V3 v, float w) {
    return mk3(m[0] * v.x + m[1] * v.y + m[2] * v.z + m[3] * w,
               m[4] * v.x + m[5] * v.y + m[6] * v.z + m[7] * w,
               m[8] * v.x + m[9] * v.y + m[10] * v.z + m[11] * w);
}
