/*
 * kernels.hip - the gfx950 wavefront pipeline of the path tracer.
 *
 *   k_ctl_init     clear the chunk's queue counters (no memset: graph-safe)
 *   k_raygen       camera rays for (request, pass, pixel) work, PCG32 seeded per path
 *                  (src/renderer.cu:61-69,81-110)
 *   k_extend       closest hit (traverse.h): spheres, then a TLAS over instances and per-mesh
 *                  BLAS, stack walk in wave64 with the stack in LDS, t-pruned and ordered
 *                  (src/renderer.cu:112-137, src/primitives/mesh.cu:479-562)
 *   k_shade        medium free flight and the 9 BSDFs (src/renderer.cu:139-388); its epilogue
 *                  is the compaction: wave ballot + one atomic per wave appends survivors to
 *                  the next depth's queue (replaces the host round trip of
 *                  src/renderer.cu:390-424)
 *   k_accumulate   Kahan-compensated per-pixel sum of the chunk's path radiances, in pass order
 *   k_chunk_end    segment count (src/renderer.cu:554)
 *   k_tonemap      src/renderer.cu:426-441 with src/tonemapper.cu:42-76
 *
 * Every kernel reads its work count from device memory, so a whole chunk is a fixed launch
 * sequence (capturable as a hipGraph) with no host synchronisation between depths.
 */
#include "traverse.h"

#ifndef SHADE_BLOCK
#define SHADE_BLOCK 512                       /* 8 waves share one compaction atomic */
#endif
/* Waves per SIMD the shading kernel's registers must allow: 6 caps them at 80 without spills,
 * so three 512-thread blocks stay resident per CU instead of two */
#ifndef SHADE_WPE
#define SHADE_WPE 6
#endif
#if SHADE_WPE > 0
#define SHADE_BOUNDS __launch_bounds__(SHADE_BLOCK) __attribute__((amdgpu_waves_per_eu(SHADE_WPE, 8)))
#else
#define SHADE_BOUNDS __launch_bounds__(SHADE_BLOCK)
#endif
#define SHADE_WAVES (SHADE_BLOCK / 64)

__global__ void k_ctl_init(DCtl* ctl_, const DChunk* chunk) {
    GLOBAL DCtl* ctl = gpw(ctl_);
    for (int i = threadIdx.x; i < ILM_MAX_DEPTH + 2; i += blockDim.x) {
        ctl->count[i] = 0u;
        ctl->fetch[i] = 0u;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t total = gp(chunk)->total_paths;
        ctl->count[0] = total;
        ctl->total = total;
        ctl->segments = 0ull;
    }
}

/* src/renderer.cu:81-110 (and the per-pass seeding of :61-69): the camera rays written out, for
 * the extend modes that read the first depth's rays from memory */
__global__ __launch_bounds__(256) void k_raygen(KParams P) {
    const uint32_t total = gp(P.chunk)->total_paths;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        float4 o4, d4;
        uint2 rg;
        camera_ray(P, i, o4, d4, rg);
        gpw(P.in.o)[i] = o4;
        gpw(P.in.d)[i] = d4;
        gpw(P.in.c)[i] = make_float4(1.0f, 1.0f, 1.0f, u2f(i));
        gpw(P.in.rng)[i] = rg;
        gpw(P.contrib)[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (P.seg_count) gpw(P.seg_count)[i] = 0;
    }
}

/* Per-path results cleared for a chunk whose first depth forms its camera rays in place. */
__global__ __launch_bounds__(256) void k_path_clear(KParams P) {
    const uint32_t total = gp(P.chunk)->total_paths;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        gpw(P.contrib)[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (P.seg_count) gpw(P.seg_count)[i] = 0;
    }
}

enum { ST_ALIVE = 0, ST_DEAD = 1, ST_EMIT = 2 };

/* src/tonemapper.cu:33-76 */
__device__ inline V3 uncharted2(V3 x) {
    const float A = 0.15f, Bc = 0.50f, C = 0.10f, D = 0.20f, E = 0.02f, F = 0.30f;
    V3 num = adds3(mulv3(x, adds3(mul3(x, A), C * Bc)), D * E);
    V3 den = adds3(mulv3(x, adds3(mul3(x, A), Bc)), D * F);
    return adds3(divv3(num, den), -E / F);
}
__device__ inline V3 pow3(V3 v, float p) { return mk3(pm_powf(v.x, p), pm_powf(v.y, p), pm_powf(v.z, p)); }
__device__ inline V3 tonemap_eval(int op, V3 in) {
    const float gamma = 1.0f / 2.2f;
    switch (op) {
    case ILM_TONEMAP_REINHARD: return pow3(divv3(in, adds3(in, 1.0f)), gamma);
    case ILM_TONEMAP_FILMIC: {
        V3 x = adds3(in, -0.004f);
        x = mk3(__builtin_fmaxf(x.x, 0.0f), __builtin_fmaxf(x.y, 0.0f), __builtin_fmaxf(x.z, 0.0f));
        return divv3(mulv3(x, adds3(mul3(x, 6.2f), 0.5f)), adds3(mulv3(x, adds3(mul3(x, 6.2f), 1.7f)), 0.06f));
    }
    case ILM_TONEMAP_UNCHARTED: {
        const float Wt = 11.2f, exposure_bias = 2.0f;
        V3 curr = uncharted2(mul3(in, exposure_bias));
        V3 white = divv3(mk3(1.0f, 1.0f, 1.0f), uncharted2(mk3(Wt, Wt, Wt)));
        return pow3(mulv3(curr, white), gamma);
    }
    default: return pow3(in, gamma);
    }
}

__device__ inline V3 conductor3(const GLOBAL DMat* m, float cosI) {
    return mk3(fresnel_conductor(m->eta[0], m->k[0], cosI), fresnel_conductor(m->eta[1], m->k[1], cosI),
               fresnel_conductor(m->eta[2], m->k[2], cosI));
}

/* Scene arrays a shading step reads: per path from its scene, or once per block for a chunk
 * of one scene. */
struct ShadeRefs {
    const GLOBAL DInst* inst;
    const GLOBAL DMat* mats;
    const GLOBAL DMedium* mediums;
};

/* One path segment after its closest hit: src/renderer.cu:152-386. */
__device__ inline int shade_path(const GLOBAL DScene* S, const ShadeRefs& SR, const float ray_bias, const float4 h4, int hinst, V3& ro, V3& rd, V3& color, int& medium, uint64_t& rs, uint64_t inc) {
    const float tmin = h4.x;
    const bool is = hinst != -1;
    if (medium >= 0) {
        const GLOBAL DMedium* md = SR.mediums + medium;
        if (md->active) {
            const float sig_s = md->scattering;
            const float a0 = md->absorption[0], a1 = md->absorption[1], a2 = md->absorption[2];
            float u = ilm_pcg32_uniform(&rs, inc);
            float scatter_dist = -pm_logf(u) / sig_s;
            if (scatter_dist < tmin) {
                float u1 = ilm_pcg32_uniform(&rs, inc);
                float u2 = ilm_pcg32_uniform(&rs, inc);
                V3 sdir = to_basis3(sample_hg(md->g, u1, u2), rd);
                ro = add3(ro, mul3(rd, scatter_dist));
                rd = normalize3(sdir);
                color = mulv3(color, mk3(pm_expf(-1.0f * scatter_dist * a0), pm_expf(-1.0f * scatter_dist * a1),
                                         pm_expf(-1.0f * scatter_dist * a2)));
                if ((double)len2_3(color) < 1e-4) return ST_DEAD;
                return ST_ALIVE;
            }
            color = mulv3(color, mk3(pm_expf(-1.0f * tmin * a0), pm_expf(-1.0f * tmin * a1), pm_expf(-1.0f * tmin * a2)));
        }
    }
    if (!is) {
        /* src/renderer.cu:378-386 */
        float u = (float)((double)(pm_atan2f(rd.z, rd.x) * 0.5f) * ILM_INV_PI_D + 0.5f);
        float v = (float)(0.5f - (double)pm_asinf(rd.y) * ILM_INV_PI_D);
        color = mulv3(color, tex_eval(&S->env, u, v));
        return ST_EMIT;
    }
    /* surface attributes: src/primitives/mesh.cu:532-561, mesh_instance.cu:102-107, sphere.cu:71-76 */
    V3 n;
    float hu, hv;
    int mat;
    if (hinst >= 0) {
        /* the instance record's flags, material and shading pointers are requested together
         * (bytes 96..111 and 144..159), and the triangle's three shading vertices then come in
         * one request: the normal waits on two round trips, not on the record, a mesh record,
         * the index array and the vertex arrays */
        const GLOBAL DInst* I = SR.inst + hinst;
        const GLOBAL int4* iw = (const GLOBAL int4*)I;
        const int4 misc = iw[6], vp0 = iw[9];
        asm volatile("" ::"v"(misc.x), "v"(vp0.x));
        const int flags = misc.z;
        const GLOBAL float4* sv = (const GLOBAL float4*)(((uint64_t)(uint32_t)vp0.y << 32) | (uint32_t)vp0.x);
        const int prim = __float_as_int(h4.w);
        const float bu = h4.y, bv = h4.z;
        const float4 s0 = sv[3 * prim], s1 = sv[3 * prim + 1], s2 = sv[3 * prim + 2];
        if (flags & ILM_INST_FACE_NORMALS) {
            /* cross(v1 - v0, v2 - v0) of the triangle's vertices, as src/primitives/mesh.cu:541 */
            V3 p0 = mk3(s0.x, s0.y, s0.z);
            V3 p1 = mk3(s1.x, s1.y, s1.z);
            V3 p2 = mk3(s2.x, s2.y, s2.z);
            n = normalize3(cross3(sub3(p1, p0), sub3(p2, p0)));
        } else {
            V3 n0 = mk3(s0.x, s0.y, s0.z);
            V3 n1 = mk3(s1.x, s1.y, s1.z);
            V3 n2 = mk3(s2.x, s2.y, s2.z);
            n = normalize3(add3(mul3(n0, 1.0f - bu - bv), add3(mul3(n1, bu), mul3(n2, bv))));
        }
        if (flags & ILM_INST_TEXCOORDS) {
            const GLOBAL float* uvp = (const GLOBAL float*)(((uint64_t)(uint32_t)vp0.w << 32) | (uint32_t)vp0.z);
            const GLOBAL int* idx = gp(I->idx);
            const int i0 = idx[3 * prim], i1 = idx[3 * prim + 1], i2 = idx[3 * prim + 2];
            float w = 1.0f - bu - bv;
            hu = (uvp[2 * i0] * w + uvp[2 * i1] * bu) + uvp[2 * i2] * bv;
            hv = (uvp[2 * i0 + 1] * w + uvp[2 * i1 + 1] * bu) + uvp[2 * i2 + 1] * bv;
        } else {
            hu = bu;
            hv = bv;
        }
        float ti[12];
        load_rows12(I->tinv, ti);
        n = normalize3(xform3(ti, n, 0.0f));
        mat = misc.y;
    } else {
        const GLOBAL DSphere* sp = gp(S->spheres) + (-2 - hinst);
        V3 pos = add3(ro, mul3(rd, tmin));
        n = normalize3(sub3(pos, mk3(sp->c[0], sp->c[1], sp->c[2])));
        hu = (float)((double)(pm_atan2f(n.z, n.x) * 0.5f) * ILM_INV_PI_D + 0.5f);
        hv = (float)(0.5f - (double)pm_asinf(n.y) * ILM_INV_PI_D);
        mat = sp->material;
    }
    const GLOBAL DMat* m = SR.mats + mat;
    const int type = m->type;
    const V3 d = rd;
    V3 new_dir;
    V3 norm_o = mul3(n, dot3(n, d) > 0 ? -1.0f : 1.0f);
    V3 new_origin = add3(ro, mul3(d, tmin));
    V3 albedo = tex_eval(&m->albedo, hu, hv);
    const float bias = ray_bias;
    switch (type) {
    case ILM_MAT_EMISSIVE:
        color = mulv3(color, albedo);
        return ST_EMIT;
    case ILM_MAT_DIFFUSE: {
        color = mulv3(color, albedo);
        float u1 = ilm_pcg32_uniform(&rs, inc);
        float u2 = ilm_pcg32_uniform(&rs, inc);
        new_dir = to_basis3(sample_hemisphere_cosine(u1, u2), norm_o);
        new_origin = add3(new_origin, mul3(norm_o, bias));
        break;
    }
    case ILM_MAT_SPECULAR:
        color = mulv3(color, mk3(0.99f, 0.99f, 0.99f));
        new_dir = reflect3(d, norm_o);
        new_origin = add3(new_origin, mul3(norm_o, bias));
        break;
    case ILM_MAT_REFRACTIVE: {
        const float ior = m->ior;
        float cosI = -dot3(d, n);
        float cosT = 0.0f;
        float F = fresnel_dielectric(cosI, ior, cosT);
        if (F == 1.0f || ilm_pcg32_uniform(&rs, inc) < F) {
            new_dir = reflect3(d, n);
            new_origin = add3(new_origin, mul3(norm_o, bias));
        } else {
            float eta = cosI < 0.0f ? ior : 1.0f / ior;
            medium = cosI > 0.0f ? m->medium : -1;
            new_dir = add3(mul3(d, eta), mul3(norm_o, eta * pm_fabs(cosI) - cosT));
            new_origin = add3(new_origin, mul3(norm_o, -bias));
        }
        color = mulv3(color, albedo);
        break;
    }
    case ILM_MAT_ROUGHREFLECTIVE: {
        const int dist = m->distribution;
        const float roughness = tex_eval(&m->roughness, hu, hv).x;
        V3 wi = mul3(d, -1.0f);
        float wiDotN = dot3(wi, n);
        float a = mf_alpha(dist, roughness, wiDotN);
        float u1 = ilm_pcg32_uniform(&rs, inc);
        float u2 = ilm_pcg32_uniform(&rs, inc);
        V3 mm = to_basis3(mf_sample(dist, a, u1, u2), n);
        float wiDotT = 0.0f;
        float wiDotM = dot3(wi, mm);
        float F = fresnel_dielectric(wiDotM, m->ior, wiDotT);
        new_dir = reflect3(d, mm);
        if (wiDotN * dot3(new_dir, n) <= 0.0f) return ST_DEAD;
        new_origin = add3(new_origin, mul3(norm_o, bias));
        float G = mf_G(dist, wi, new_dir, mm, n, a);
        float weight = (F * G * pm_fabs(wiDotM)) / (pm_fabs(wiDotN) * pm_fabs(dot3(mm, n)));
        color = mulv3(color, mul3(albedo, weight));
        break;
    }
    case ILM_MAT_ROUGHREFRACTIVE: {
        const int dist = m->distribution;
        const float ior = m->ior;
        const float roughness = tex_eval(&m->roughness, hu, hv).x;
        V3 wi = mul3(d, -1.0f);
        float wiDotN = dot3(wi, n);
        float a = mf_alpha(dist, roughness, wiDotN);
        float u1 = ilm_pcg32_uniform(&rs, inc);
        float u2 = ilm_pcg32_uniform(&rs, inc);
        V3 mm = to_basis3(mf_sample(dist, a, u1, u2), n);
        float wiDotT = 0.0f;
        float wiDotM = dot3(wi, mm);
        float F = fresnel_dielectric(wiDotM, ior, wiDotT);
        if (F == 1.0f || ilm_pcg32_uniform(&rs, inc) < F) {
            new_dir = reflect3(d, mm);
            if (wiDotN * dot3(new_dir, n) <= 0.0f) return ST_DEAD;
            new_origin = add3(new_origin, mul3(norm_o, bias));
        } else {
            float eta = wiDotM < 0.0f ? ior : 1.0f / ior;
            medium = wiDotM > 0.0f ? m->medium : -1;
            new_dir = sub3(mul3(mm, wiDotM * eta - (wiDotM > 0.0f ? 1.0f : -1.0f) * wiDotT), mul3(wi, eta));
            if (wiDotN * dot3(new_dir, n) >= 0.0f) return ST_DEAD;
            new_origin = add3(new_origin, mul3(norm_o, -bias));
        }
        float G = mf_G(dist, wi, new_dir, mm, n, a);
        float weight = (G * pm_fabs(wiDotM)) / (pm_fabs(wiDotN) * pm_fabs(dot3(mm, n)));
        color = mulv3(color, mul3(albedo, weight));
        break;
    }
    case ILM_MAT_CONDUCTOR: {
        float cosI = -dot3(d, n);
        if (cosI <= 0) return ST_DEAD;
        new_dir = reflect3(d, norm_o);
        new_origin = add3(new_origin, mul3(norm_o, bias));
        color = mulv3(color, conductor3(m, cosI));
        break;
    }
    case ILM_MAT_ROUGHCONDUCTOR: {
        const int dist = m->distribution;
        V3 wi = mul3(d, -1.0f);
        float wiDotN = dot3(wi, n);
        if (wiDotN <= 0) return ST_DEAD;
        const float roughness = tex_eval(&m->roughness, hu, hv).x;
        float a = mf_alpha(dist, roughness, wiDotN);
        float u1 = ilm_pcg32_uniform(&rs, inc);
        float u2 = ilm_pcg32_uniform(&rs, inc);
        V3 mm = to_basis3(mf_sample(dist, a, u1, u2), n);
        V3 F = conductor3(m, wiDotN);
        new_dir = reflect3(d, mm);
        if (wiDotN * dot3(new_dir, n) <= 0.0f) return ST_DEAD;
        float wiDotM = dot3(wi, mm);
        float G = mf_G(dist, wi, new_dir, mm, n, a);
        float weight = (G * pm_fabs(wiDotM)) / (pm_fabs(wiDotN) * pm_fabs(dot3(mm, n)));
        color = mulv3(color, mul3(F, weight));
        new_origin = add3(new_origin, mul3(mm, bias));
        break;
    }
    case ILM_MAT_PLASTIC: {
        const float ior = m->ior;
        float cosI = -dot3(d, n);
        float cosT = 0.0f;
        float F = fresnel_dielectric(cosI, ior, cosT);
        float u1 = ilm_pcg32_uniform(&rs, inc);
        if (F == 1.0f || u1 < F) {
            new_dir = reflect3(d, n);
            color = mulv3(color, mk3(0.99f, 0.99f, 0.99f));
        } else {
            float u2 = ilm_pcg32_uniform(&rs, inc);
            float uu = (float)((double)(u1 - F) / (1.0 - (double)F));
            new_dir = to_basis3(sample_hemisphere_cosine(uu, u2), norm_o);
            float cosO = dot3(new_dir, n);
            float Fo = fresnel_dielectric(cosO, ior, cosT);
            V3 diff = divs3(albedo, 1.0f - m->diffuse_fresnel);
            float inv_eta = 1.0f / ior;
            color = mulv3(color, mul3(diff, inv_eta * inv_eta * (1 - Fo)));
        }
        new_origin = add3(new_origin, mul3(norm_o, bias));
        break;
    }
    default:
        return ST_DEAD;
    }
    ro = new_origin;
    rd = normalize3(new_dir);
    return ST_ALIVE;
}

/* Shading key for the block-local sort: paths inside a medium, misses, spheres and instances
 * each get their own buckets (an instance stands in for its material), so a wave mostly runs
 * one material's code. INVALID sorts the block's empty slots last. */
/* Stable split of a shading block's threads by key (0 <= key < NK): the thread's position when
 * the block is ordered by key, then wave, then lane. Each wave counts its keys with one ballot
 * per distinct key present (usually one to three); wave 0 turns the waves x keys counts into
 * start positions. Every thread of the block calls it with all lanes active. On return
 * s_wk[k] (wave 0's entry for key k) is the number of threads with a key below k. */
template <int NK>
__device__ inline uint32_t block_split(uint32_t key, uint32_t* s_wk) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const unsigned long long lt = (1ull << lane) - 1ull;
    if (lane < NK) s_wk[wid * NK + lane] = 0u;
    unsigned long long rem = __ballot(1);
    uint32_t rank = 0;
    while (rem) {
        const int leader = __builtin_ctzll(rem);
        const uint32_t k = (uint32_t)__builtin_amdgcn_readlane((int)key, leader);
        const unsigned long long m = __ballot(key == k);
        if (key == k) rank = (uint32_t)__popcll(m & lt);
        if (lane == leader) s_wk[wid * NK + k] = (uint32_t)__popcll(m);
        rem &= ~m;
    }
    __syncthreads();
    if (wid == 0) {
        uint32_t col[SHADE_WAVES];
        uint32_t tot = 0;
#pragma unroll
        for (int w = 0; w < SHADE_WAVES; w++) {
            col[w] = lane < NK ? s_wk[w * NK + lane] : 0u;
            tot += col[w];
        }
        uint32_t incl = tot;
#pragma unroll
        for (int d = 1; d < 32; d <<= 1) {
            const uint32_t v = (uint32_t)__shfl_up((int)incl, d);
            if (lane >= d) incl += v;
        }
        uint32_t run = incl - tot;
        if (lane < NK) {
#pragma unroll
            for (int w = 0; w < SHADE_WAVES; w++) { s_wk[w * NK + lane] = run; run += col[w]; }
        }
    }
    __syncthreads();
    return s_wk[wid * NK + key] + rank;
}

#define SHADE_KEYS 18
#define SHADE_KEY_INVALID 17
__device__ inline uint32_t shade_key(int medium, int hinst) {
    /* inside a medium, a ray whose capped walk found no surface scatters (the cap is the
     * scatter distance), one that found a surface usually reaches it: two different code paths */
    if (medium >= 0) return hinst == -1 ? 15u : 16u;
    if (hinst == -1) return 14u;
    if (hinst < -1) return 9u + ((uint32_t)(-2 - hinst) & 3u);
    return (uint32_t)hinst % 9u;
}

template <int SORT, int CAM>
__global__ SHADE_BOUNDS void k_shade(KParams P, int depth) {
    /* survivors of the block's 512 paths get their queue slots from one atomic: per-wave counts
     * meet in LDS, wave 0 lane 0 reserves the block's range, each wave takes its offset */
    __shared__ uint32_t wave_n[SHADE_WAVES];
    __shared__ uint32_t block_base;
    /* SORT: the block's path states are staged in LDS and shaded in key order */
    constexpr int SB = SORT ? SHADE_BLOCK : 1;
    __shared__ float4 s_o[SB], s_d[SB], s_c[SB], s_h[SB];
    __shared__ uint2 s_r[SB];
    __shared__ int s_i[SB];
    __shared__ uint16_t s_ord[SB];
    /* survivors leave grouped by direction octant when sorting, and at the first depth */
    constexpr bool OCT = SORT || CAM;
    __shared__ uint32_t s_wk[OCT ? SHADE_WAVES * SHADE_KEYS : 1];
    GLOBAL DCtl* ctl = gpw(P.ctl);
    const uint32_t count = ctl->count[depth];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const GLOBAL DReq* reqs = gp(P.chunk)->reqs;
    /* a chunk of one scene: its arrays are read once per block */
    const GLOBAL DScene* Su = nullptr;
    ShadeRefs SRu = {nullptr, nullptr, nullptr};
    if (P.uni_scene) {
        Su = uni_p(gp(reqs[0].scene));
        SRu.inst = uni_p(gp(Su->inst));
        SRu.mats = uni_p(gp(Su->mats));
        SRu.mediums = uni_p(gp(Su->mediums));
    }
    /* first depth of a chunk of one work item: its camera values are read once per wave */
    CamSrc cs;
    const bool cu = CAM && camsrc_load(P, cs);
    const GLOBAL float4* ino = gp(P.in.o);
    const GLOBAL float4* ind = gp(P.in.d);
    const GLOBAL float4* inc4 = gp(P.in.c);
    const GLOBAL uint2* inr = gp(P.in.rng);
    const GLOBAL float4* hh = gp(P.hits.h);
    const GLOBAL int* hinst_in = gp(P.hits.inst);
    GLOBAL float4* outo = gpw(P.out.o);
    GLOBAL float4* outd = gpw(P.out.d);
    GLOBAL float4* outc = gpw(P.out.c);
    GLOBAL uint2* outr = gpw(P.out.rng);
    /* the loop bound depends only on blockIdx and count, so every thread of the block runs the
     * same trips and reaches every barrier */
    for (uint32_t tbase = blockIdx.x * (uint32_t)SHADE_BLOCK; tbase < count; tbase += gridDim.x * (uint32_t)SHADE_BLOCK) {
        uint32_t i = tbase + threadIdx.x;
        bool valid = i < count;
        float4 o4 = make_float4(0, 0, 0, 0), d4 = o4, c4 = o4, h4 = o4;
        uint2 rg = make_uint2(0, 0);
        int hinst = -1;
        if (valid) {
            if (CAM) {
                if (cu) camera_ray_u(camu_fetch(cs), i, o4, d4, rg);
                else camera_ray(P, i, o4, d4, rg);
                c4 = make_float4(1.0f, 1.0f, 1.0f, u2f(i));
            } else {
                o4 = ino[i]; d4 = ind[i]; c4 = inc4[i];
                rg = inr[i];
            }
            h4 = hh[i];
            hinst = hinst_in[i];
        }
        if (SORT) {
            const uint32_t key = valid ? shade_key((int)(f2u(o4.w) & 0xffffu) - 1, hinst) : SHADE_KEY_INVALID;
            const int t = threadIdx.x;
            s_o[t] = o4; s_d[t] = d4; s_c[t] = c4; s_h[t] = h4; s_i[t] = hinst; s_r[t] = rg;
            s_ord[block_split<SHADE_KEYS>(key, s_wk)] = (uint16_t)t;
            __syncthreads();
            const int j = s_ord[t];
            o4 = s_o[j]; d4 = s_d[j]; c4 = s_c[j]; h4 = s_h[j]; hinst = s_i[j]; rg = s_r[j];
            valid = tbase + (uint32_t)j < count;
        }
        int status = ST_DEAD;
        float4 oo = make_float4(0, 0, 0, 0), od = oo, oc = oo;
        uint2 org = make_uint2(0, 0);
        if (valid) {
            uint32_t tag = f2u(o4.w);
            int req = (int)(tag >> 16);
            int medium = (int)(tag & 0xffffu) - 1;
            uint32_t pixel = f2u(d4.w);
            uint32_t pid = f2u(c4.w);
            const GLOBAL DReq* R = reqs + req;
            const GLOBAL DScene* S = P.uni_scene ? Su : gp(R->scene);
            ShadeRefs SR = SRu;
            if (!P.uni_scene) { SR.inst = gp(S->inst); SR.mats = gp(S->mats); SR.mediums = gp(S->mediums); }
            uint64_t rs = ((uint64_t)rg.y << 32) | (uint64_t)rg.x;
            V3 ro = mk3(o4.x, o4.y, o4.z), rd = mk3(d4.x, d4.y, d4.z), color = mk3(c4.x, c4.y, c4.z);
            if (P.seg_count) gpw(P.seg_count)[pid] += 1;
            status = shade_path(S, SR, R->ray_bias, h4, hinst, ro, rd, color, medium, rs, ilm_pcg32_inc(pixel));
            if (status == ST_EMIT) gpw(P.contrib)[pid] = make_float4(color.x, color.y, color.z, 0.0f);
            if (status == ST_ALIVE && depth + 1 >= R->max_depth) status = ST_DEAD;   /* src/renderer.cu:549 */
            oo = make_float4(ro.x, ro.y, ro.z, u2f(((uint32_t)req << 16) | (uint32_t)(medium + 1)));
            od = make_float4(rd.x, rd.y, rd.z, u2f(pixel));
            oc = make_float4(color.x, color.y, color.z, u2f(pid));
            org = make_uint2((uint32_t)rs, (uint32_t)(rs >> 32));
        }
        const bool alive = status == ST_ALIVE;
        uint32_t slot;
        if constexpr (OCT) {
            /* survivors leave the block grouped by direction octant, so the extend waves that
             * take them 64 at a time trace rays from nearby origins in similar directions; one
             * atomic reserves the block's range */
            const int t = threadIdx.x;
            const uint32_t okey = alive ? ((od.x < 0.0f ? 1u : 0u) | (od.y < 0.0f ? 2u : 0u) | (od.z < 0.0f ? 4u : 0u)) : 8u;
            const uint32_t opos = block_split<9>(okey, s_wk);
            if (t == 0) {
                const uint32_t run = s_wk[8];   /* survivors: keys 0..7 */
                block_base = run ? atomicAdd(&ctl->count[depth + 1], run) : 0u;
            }
            __syncthreads();
            slot = alive ? block_base + opos : 0u;
            __syncthreads();   /* s_wk, block_base and the staging arrays are rewritten next trip */
        } else {
            /* compaction: ballot + prefix within the wave, wave offsets within the block, one
             * atomic per block */
            const unsigned long long mask = __ballot(alive);
            const uint32_t n = (uint32_t)__popcll(mask);
            if (lane == 0) wave_n[wid] = n;
            __syncthreads();
            if (threadIdx.x == 0) {
                uint32_t total = 0;
                for (int w = 0; w < SHADE_WAVES; w++) total += wave_n[w];
                block_base = total ? atomicAdd(&ctl->count[depth + 1], total) : 0u;
            }
            __syncthreads();
            uint32_t base = block_base;
            for (int w = 0; w < wid; w++) base += wave_n[w];
            __syncthreads();   /* wave_n and block_base are rewritten by the next trip */
            slot = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
        }
        if (alive) {
            outo[slot] = oo;
            outd[slot] = od;
            outc[slot] = oc;
            outr[slot] = org;
        }
    }
}

/* Kahan-compensated sum of the chunk's path radiances, pass order per pixel. One wave per
 * block takes 64 consecutive accumulate indices (pixels) at a time. Path ids come in runs of
 * `sgroup` passes per pixel, so when a wave's 64 pixels lie in one work item and the runs are at
 * least 8 long, the wave reads each 8-pass batch of its pixels cooperatively: 8 lanes per pixel
 * read that pixel's 128 contiguous bytes, so a load instruction touches 8 cache lines rather than
 * 64, and the batch is handed to the pixel's lane through LDS. Otherwise each lane reads its own
 * pixel's radiances. Either way a lane adds its pixel's passes in pass order. */
#define ACC_BATCH 8
#define ACC_ROW (ACC_BATCH + 1)   /* float4 per pixel row in LDS: one pad against bank conflicts */
__device__ inline void kahan_add(float4& s, float4& c, const float4 x) {
    float y, t;
    y = x.x - c.x; t = s.x + y; c.x = (t - s.x) - y; s.x = t;
    y = x.y - c.y; t = s.y + y; c.y = (t - s.y) - y; s.y = t;
    y = x.z - c.z; t = s.z + y; c.z = (t - s.z) - y; s.z = t;
}
__global__ __launch_bounds__(64) void k_accumulate(KParams P) {
    __shared__ float4 stage[64 * ACC_ROW];
    const GLOBAL DChunk* C = gp(P.chunk);
    const uint32_t total_px = C->total_px;
    const GLOBAL float4* contrib = gp(P.contrib);
    const uint32_t lane = threadIdx.x;
    /* the loop bound is uniform over the wave: every lane reaches the LDS barriers */
    for (uint32_t g0 = blockIdx.x * 64u; g0 < total_px; g0 += gridDim.x * 64u) {
        const uint32_t g = g0 + lane;
        const bool valid = g < total_px;
        const uint32_t gq = valid ? g : g0;
        int lo = 0, hi = C->n_work - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (C->work[mid].acc_base <= gq) lo = mid; else hi = mid - 1;
        }
        const GLOBAL DWork* W = C->work + lo;
        const GLOBAL DReq* R = C->reqs + W->req;
        const uint32_t q = gq - W->acc_base;
        const uint32_t p = tile_pixel_r(W->px0 + q, R);
        GLOBAL float4* acc = gpw(R->accum);
        GLOBAL float4* cmp = gpw(R->comp);
        float4 s = make_float4(0.0f, 0.0f, 0.0f, 0.0f), c = s;
        if (valid) { s = acc[p]; c = cmp[p]; }
        const int npass = W->npass;
        const uint32_t npx = W->npx, pid0 = W->pid_base, sg = W->sgroup;
        /* one work item for the whole wave, all 64 pixels present, runs of at least a batch */
        const bool coop = __ballot(valid && lo == __builtin_amdgcn_readfirstlane(lo)) == ~0ull && sg >= ACC_BATCH;
        if (coop) {
            const uint32_t q0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)q);
            const uint32_t sub = lane >> 3, j = lane & 7u;
            for (int k0 = 0; k0 < npass; k0 += ACC_BATCH) {
                /* pass k0 + j of pixel q0 + 8 i + sub; a run of sg passes never splits a batch */
                const uint32_t base = pid0 + ((uint32_t)k0 / sg) * npx * sg + ((uint32_t)k0 % sg) + j;
                float4 xs[8];
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    xs[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    if (k0 + (int)j < npass) xs[i] = contrib[base + (q0 + 8u * i + sub) * sg];
                }
#pragma unroll
                for (int i = 0; i < 8; i++) stage[(8u * i + sub) * ACC_ROW + j] = xs[i];
                __syncthreads();
#pragma unroll
                for (int jj = 0; jj < ACC_BATCH; jj++) {
                    if (k0 + jj >= npass) break;
                    kahan_add(s, c, stage[lane * ACC_ROW + jj]);
                }
                __syncthreads();   /* the next batch rewrites the stage */
            }
        } else if (valid) {
            /* eight loads in flight before the dependent Kahan chain */
            for (int k0 = 0; k0 < npass; k0 += 8) {
                float4 xs[8];
#pragma unroll
                for (int jj = 0; jj < 8; jj++)
                    if (k0 + jj < npass) xs[jj] = contrib[pid0 + slice_local((uint32_t)(k0 + jj), q, npx, sg)];
                    else xs[jj] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
                for (int jj = 0; jj < 8; jj++) {
                    if (k0 + jj >= npass) break;
                    kahan_add(s, c, xs[jj]);
                }
            }
        }
        if (valid) {
            acc[p] = s;
            cmp[p] = c;
        }
    }
}

__global__ void k_chunk_end(DCtl* ctl, const DChunk* chunk, unsigned long long* seg_total) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        unsigned long long s = 0;
        const int D = gp(chunk)->depth_steps;
        for (int d = 0; d < D; d++) s += gp(ctl)->count[d];
        gpw(ctl)->segments = s;
        atomicAdd(gpw(seg_total), s);   /* chunks of the two pipelines may end together */
    }
}

/* src/renderer.cu:426-441 */
__global__ __launch_bounds__(256) void k_tonemap(const float4* accum, ilm_pixel* out, int n, int spp, int op) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 a = gp(accum)[i];
    float samples = (float)spp;
    V3 avg = mul3(mk3(a.x, a.y, a.z), 1 / samples);
    avg = tonemap_eval(op, avg);
    avg = mk3(__builtin_fmaxf(__builtin_fminf(avg.x, 1.0f), 0.0f), __builtin_fmaxf(__builtin_fminf(avg.y, 1.0f), 0.0f),
              __builtin_fmaxf(__builtin_fminf(avg.z, 1.0f), 0.0f));
    ilm_pixel px;
    px.r = (uint64_t)(int)(255 * avg.x);
    px.g = (uint64_t)(int)(255 * avg.y);
    px.b = (uint64_t)(int)(255 * avg.z);
    gpw(out)[i] = px;
}

/* Kahan sums (float4 per pixel) to packed RGB floats, the layout of ilm_render's accum_rgb. */
__global__ __launch_bounds__(256) void k_sums_to_rgb(const float4* sums, float* rgb, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 a = gp(sums)[i];
    GLOBAL float* o = gpw(rgb) + 3 * (size_t)i;
    o[0] = a.x;
    o[1] = a.y;
    o[2] = a.z;
}

__global__ void k_fill_zero4(float4* p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        gpw(p)[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

/* ---------------- host-side launchers ---------------- */
#include "kernels.h"

/* A launch reports its own error only: the thread's sticky error is cleared first, so a failed
 * call elsewhere (whose status its caller already handled) is not blamed on the launch. */
#define ILM_LAUNCH(...)                                                                            \
    do {                                                                                           \
        (void)hipGetLastError();                                                                   \
        hipLaunchKernelGGL(__VA_ARGS__);                                                           \
    } while (0)

hipError_t ilm_k_ctl_init(const KParams& P, hipStream_t s) {
    ILM_LAUNCH(k_ctl_init, dim3(1), dim3(256), 0, s, P.ctl, P.chunk);
    return hipGetLastError();
}
hipError_t ilm_k_raygen(const KParams& P, int grid, hipStream_t s) {
    ILM_LAUNCH(k_raygen, dim3(grid), dim3(256), 0, s, P);
    return hipGetLastError();
}
hipError_t ilm_k_path_clear(const KParams& P, int grid, hipStream_t s) {
    ILM_LAUNCH(k_path_clear, dim3(grid), dim3(256), 0, s, P);
    return hipGetLastError();
}

template <int T, int Q, int C, int O>
static void launch_extend(const KParams& P, int depth, int* ovf, int grid, hipStream_t s) {
    if (P.extend_mode == ILM_EXTEND_STEP_REFILL && P.cam0 && P.packet0 && depth == 0)
        ILM_LAUNCH((k_extend_packet<T, Q, C>), dim3(grid), dim3(EXT_BLOCK), 0, s, P, depth);
    else if (P.extend_mode == ILM_EXTEND_STEP_REFILL && P.cam0 && depth == 0)
        ILM_LAUNCH((k_extend_pers<T, Q, C, O, 1, 0>), dim3(grid), dim3(EXT_BLOCK), 0, s, P, depth, ovf);
    else if (P.extend_mode == ILM_EXTEND_STEP_REFILL) {
        if constexpr (C == 0) {
            if (P.uni_scene) {
                ILM_LAUNCH((k_extend_pers<T, Q, 0, O, 0, 1>), dim3(grid), dim3(EXT_BLOCK), 0, s, P, depth, ovf);
                return;
            }
        }
        ILM_LAUNCH((k_extend_pers<T, Q, C, O, 0, 0>), dim3(grid), dim3(EXT_BLOCK), 0, s, P, depth, ovf);
    }
    else if (P.extend_mode == ILM_EXTEND_LEAF_REFILL)
        ILM_LAUNCH((k_extend_refill<T, Q, C, O>), dim3(grid), dim3(EXT_BLOCK), 0, s, P, depth, ovf);
    else
        ILM_LAUNCH((k_extend<T, Q, C, O>), dim3(grid), dim3(EXT_BLOCK), 0, s, P, depth, ovf);
}
template <int T, int Q, int C>
static void launch_extend_o(const KParams& P, int depth, int* ovf, int grid, hipStream_t s) {
    if (P.use_ovf) launch_extend<T, Q, C, 1>(P, depth, ovf, grid, s);
    else launch_extend<T, Q, C, 0>(P, depth, ovf, grid, s);
}
hipError_t ilm_k_extend(const KParams& P, int depth, int* ovf, int grid, hipStream_t s) {
    const int T = P.tri_test == ILM_TRI_WATERTIGHT;
    const int Q = P.bvh_layout == ILM_BVH_QUANT8;
    const int C = P.trav_stats != nullptr;
    switch (T * 4 + Q * 2 + C) {
    case 0: launch_extend_o<0, 0, 0>(P, depth, ovf, grid, s); break;
    case 1: launch_extend_o<0, 0, 1>(P, depth, ovf, grid, s); break;
    case 2: launch_extend_o<0, 1, 0>(P, depth, ovf, grid, s); break;
    case 3: launch_extend_o<0, 1, 1>(P, depth, ovf, grid, s); break;
    case 4: launch_extend_o<1, 0, 0>(P, depth, ovf, grid, s); break;
    case 5: launch_extend_o<1, 0, 1>(P, depth, ovf, grid, s); break;
    case 6: launch_extend_o<1, 1, 0>(P, depth, ovf, grid, s); break;
    default: launch_extend_o<1, 1, 1>(P, depth, ovf, grid, s); break;
    }
    return hipGetLastError();
}
hipError_t ilm_k_shade(const KParams& P, int depth, int grid, hipStream_t s) {
    /* the first depth's paths are in path-id order, consecutive ids being samples of one pixel
     * (or of neighbouring pixels), so a wave already meets one or two materials: sorting them
     * through LDS costs more than it saves there (measured), and only later depths sort */
    const bool cam = P.cam0 && depth == 0;
    if (cam) ILM_LAUNCH((k_shade<0, 1>), dim3(grid), dim3(SHADE_BLOCK), 0, s, P, depth);
    else if (P.shade_sort) ILM_LAUNCH((k_shade<1, 0>), dim3(grid), dim3(SHADE_BLOCK), 0, s, P, depth);
    else ILM_LAUNCH((k_shade<0, 0>), dim3(grid), dim3(SHADE_BLOCK), 0, s, P, depth);
    return hipGetLastError();
}
hipError_t ilm_k_accumulate(const KParams& P, int grid, hipStream_t s) {
    ILM_LAUNCH(k_accumulate, dim3(grid * 4), dim3(64), 0, s, P);
    return hipGetLastError();
}
hipError_t ilm_k_chunk_end(const KParams& P, unsigned long long* seg_total, hipStream_t s) {
    ILM_LAUNCH(k_chunk_end, dim3(1), dim3(64), 0, s, P.ctl, P.chunk, seg_total);
    return hipGetLastError();
}
hipError_t ilm_k_tonemap(const float4* accum, ilm_pixel* out, int n, int spp, int op, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    ILM_LAUNCH(k_tonemap, dim3((n + 255) / 256), dim3(256), 0, s, accum, out, n, spp, op);
    return hipGetLastError();
}
hipError_t ilm_k_sums_to_rgb(const float4* sums, float* rgb, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    ILM_LAUNCH(k_sums_to_rgb, dim3((n + 255) / 256), dim3(256), 0, s, sums, rgb, n);
    return hipGetLastError();
}
hipError_t ilm_k_fill_zero4(float4* p, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    size_t g = (n + 255) / 256;
    if (g > 65535) g = 65535;
    ILM_LAUNCH(k_fill_zero4, dim3((unsigned)g), dim3(256), 0, s, p, n);
    return hipGetLastError();
}
template <class K>
static int occ_of(K kernel) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, EXT_BLOCK, 0) != hipSuccess || nb <= 0) nb = 8;
    return nb;
}
/* Blocks per CU of the extend kernel (the smallest over the variants a context may launch). */
int ilm_k_extend_occupancy(int tri_test, int bvh_layout, int extend_mode) {
    int a, b;
    if (extend_mode == ILM_EXTEND_STEP_REFILL) {
        if (!tri_test && !bvh_layout) { a = occ_of(k_extend_pers<0, 0, 0, 0, 0, 0>); b = occ_of(k_extend_pers<0, 0, 0, 1, 0, 0>); }
        else if (tri_test && !bvh_layout) { a = occ_of(k_extend_pers<1, 0, 0, 0, 0, 0>); b = occ_of(k_extend_pers<1, 0, 0, 1, 0, 0>); }
        else if (!tri_test) { a = occ_of(k_extend_pers<0, 1, 0, 0, 0, 0>); b = occ_of(k_extend_pers<0, 1, 0, 1, 0, 0>); }
        else { a = occ_of(k_extend_pers<1, 1, 0, 0, 0, 0>); b = occ_of(k_extend_pers<1, 1, 0, 1, 0, 0>); }
    } else if (extend_mode == ILM_EXTEND_LEAF_REFILL) {
        if (!tri_test && !bvh_layout) { a = occ_of(k_extend_refill<0, 0, 0, 0>); b = occ_of(k_extend_refill<0, 0, 0, 1>); }
        else if (tri_test && !bvh_layout) { a = occ_of(k_extend_refill<1, 0, 0, 0>); b = occ_of(k_extend_refill<1, 0, 0, 1>); }
        else if (!tri_test) { a = occ_of(k_extend_refill<0, 1, 0, 0>); b = occ_of(k_extend_refill<0, 1, 0, 1>); }
        else { a = occ_of(k_extend_refill<1, 1, 0, 0>); b = occ_of(k_extend_refill<1, 1, 0, 1>); }
    } else {
        if (!tri_test && !bvh_layout) { a = occ_of(k_extend<0, 0, 0, 0>); b = occ_of(k_extend<0, 0, 0, 1>); }
        else if (tri_test && !bvh_layout) { a = occ_of(k_extend<1, 0, 0, 0>); b = occ_of(k_extend<1, 0, 0, 1>); }
        else if (!tri_test) { a = occ_of(k_extend<0, 1, 0, 0>); b = occ_of(k_extend<0, 1, 0, 1>); }
        else { a = occ_of(k_extend<1, 1, 0, 0>); b = occ_of(k_extend<1, 1, 0, 1>); }
    }
    return a > b ? a : b;
}
/* Resident 512-thread shading blocks per CU (the LDS staging of the sorted variant limits it). */
int ilm_k_shade_occupancy(int sort) {
    int nb = 0;
    hipError_t e = sort ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_shade<1, 0>, SHADE_BLOCK, 0)
                        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_shade<0, 0>, SHADE_BLOCK, 0);
    return (e != hipSuccess || nb <= 0) ? 2 : nb;
}
int ilm_k_ovf_stack() { return OVF_STACK; }
int ilm_k_lds_stack() { return LDS_STACK; }
int ilm_k_ext_block() { return EXT_BLOCK; }
