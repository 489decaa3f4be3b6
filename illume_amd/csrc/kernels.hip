/*
 * kernels.hip - the gfx950 wavefront pipeline of the path tracer.
 *
 *   k_ctl_init     clear the chunk's queue counters (no memset: graph-safe)
 *   k_raygen       camera rays for (request, pass, pixel) work, PCG32 seeded per path
 *                  (src/renderer.cu:61-69,81-110)
 *   k_extend       closest hit: spheres, then a TLAS over instances and per-mesh BLAS, stack
 *                  walk in wave64 with the stack in LDS (bounded global spill), t-pruned and
 *                  ordered (src/renderer.cu:112-137, src/primitives/mesh.cu:479-562)
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
#include "shade.h"

#define EXT_BLOCK 64            /* one wave per workgroup: the LDS stack belongs to one wave */
#define LDS_STACK 32
#define OVF_STACK 96            /* spill depth: scene upload checks the two-level stack bound fits 128 */
#define STACK_SENTINEL ((int)0x7ffffffe)
#define SHADE_BLOCK 256

/* conservative slab-test margins (relative); cover rcp and fma rounding of the box test */
#define BOX_LO 0.99999f
#define BOX_HI 1.00001f

__device__ inline uint32_t f2u(float f) { return __float_as_uint(f); }
__device__ inline float u2f(uint32_t u) { return __uint_as_float(u); }

struct Best {
    float t, u, v;
    int inst, prim;
};

__device__ inline V3 rcp3(V3 d) {
    return mk3(__builtin_amdgcn_rcpf(d.x), __builtin_amdgcn_rcpf(d.y), __builtin_amdgcn_rcpf(d.z));
}

/* Two child boxes of an FP32 BVH2 node against the ray (o, 1/d); noid = -o/d. */
__device__ inline void box2(const float4 a, const float4 b, const float4 c, V3 id, V3 noid, float tmax,
                            float& t0, float& t1, bool& h0, bool& h1) {
    float x0 = __builtin_fmaf(a.x, id.x, noid.x), x1 = __builtin_fmaf(a.y, id.x, noid.x);
    float y0 = __builtin_fmaf(a.z, id.y, noid.y), y1 = __builtin_fmaf(a.w, id.y, noid.y);
    float z0 = __builtin_fmaf(b.x, id.z, noid.z), z1 = __builtin_fmaf(b.y, id.z, noid.z);
    float tn = __builtin_fmaxf(__builtin_fmaxf(__builtin_fminf(x0, x1), __builtin_fminf(y0, y1)),
                               __builtin_fmaxf(__builtin_fminf(z0, z1), 0.0f));
    float tf = __builtin_fminf(__builtin_fminf(__builtin_fmaxf(x0, x1), __builtin_fmaxf(y0, y1)),
                               __builtin_fmaxf(z0, z1));
    t0 = tn * BOX_LO;
    h0 = (t0 <= tf * BOX_HI) && (t0 <= tmax);

    float X0 = __builtin_fmaf(b.z, id.x, noid.x), X1 = __builtin_fmaf(b.w, id.x, noid.x);
    float Y0 = __builtin_fmaf(c.x, id.y, noid.y), Y1 = __builtin_fmaf(c.y, id.y, noid.y);
    float Z0 = __builtin_fmaf(c.z, id.z, noid.z), Z1 = __builtin_fmaf(c.w, id.z, noid.z);
    float Tn = __builtin_fmaxf(__builtin_fmaxf(__builtin_fminf(X0, X1), __builtin_fminf(Y0, Y1)),
                               __builtin_fmaxf(__builtin_fminf(Z0, Z1), 0.0f));
    float Tf = __builtin_fminf(__builtin_fminf(__builtin_fmaxf(X0, X1), __builtin_fmaxf(Y0, Y1)),
                               __builtin_fmaxf(Z0, Z1));
    t1 = Tn * BOX_LO;
    h1 = (t1 <= Tf * BOX_HI) && (t1 <= tmax);
}

/* Up to four dequantised child boxes of a QNode. */
__device__ inline void box4q(const QNode& n, V3 id, V3 noid, float tmax, float* tn_out, uint32_t& hitmask) {
    float sx = u2f((uint32_t)n.ex << 23), sy = u2f((uint32_t)n.ey << 23), sz = u2f((uint32_t)n.ez << 23);
    hitmask = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (k >= (int)n.nchild) break;
        float lx = n.ox + (float)n.qlo[0][k] * sx, hx = n.ox + (float)n.qhi[0][k] * sx;
        float ly = n.oy + (float)n.qlo[1][k] * sy, hy = n.oy + (float)n.qhi[1][k] * sy;
        float lz = n.oz + (float)n.qlo[2][k] * sz, hz = n.oz + (float)n.qhi[2][k] * sz;
        float x0 = __builtin_fmaf(lx, id.x, noid.x), x1 = __builtin_fmaf(hx, id.x, noid.x);
        float y0 = __builtin_fmaf(ly, id.y, noid.y), y1 = __builtin_fmaf(hy, id.y, noid.y);
        float z0 = __builtin_fmaf(lz, id.z, noid.z), z1 = __builtin_fmaf(hz, id.z, noid.z);
        float tn = __builtin_fmaxf(__builtin_fmaxf(__builtin_fminf(x0, x1), __builtin_fminf(y0, y1)),
                                   __builtin_fmaxf(__builtin_fminf(z0, z1), 0.0f));
        float tf = __builtin_fminf(__builtin_fminf(__builtin_fmaxf(x0, x1), __builtin_fmaxf(y0, y1)),
                                   __builtin_fmaxf(z0, z1));
        tn = tn * BOX_LO;
        tn_out[k] = tn;
        if (tn <= tf * BOX_HI && tn <= tmax) hitmask |= 1u << k;
    }
}

struct Stack {
    int* lds;       /* &stk[lane]; entry s at lds[s * EXT_BLOCK] */
    int* ovf;       /* this lane's spill area */
    int sp;
    __device__ inline void push(int v) {
        if (sp < LDS_STACK) lds[sp * EXT_BLOCK] = v;
        else if (sp < LDS_STACK + OVF_STACK) ovf[sp - LDS_STACK] = v;
        sp = sp < LDS_STACK + OVF_STACK ? sp + 1 : sp;
    }
    __device__ inline int pop() {
        --sp;
        return sp < LDS_STACK ? lds[sp * EXT_BLOCK] : ovf[sp - LDS_STACK];
    }
};

struct TravCount {
    uint32_t boxes, tris;
};

template <int TRI, int QUANT, int COUNT>
__device__ inline void trace_ray(const DScene& S, V3 O, V3 D, Best& B, Stack& st, TravCount* tc) {
    B.t = ILM_FLT_MAX; B.u = 0.0f; B.v = 0.0f; B.inst = -1; B.prim = -1;
    for (int s = 0; s < S.n_spheres; s++) {
        float t;
        if (sphere_hit(S.spheres[s], O, D, t) && t < B.t) { B.t = t; B.inst = -2 - s; }
    }
    int node = QUANT ? S.qtlas_root : S.tlas_root;
    if (node == ILM_NO_CHILD) return;
    const BNode* nodes = S.tlas;
    const QNode* qnodes = S.qtlas;
    const float4* tris = nullptr;
    bool blas = false;
    int cur = -1;
    V3 o = O, d = D;
    V3 id = rcp3(d);
    V3 noid = mk3(-o.x * id.x, -o.y * id.y, -o.z * id.z);
    WTRay wt;
    st.sp = 0;
    for (;;) {
        if (node >= 0) {
            if (QUANT) {
                const QNode& n = qnodes[node];
                float tn[4];
                uint32_t hm;
                box4q(n, id, noid, B.t, tn, hm);
                if (COUNT) tc->boxes += n.nchild;
                if (hm == 0) goto pop;
                /* visit nearest first: push the others far-to-near */
                int order[4];
                int cnt = 0;
#pragma unroll
                for (int k = 0; k < 4; k++) if (hm & (1u << k)) order[cnt++] = k;
                for (int a1 = 1; a1 < cnt; a1++) {
                    int key = order[a1];
                    int b1 = a1 - 1;
                    while (b1 >= 0 && tn[order[b1]] > tn[key]) { order[b1 + 1] = order[b1]; b1--; }
                    order[b1 + 1] = key;
                }
                for (int k = cnt - 1; k >= 1; k--) st.push(n.child[order[k]]);
                node = n.child[order[0]];
                continue;
            } else {
                const BNode* n = nodes + node;
                float4 a = n->a, b = n->b, c = n->c;
                int4 ch = n->d;
                float t0, t1;
                bool h0, h1;
                box2(a, b, c, id, noid, B.t, t0, t1, h0, h1);
                if (COUNT) tc->boxes += 2;
                if (h0 && h1) {
                    int nearc = ch.x, farc = ch.y;
                    if (t1 < t0) { nearc = ch.y; farc = ch.x; }
                    st.push(farc);
                    node = nearc;
                    continue;
                }
                if (h0) { node = ch.x; continue; }
                if (h1) { node = ch.y; continue; }
                goto pop;
            }
        } else if (!blas) {
            /* TLAS leaf: one instance; enter its BLAS in object space */
            int k = S.tlas_prims[ilm_leaf_first(node)];
            const DInst& I = S.inst[k];
            const DMesh& M = S.meshes[I.mesh];
            const int broot = QUANT ? M.qroot : M.root;
            if (broot == ILM_NO_CHILD) goto pop;   /* empty mesh */
            o = xform3(I.inv, O, 1.0f);
            d = xform3(I.inv, D, 0.0f);
            id = rcp3(d);
            noid = mk3(-o.x * id.x, -o.y * id.y, -o.z * id.z);
            if (TRI) wt = wt_setup(d);
            nodes = M.nodes;
            qnodes = M.qnodes;
            tris = M.tris;
            st.push(STACK_SENTINEL);
            blas = true;
            cur = k;
            node = broot;
            continue;
        } else {
            uint32_t f = ilm_leaf_first(node), cnt = ilm_leaf_count(node);
            if (COUNT) tc->tris += cnt;
            for (uint32_t q = 0; q < cnt; q++) {
                const float4* tp = tris + 3 * (f + q);
                float4 A = tp[0], Bq = tp[1], C = tp[2];
                float u = 0.0f, v = 0.0f, t;
                if (TRI) t = tri_wt(wt, mk3(A.x, A.y, A.z), mk3(Bq.x, Bq.y, Bq.z), mk3(C.x, C.y, C.z), o, u, v);
                else t = tri_mt(mk3(A.x, A.y, A.z), mk3(Bq.x, Bq.y, Bq.z), mk3(C.x, C.y, C.z), o, d, u, v);
                if (t > 0.0f && t < B.t) { B.t = t; B.u = u; B.v = v; B.inst = cur; B.prim = __float_as_int(A.w); }
            }
        }
    pop:
        for (;;) {
            if (st.sp == 0) return;
            node = st.pop();
            if (node != STACK_SENTINEL) break;
            blas = false;
            nodes = S.tlas;
            qnodes = S.qtlas;
            o = O; d = D;
            id = rcp3(d);
            noid = mk3(-o.x * id.x, -o.y * id.y, -o.z * id.z);
        }
    }
}

__global__ void k_ctl_init(DCtl* ctl, const DChunk* chunk) {
    for (int i = threadIdx.x; i < ILM_MAX_DEPTH + 2; i += blockDim.x) {
        ctl->count[i] = 0u;
        ctl->fetch[i] = 0u;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        ctl->count[0] = chunk->total_paths;
        ctl->total = chunk->total_paths;
        ctl->segments = 0ull;
    }
}

__device__ inline int find_work_pid(const DWork* work, int n, uint32_t i) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (work[mid].pid_base <= i) lo = mid; else hi = mid - 1;
    }
    return lo;
}

/* src/renderer.cu:81-110 (and the per-pass seeding of :61-69) */
__global__ __launch_bounds__(256) void k_raygen(KParams P) {
    const uint32_t total = P.chunk->total_paths;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        int req;
        uint32_t pixel, pass;
        if (P.list) {
            req = 0;
            pixel = P.list[2 * i];
            pass = P.list[2 * i + 1];
        } else {
            const DWork& W = P.chunk->work[find_work_pid(P.chunk->work, P.chunk->n_work, i)];
            uint32_t local = i - W.pid_base;
            req = W.req;
            pass = (uint32_t)W.pass0 + local / W.npx;
            pixel = W.px0 + local % W.npx;
        }
        const DReq& R = P.chunk->reqs[req];
        const DScene& S = *R.scene;
        uint64_t rs = ilm_path_seed(pass, pixel);
        const uint64_t inc = ilm_pcg32_inc(pixel);
        int Wd = (int)R.image_width;
        int x = (int)pixel % Wd, y = (int)pixel / Wd;
        float left_edge = R.left + R.pixel_size * (float)x;
        float top_edge = R.top - R.pixel_size * (float)y;
        float rx = left_edge + R.pixel_size * ilm_pcg32_uniform(&rs, inc);
        float ry = top_edge - R.pixel_size * ilm_pcg32_uniform(&rs, inc);
        V3 pos = mk3(0.0f, 0.0f, 0.0f);
        if (S.aperture > ILM_FLT_EPSILON) {
            float u1 = ilm_pcg32_uniform(&rs, inc);
            float u2 = ilm_pcg32_uniform(&rs, inc);
            pos = mul3(sample_circle(u1, u2), S.aperture);
        }
        V3 origin = xform3(S.cam, pos, 1.0f);
        V3 image_pos = xform3(S.cam, mk3(rx, ry, S.dof), 1.0f);
        V3 dir = normalize3(sub3(image_pos, origin));
        P.in.o[i] = make_float4(origin.x, origin.y, origin.z, u2f((uint32_t)req << 16));
        P.in.d[i] = make_float4(dir.x, dir.y, dir.z, u2f(pixel));
        P.in.c[i] = make_float4(1.0f, 1.0f, 1.0f, u2f(i));
        P.in.rng[i] = make_uint2((uint32_t)rs, (uint32_t)(rs >> 32));
        P.contrib[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (P.seg_count) P.seg_count[i] = 0;
    }
}

template <int TRI, int QUANT, int COUNT>
__global__ __launch_bounds__(EXT_BLOCK) void k_extend(KParams P, int depth, int* ovf_base) {
    __shared__ int stk[LDS_STACK * EXT_BLOCK];
    const int lane = threadIdx.x;
    Stack st;
    st.lds = stk + lane;
    st.ovf = ovf_base + ((size_t)blockIdx.x * EXT_BLOCK + lane) * OVF_STACK;
    st.sp = 0;
    const uint32_t count = P.ctl->count[depth];
    for (;;) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&P.ctl->fetch[depth], (uint32_t)EXT_BLOCK);
        base = __shfl(base, 0);
        if (base >= count) break;
        uint32_t i = base + lane;
        if (i < count) {
            float4 o4 = P.in.o[i];
            float4 d4 = P.in.d[i];
            int req = (int)(f2u(o4.w) >> 16);
            const DScene& S = *P.chunk->reqs[req].scene;
            Best B;
            if (COUNT) {
                TravCount tc = {0u, 0u};
                trace_ray<TRI, QUANT, 1>(S, mk3(o4.x, o4.y, o4.z), mk3(d4.x, d4.y, d4.z), B, st, &tc);
                atomicAdd(&P.trav_stats[0], (unsigned long long)tc.boxes);
                atomicAdd(&P.trav_stats[1], (unsigned long long)tc.tris);
            } else {
                trace_ray<TRI, QUANT, 0>(S, mk3(o4.x, o4.y, o4.z), mk3(d4.x, d4.y, d4.z), B, st, nullptr);
            }
            P.hits.h[i] = make_float4(B.t, B.u, B.v, __int_as_float(B.prim));
            P.hits.inst[i] = B.inst;
        }
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

/* One path segment after its closest hit: src/renderer.cu:152-386. */
__device__ inline int shade_path(const KParams& P, const DScene& S, const DReq& R, const float4 h4, int hinst,
                                 V3& ro, V3& rd, V3& color, int& medium, uint64_t& rs, uint64_t inc) {
    float tmin = h4.x;
    const bool is = hinst != -1;
    if (medium >= 0 && S.mediums[medium].active) {
        const DMedium& md = S.mediums[medium];
        float u = ilm_pcg32_uniform(&rs, inc);
        float scatter_dist = -pm_logf(u) / md.scattering;
        if (scatter_dist < tmin) {
            float u1 = ilm_pcg32_uniform(&rs, inc);
            float u2 = ilm_pcg32_uniform(&rs, inc);
            V3 sdir = to_basis3(sample_hg(md.g, u1, u2), rd);
            ro = add3(ro, mul3(rd, scatter_dist));
            rd = normalize3(sdir);
            color = mulv3(color, mk3(pm_expf(-1.0f * scatter_dist * md.absorption[0]),
                                     pm_expf(-1.0f * scatter_dist * md.absorption[1]),
                                     pm_expf(-1.0f * scatter_dist * md.absorption[2])));
            if ((double)len2_3(color) < 1e-4) return ST_DEAD;
            return ST_ALIVE;
        }
        color = mulv3(color, mk3(pm_expf(-1.0f * tmin * md.absorption[0]),
                                 pm_expf(-1.0f * tmin * md.absorption[1]),
                                 pm_expf(-1.0f * tmin * md.absorption[2])));
    }
    if (!is) {
        /* src/renderer.cu:378-386 */
        float u = (float)((double)(pm_atan2f(rd.z, rd.x) * 0.5f) * ILM_INV_PI_D + 0.5f);
        float v = (float)(0.5f - (double)pm_asinf(rd.y) * ILM_INV_PI_D);
        color = mulv3(color, tex_eval(S.env, u, v));
        return ST_EMIT;
    }
    /* surface attributes: src/primitives/mesh.cu:532-561, mesh_instance.cu:102-107, sphere.cu:207-212 */
    V3 n;
    float hu, hv;
    int mat;
    if (hinst >= 0) {
        const DInst& I = S.inst[hinst];
        const DMesh& M = S.meshes[I.mesh];
        int prim = __float_as_int(h4.w);
        int i0 = M.idx[3 * prim], i1 = M.idx[3 * prim + 1], i2 = M.idx[3 * prim + 2];
        float bu = h4.y, bv = h4.z;
        if (M.face_normals) {
            V3 v0 = mk3(M.pos[3 * i0], M.pos[3 * i0 + 1], M.pos[3 * i0 + 2]);
            V3 v1 = mk3(M.pos[3 * i1], M.pos[3 * i1 + 1], M.pos[3 * i1 + 2]);
            V3 v2 = mk3(M.pos[3 * i2], M.pos[3 * i2 + 1], M.pos[3 * i2 + 2]);
            n = normalize3(cross3(sub3(v1, v0), sub3(v2, v0)));
        } else {
            V3 n0 = mk3(M.nrm[3 * i0], M.nrm[3 * i0 + 1], M.nrm[3 * i0 + 2]);
            V3 n1 = mk3(M.nrm[3 * i1], M.nrm[3 * i1 + 1], M.nrm[3 * i1 + 2]);
            V3 n2 = mk3(M.nrm[3 * i2], M.nrm[3 * i2 + 1], M.nrm[3 * i2 + 2]);
            n = normalize3(add3(mul3(n0, 1.0f - bu - bv), add3(mul3(n1, bu), mul3(n2, bv))));
        }
        if (M.has_texcoords) {
            float w = 1.0f - bu - bv;
            hu = (M.uv[2 * i0] * w + M.uv[2 * i1] * bu) + M.uv[2 * i2] * bv;
            hv = (M.uv[2 * i0 + 1] * w + M.uv[2 * i1 + 1] * bu) + M.uv[2 * i2 + 1] * bv;
        } else {
            hu = bu;
            hv = bv;
        }
        n = normalize3(xform3(I.tinv, n, 0.0f));
        mat = I.material;
    } else {
        const DSphere& sp = S.spheres[-2 - hinst];
        V3 pos = add3(ro, mul3(rd, tmin));
        n = normalize3(sub3(pos, mk3(sp.c[0], sp.c[1], sp.c[2])));
        hu = (float)((double)(pm_atan2f(n.z, n.x) * 0.5f) * ILM_INV_PI_D + 0.5f);
        hv = (float)(0.5f - (double)pm_asinf(n.y) * ILM_INV_PI_D);
        mat = sp.material;
    }
    const DMat& m = S.mats[mat];
    const V3 d = rd;
    V3 new_dir;
    V3 norm_o = mul3(n, dot3(n, d) > 0 ? -1.0f : 1.0f);
    V3 new_origin = add3(ro, mul3(d, tmin));
    V3 albedo = tex_eval(m.albedo, hu, hv);
    float roughness = tex_eval(m.roughness, hu, hv).x;
    const float bias = R.ray_bias;
    switch (m.type) {
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
        float cosI = -dot3(d, n);
        float cosT = 0.0f;
        float F = fresnel_dielectric(cosI, m.ior, cosT);
        if (F == 1.0f || ilm_pcg32_uniform(&rs, inc) < F) {
            new_dir = reflect3(d, n);
            new_origin = add3(new_origin, mul3(norm_o, bias));
        } else {
            float eta = cosI < 0.0f ? m.ior : 1.0f / m.ior;
            medium = cosI > 0.0f ? m.medium : -1;
            new_dir = add3(mul3(d, eta), mul3(norm_o, eta * pm_fabs(cosI) - cosT));
            new_origin = add3(new_origin, mul3(norm_o, -bias));
        }
        color = mulv3(color, albedo);
        break;
    }
    case ILM_MAT_ROUGHREFLECTIVE: {
        V3 wi = mul3(d, -1.0f);
        float wiDotN = dot3(wi, n);
        float a = mf_alpha(m.distribution, roughness, wiDotN);
        float u1 = ilm_pcg32_uniform(&rs, inc);
        float u2 = ilm_pcg32_uniform(&rs, inc);
        V3 mm = to_basis3(mf_sample(m.distribution, a, u1, u2), n);
        float wiDotT = 0.0f;
        float wiDotM = dot3(wi, mm);
        float F = fresnel_dielectric(wiDotM, m.ior, wiDotT);
        new_dir = reflect3(d, mm);
        if (wiDotN * dot3(new_dir, n) <= 0.0f) return ST_DEAD;
        new_origin = add3(new_origin, mul3(norm_o, bias));
        float G = mf_G(m.distribution, wi, new_dir, mm, n, a);
        float weight = (F * G * pm_fabs(wiDotM)) / (pm_fabs(wiDotN) * pm_fabs(dot3(mm, n)));
        color = mulv3(color, mul3(albedo, weight));
        break;
    }
    case ILM_MAT_ROUGHREFRACTIVE: {
        V3 wi = mul3(d, -1.0f);
        float wiDotN = dot3(wi, n);
        float a = mf_alpha(m.distribution, roughness, wiDotN);
        float u1 = ilm_pcg32_uniform(&rs, inc);
        float u2 = ilm_pcg32_uniform(&rs, inc);
        V3 mm = to_basis3(mf_sample(m.distribution, a, u1, u2), n);
        float wiDotT = 0.0f;
        float wiDotM = dot3(wi, mm);
        float F = fresnel_dielectric(wiDotM, m.ior, wiDotT);
        if (F == 1.0f || ilm_pcg32_uniform(&rs, inc) < F) {
            new_dir = reflect3(d, mm);
            if (wiDotN * dot3(new_dir, n) <= 0.0f) return ST_DEAD;
            new_origin = add3(new_origin, mul3(norm_o, bias));
        } else {
            float eta = wiDotM < 0.0f ? m.ior : 1.0f / m.ior;
            medium = wiDotM > 0.0f ? m.medium : -1;
            new_dir = sub3(mul3(mm, wiDotM * eta - (wiDotM > 0.0f ? 1.0f : -1.0f) * wiDotT), mul3(wi, eta));
            if (wiDotN * dot3(new_dir, n) >= 0.0f) return ST_DEAD;
            new_origin = add3(new_origin, mul3(norm_o, -bias));
        }
        float G = mf_G(m.distribution, wi, new_dir, mm, n, a);
        float weight = (G * pm_fabs(wiDotM)) / (pm_fabs(wiDotN) * pm_fabs(dot3(mm, n)));
        color = mulv3(color, mul3(albedo, weight));
        break;
    }
    case ILM_MAT_CONDUCTOR: {
        float cosI = -dot3(d, n);
        if (cosI <= 0) return ST_DEAD;
        new_dir = reflect3(d, norm_o);
        new_origin = add3(new_origin, mul3(norm_o, bias));
        color = mulv3(color, mk3(fresnel_conductor(m.eta[0], m.k[0], cosI), fresnel_conductor(m.eta[1], m.k[1], cosI),
                                 fresnel_conductor(m.eta[2], m.k[2], cosI)));
        break;
    }
    case ILM_MAT_ROUGHCONDUCTOR: {
        V3 wi = mul3(d, -1.0f);
        float wiDotN = dot3(wi, n);
        if (wiDotN <= 0) return ST_DEAD;
        float a = mf_alpha(m.distribution, roughness, wiDotN);
        float u1 = ilm_pcg32_uniform(&rs, inc);
        float u2 = ilm_pcg32_uniform(&rs, inc);
        V3 mm = to_basis3(mf_sample(m.distribution, a, u1, u2), n);
        V3 F = mk3(fresnel_conductor(m.eta[0], m.k[0], wiDotN), fresnel_conductor(m.eta[1], m.k[1], wiDotN),
                   fresnel_conductor(m.eta[2], m.k[2], wiDotN));
        new_dir = reflect3(d, mm);
        if (wiDotN * dot3(new_dir, n) <= 0.0f) return ST_DEAD;
        float wiDotM = dot3(wi, mm);
        float G = mf_G(m.distribution, wi, new_dir, mm, n, a);
        float weight = (G * pm_fabs(wiDotM)) / (pm_fabs(wiDotN) * pm_fabs(dot3(mm, n)));
        color = mulv3(color, mul3(F, weight));
        new_origin = add3(new_origin, mul3(mm, bias));
        break;
    }
    case ILM_MAT_PLASTIC: {
        float cosI = -dot3(d, n);
        float cosT = 0.0f;
        float F = fresnel_dielectric(cosI, m.ior, cosT);
        float u1 = ilm_pcg32_uniform(&rs, inc);
        if (F == 1.0f || u1 < F) {
            new_dir = reflect3(d, n);
            color = mulv3(color, mk3(0.99f, 0.99f, 0.99f));
        } else {
            float u2 = ilm_pcg32_uniform(&rs, inc);
            float uu = (float)((double)(u1 - F) / (1.0 - (double)F));
            new_dir = to_basis3(sample_hemisphere_cosine(uu, u2), norm_o);
            float cosO = dot3(new_dir, n);
            float Fo = fresnel_dielectric(cosO, m.ior, cosT);
            V3 diff = divs3(albedo, 1.0f - m.diffuse_fresnel);
            float inv_eta = 1.0f / m.ior;
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

__global__ __launch_bounds__(SHADE_BLOCK) void k_shade(KParams P, int depth) {
    const uint32_t count = P.ctl->count[depth];
    const int lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t wbase = wave * 64u; wbase < count; wbase += nwaves * 64u) {
        uint32_t i = wbase + lane;
        int status = ST_DEAD;
        float4 oo = make_float4(0, 0, 0, 0), od = oo, oc = oo;
        uint2 org = make_uint2(0, 0);
        if (i < count) {
            float4 o4 = P.in.o[i], d4 = P.in.d[i], c4 = P.in.c[i];
            uint2 rg = P.in.rng[i];
            float4 h4 = P.hits.h[i];
            int hinst = P.hits.inst[i];
            uint32_t tag = f2u(o4.w);
            int req = (int)(tag >> 16);
            int medium = (int)(tag & 0xffffu) - 1;
            uint32_t pixel = f2u(d4.w);
            uint32_t pid = f2u(c4.w);
            const DReq& R = P.chunk->reqs[req];
            const DScene& S = *R.scene;
            uint64_t rs = ((uint64_t)rg.y << 32) | (uint64_t)rg.x;
            V3 ro = mk3(o4.x, o4.y, o4.z), rd = mk3(d4.x, d4.y, d4.z), color = mk3(c4.x, c4.y, c4.z);
            if (P.seg_count) P.seg_count[pid] += 1;
            status = shade_path(P, S, R, h4, hinst, ro, rd, color, medium, rs, ilm_pcg32_inc(pixel));
            if (status == ST_EMIT) P.contrib[pid] = make_float4(color.x, color.y, color.z, 0.0f);
            if (status == ST_ALIVE && depth + 1 >= R.max_depth) status = ST_DEAD;   /* src/renderer.cu:549 */
            oo = make_float4(ro.x, ro.y, ro.z, u2f(((uint32_t)req << 16) | (uint32_t)(medium + 1)));
            od = make_float4(rd.x, rd.y, rd.z, u2f(pixel));
            oc = make_float4(color.x, color.y, color.z, u2f(pid));
            org = make_uint2((uint32_t)rs, (uint32_t)(rs >> 32));
        }
        /* compaction: ballot + prefix within the wave, one atomic per wave */
        const bool alive = status == ST_ALIVE;
        const unsigned long long mask = __ballot(alive);
        const uint32_t n = (uint32_t)__popcll(mask);
        uint32_t base = 0;
        if (lane == 0 && n) base = atomicAdd(&P.ctl->count[depth + 1], n);
        base = __shfl(base, 0);
        if (alive) {
            const uint32_t rank = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
            const uint32_t slot = base + rank;
            P.out.o[slot] = oo;
            P.out.d[slot] = od;
            P.out.c[slot] = oc;
            P.out.rng[slot] = org;
        }
    }
}

/* Kahan-compensated sum of the chunk's path radiances, pass order per pixel. */
__global__ __launch_bounds__(256) void k_accumulate(KParams P) {
    const uint32_t total_px = P.chunk->total_px;
    const DWork* work = P.chunk->work;
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < total_px; g += gridDim.x * blockDim.x) {
        int lo = 0, hi = P.chunk->n_work - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (work[mid].acc_base <= g) lo = mid; else hi = mid - 1;
        }
        const DWork& W = work[lo];
        const DReq& R = P.chunk->reqs[W.req];
        const uint32_t q = g - W.acc_base;
        const uint32_t p = W.px0 + q;
        float4 s = R.accum[p], c = R.comp[p];
        for (int k = 0; k < W.npass; k++) {
            float4 x = P.contrib[W.pid_base + (uint32_t)k * W.npx + q];
            float y, t;
            y = x.x - c.x; t = s.x + y; c.x = (t - s.x) - y; s.x = t;
            y = x.y - c.y; t = s.y + y; c.y = (t - s.y) - y; s.y = t;
            y = x.z - c.z; t = s.z + y; c.z = (t - s.z) - y; s.z = t;
        }
        R.accum[p] = s;
        R.comp[p] = c;
    }
}

__global__ void k_chunk_end(DCtl* ctl, const DChunk* chunk, unsigned long long* seg_total) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        unsigned long long s = 0;
        for (int d = 0; d < chunk->depth_steps; d++) s += ctl->count[d];
        ctl->segments = s;
        *seg_total += s;
    }
}

/* src/renderer.cu:426-441 */
__global__ __launch_bounds__(256) void k_tonemap(const float4* accum, ilm_pixel* out, int n, int spp, int op) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 a = accum[i];
    float samples = (float)spp;
    V3 avg = mul3(mk3(a.x, a.y, a.z), 1 / samples);
    avg = tonemap_eval(op, avg);
    avg = mk3(__builtin_fmaxf(__builtin_fminf(avg.x, 1.0f), 0.0f), __builtin_fmaxf(__builtin_fminf(avg.y, 1.0f), 0.0f),
              __builtin_fmaxf(__builtin_fminf(avg.z, 1.0f), 0.0f));
    ilm_pixel px;
    px.r = (uint64_t)(int)(255 * avg.x);
    px.g = (uint64_t)(int)(255 * avg.y);
    px.b = (uint64_t)(int)(255 * avg.z);
    out[i] = px;
}

__global__ void k_fill_zero4(float4* p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

/* ---------------- host-side launchers ---------------- */
#include "kernels.h"

hipError_t ilm_k_ctl_init(const KParams& P, hipStream_t s) {
    hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(256), 0, s, P.ctl, P.chunk);
    return hipGetLastError();
}
hipError_t ilm_k_raygen(const KParams& P, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_raygen, dim3(grid), dim3(256), 0, s, P);
    return hipGetLastError();
}
#define EXT_VARIANT(T, Q, C) hipLaunchKernelGGL((k_extend<T, Q, C>), dim3(grid), dim3(EXT_BLOCK), 0, s, P, depth, ovf)
hipError_t ilm_k_extend(const KParams& P, int depth, int* ovf, int grid, hipStream_t s) {
    const int T = P.tri_test == ILM_TRI_WATERTIGHT;
    const int Q = P.bvh_layout == ILM_BVH_QUANT8;
    const int C = P.trav_stats != nullptr;
    const int v = T * 4 + Q * 2 + C;
    switch (v) {
    case 0: EXT_VARIANT(0, 0, 0); break;
    case 1: EXT_VARIANT(0, 0, 1); break;
    case 2: EXT_VARIANT(0, 1, 0); break;
    case 3: EXT_VARIANT(0, 1, 1); break;
    case 4: EXT_VARIANT(1, 0, 0); break;
    case 5: EXT_VARIANT(1, 0, 1); break;
    case 6: EXT_VARIANT(1, 1, 0); break;
    default: EXT_VARIANT(1, 1, 1); break;
    }
    return hipGetLastError();
}
hipError_t ilm_k_shade(const KParams& P, int depth, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_shade, dim3(grid), dim3(SHADE_BLOCK), 0, s, P, depth);
    return hipGetLastError();
}
hipError_t ilm_k_accumulate(const KParams& P, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_accumulate, dim3(grid), dim3(256), 0, s, P);
    return hipGetLastError();
}
hipError_t ilm_k_chunk_end(const KParams& P, unsigned long long* seg_total, hipStream_t s) {
    hipLaunchKernelGGL(k_chunk_end, dim3(1), dim3(64), 0, s, P.ctl, P.chunk, seg_total);
    return hipGetLastError();
}
hipError_t ilm_k_tonemap(const float4* accum, ilm_pixel* out, int n, int spp, int op, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tonemap, dim3((n + 255) / 256), dim3(256), 0, s, accum, out, n, spp, op);
    return hipGetLastError();
}
hipError_t ilm_k_fill_zero4(float4* p, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    size_t g = (n + 255) / 256;
    if (g > 65535) g = 65535;
    hipLaunchKernelGGL(k_fill_zero4, dim3((unsigned)g), dim3(256), 0, s, p, n);
    return hipGetLastError();
}
int ilm_k_extend_occupancy(int tri_test, int bvh_layout) {
    int nb = 0;
    hipError_t e;
    if (!tri_test && !bvh_layout) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_extend<0, 0, 0>, EXT_BLOCK, 0);
    else if (tri_test && !bvh_layout) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_extend<1, 0, 0>, EXT_BLOCK, 0);
    else if (!tri_test && bvh_layout) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_extend<0, 1, 0>, EXT_BLOCK, 0);
    else e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_extend<1, 1, 0>, EXT_BLOCK, 0);
    if (e != hipSuccess || nb <= 0) nb = 8;
    return nb;
}
int ilm_k_ovf_stack() { return OVF_STACK; }
int ilm_k_lds_stack() { return LDS_STACK; }
int ilm_k_ext_block() { return EXT_BLOCK; }
