/*
 * engine.hip - host runtime behind the C ABI: context, device-resident scenes, the chunked
 * wavefront driver (direct launches or one hipGraph per depth count), the plain, batched and
 * served forms, and the parity hooks.
 *
 * Reference counterparts: Renderer::render_to_bitmap (src/renderer.cu:498-597) and
 * SceneRef (src/scene/sceneref.cpp:9-148). Differences by design: scenes stay on the device
 * across renders, path state is struct-of-arrays and compacted on the device, many passes
 * (and many requests) share one chunk of launches, and nothing synchronises with the host
 * between depths.
 */
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "bvh.h"
#include "bvh_cache.h"
#include "engine.h"
#include "hostmath.h"
#include "kernels.h"

using namespace ilm;

static thread_local std::string g_err;

/* ======================= no-allocation scratch ======================= */

/* Device memory of ilm_render_noalloc: path state, hits, per-path results, control block, chunk
 * descriptor, traversal-stack spill area, Kahan sums and pixels of one render, all carved from
 * this array. Each GPU the library runs on holds its own copy until the process ends. In host
 * code the name is a host-side stand-in, never a device address: kernels and copies use the
 * address hipGetSymbolAddress gives for the current device (noalloc_scratch_of). */
static __device__ __align__(256) unsigned char scratch[ILM_NOALLOC_SCRATCH_BYTES];

#define ILM_MAX_DEVICES 64
static std::mutex g_scratch_mu;                     /* g_scratch_ptr */
static void* g_scratch_ptr[ILM_MAX_DEVICES];
/* held for the whole of an ilm_render_noalloc call: calls on one GPU share its array */
static std::mutex g_noalloc_mu[ILM_MAX_DEVICES];
/* held while a graph is captured on the device, and around the calls that wait for the whole
 * device (hipFree): the runtime refuses those while any stream of the device is capturing, and
 * contexts on one device may capture and free from different threads. A leaf lock. */
static std::mutex g_capture_mu[ILM_MAX_DEVICES];
static std::mutex& capture_mu(int device) { return g_capture_mu[device >= 0 && device < ILM_MAX_DEVICES ? device : 0]; }
static void dfree(int device, void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(capture_mu(device));
    (void)hipFree(p);
}

/* This device's address of the scratch array, looked up once per device with the device
 * current (the caller has made it current). */
static hipError_t noalloc_scratch_of(int device, void** out) {
    *out = nullptr;
    if (device < 0 || device >= ILM_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!ptr) return hipErrorInvalidSymbol;
        g_scratch_ptr[device] = ptr;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

const char* ilm_global_error(void) { return g_err.c_str(); }
int ilm_version(void) { return ILM_ABI_VERSION; }

static int fail(ilm_ctx* ctx, int code, const std::string& msg) {
    g_err = msg;
    if (ctx) {
        std::lock_guard<std::mutex> lk(ctx->err_mu);
        ctx->err = msg;
    }
    return code;
}
#define HIPTRY(ctx, call)                                                                          \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail((ctx), ILM_E_DEVICE, std::string(#call ": ") + hipGetErrorString(e_));     \
    } while (0)

static double now_ms(const ilm_ctx* ctx) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ctx->t0).count();
}

/* ======================= scenes ======================= */

static hipError_t dalloc(ilm_scene* s, void** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) return hipSuccess;
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipSuccess) {
        s->allocs.push_back(*p);
        s->bytes += (int64_t)bytes;
    }
    return e;
}
/* Blocking copy through the context's copy stream (never the legacy stream). */
static hipError_t copy_sync(ilm_ctx* ctx, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, ctx->copy_stream);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(ctx->copy_stream);
}
template <class T>
static hipError_t dupload(ilm_scene* s, T** dst, const T* src, size_t count) {
    *dst = nullptr;
    if (count == 0) return hipSuccess;
    hipError_t e = dalloc(s, (void**)dst, count * sizeof(T));
    if (e != hipSuccess) return e;
    return copy_sync(s->ctx, *dst, src, count * sizeof(T), hipMemcpyHostToDevice);
}

static int upload_texture(ilm_ctx* ctx, ilm_scene* s, const ilm_texture_desc& t, DTex& d) {
    std::memset(&d, 0, sizeof(d));
    d.type = t.type;
    if (t.type == ILM_TEX_CONSTANT) {
        for (int i = 0; i < 3; i++) d.a[i] = t.color[i];
    } else if (t.type == ILM_TEX_CHECKERBOARD) {
        for (int i = 0; i < 3; i++) { d.a[i] = t.on[i]; d.b[i] = t.off[i]; }
    } else if (t.type == ILM_TEX_BITMAP) {
        if (t.width <= 0 || t.height <= 0) return fail(ctx, ILM_E_INVALID, "bitmap texture without size");
        d.w = t.width; d.h = t.height; d.hdr = t.is_hdr;
        size_t n = (size_t)t.width * t.height;
        if (t.is_hdr) {
            if (!t.texels_f32) return fail(ctx, ILM_E_INVALID, "HDR bitmap without texels");
            float4* p;
            HIPTRY(ctx, dupload(s, &p, (const float4*)t.texels_f32, n));
            d.f = p;
        } else {
            if (!t.texels_u8) return fail(ctx, ILM_E_INVALID, "LDR bitmap without texels");
            uchar4* p;
            HIPTRY(ctx, dupload(s, &p, (const uchar4*)t.texels_u8, n));
            d.u8 = p;
        }
    } else {
        return fail(ctx, ILM_E_INVALID, "unknown texture type");
    }
    d.scale[0] = t.scale[0];
    d.scale[1] = t.scale[1];
    return ILM_OK;
}

static void m4_rows(const hm::M4& m, float* out12) {
    for (int i = 0; i < 3; i++) for (int j = 0; j < 4; j++) out12[4 * i + j] = m.m[i][j];
}

/* outward rounding of a double bound to float */
static float round_down(double v) { float f = (float)v; return (double)f > v ? std::nextafter(f, -INFINITY) : f; }
static float round_up(double v) { float f = (float)v; return (double)f < v ? std::nextafter(f, INFINITY) : f; }

int ilm_scene_upload(ilm_ctx* ctx, const ilm_scene_desc* d, ilm_scene** out) {
    if (!ctx || !d || !out) return fail(ctx, ILM_E_INVALID, "null argument");
    *out = nullptr;
    std::unique_ptr<ilm_scene> s(new ilm_scene());
    s->ctx = ctx;
    HIPTRY(ctx, hipSetDevice(ctx->device));

    /* ---- validation ---- */
    for (int i = 0; i < d->num_meshes; i++) {
        const ilm_mesh_desc& m = d->meshes[i];
        if (m.num_vertices < 0 || m.num_triangles < 0 || (m.num_triangles > 0 && (!m.indices || !m.positions)))
            return fail(ctx, ILM_E_INVALID, "mesh " + std::to_string(i) + ": bad arrays");
        if (!m.face_normals && !m.normals) return fail(ctx, ILM_E_INVALID, "mesh needs normals when face_normals=0");
        if (m.has_texcoords && !m.texcoords) return fail(ctx, ILM_E_INVALID, "mesh needs texcoords");
        if ((int64_t)m.num_triangles >= (1 << 27)) return fail(ctx, ILM_E_UNSUPPORTED, "mesh too large (2^27 triangles)");
        for (int64_t k = 0; k < 3LL * m.num_triangles; k++)
            if (m.indices[k] < 0 || m.indices[k] >= m.num_vertices)
                return fail(ctx, ILM_E_INVALID, "mesh " + std::to_string(i) + ": index out of range");
    }
    for (int i = 0; i < d->num_instances; i++) {
        if (d->instances[i].mesh < 0 || d->instances[i].mesh >= d->num_meshes)
            return fail(ctx, ILM_E_INVALID, "instance mesh out of range");
        if (d->instances[i].material < 0 || d->instances[i].material >= d->num_materials)
            return fail(ctx, ILM_E_INVALID, "instance material out of range");
    }
    for (int i = 0; i < d->num_spheres; i++)
        if (d->spheres[i].material < 0 || d->spheres[i].material >= d->num_materials)
            return fail(ctx, ILM_E_INVALID, "sphere material out of range");
    for (int i = 0; i < d->num_materials; i++)
        if (d->materials[i].medium < -1 || d->materials[i].medium >= d->num_mediums)
            return fail(ctx, ILM_E_INVALID, "material medium out of range");
    if (d->num_instances >= 65536) return fail(ctx, ILM_E_UNSUPPORTED, "too many instances");

    DScene hs;
    std::memset(&hs, 0, sizeof(hs));

    /* ---- meshes: BVH, layouts, triangle slots ---- */
    std::vector<DMesh> dmeshes(d->num_meshes);
    std::vector<Box> mesh_bounds(d->num_meshes);
    int blas_depth = 0, blas_qdepth = 0;
    BuildOpts bo;
    for (int i = 0; i < d->num_meshes; i++) {
        const ilm_mesh_desc& m = d->meshes[i];
        Box mb;
        mb.reset();
        for (int64_t k = 0; k < 3LL * m.num_triangles; k++) mb.grow(&m.positions[3 * m.indices[k]]);
        mesh_bounds[i] = mb;
        MeshBVH bvh;
        bool cached = false;
        uint64_t key = 0;
        if (!ctx->bvh_cache_dir.empty()) {
            key = mesh_key(m.positions, m.num_vertices, m.indices, m.num_triangles, bo);
            cached = cache_load(ctx->bvh_cache_dir, key, m.num_triangles, bvh);
        }
        if (!cached) {
            bvh = build_mesh_bvh(m.positions, m.indices, m.num_triangles, bo);
            if (!ctx->bvh_cache_dir.empty()) (void)cache_store(ctx->bvh_cache_dir, key, bvh);
        }
        if (!ctx->bvh_cache_dir.empty()) {
            std::lock_guard<std::mutex> lk(ctx->stat_mu);
            (cached ? ctx->stat_bvh_hits : ctx->stat_bvh_misses)++;
        }
        blas_depth = std::max(blas_depth, bvh.depth);
        blas_qdepth = std::max(blas_qdepth, bvh.qdepth);
        s->quant_max_exp = std::max(s->quant_max_exp, qnodes_max_exponent(bvh.qnodes));
        std::vector<float4> tris(3 * (size_t)m.num_triangles);
        for (int k = 0; k < m.num_triangles; k++) {
            int t = bvh.order[k];
            const float* p0 = &m.positions[3 * m.indices[3 * t]];
            const float* p1 = &m.positions[3 * m.indices[3 * t + 1]];
            const float* p2 = &m.positions[3 * m.indices[3 * t + 2]];
            float tb;
            int32_t ti = t;
            std::memcpy(&tb, &ti, 4);
            tris[3 * k] = make_float4(p0[0], p0[1], p0[2], tb);
            tris[3 * k + 1] = make_float4(p1[0], p1[1], p1[2], 0.0f);
            tris[3 * k + 2] = make_float4(p2[0], p2[1], p2[2], 0.0f);
        }
        DMesh& dm = dmeshes[i];
        std::memset(&dm, 0, sizeof(dm));
        BNode* bn; QNode* qn; float4* tp; int* ip; float* pp; float* np; float* up;
        HIPTRY(ctx, dupload(s.get(), &bn, bvh.nodes.data(), bvh.nodes.size()));
        HIPTRY(ctx, dupload(s.get(), &qn, bvh.qnodes.data(), bvh.qnodes.size()));
        HIPTRY(ctx, dupload(s.get(), &tp, tris.data(), tris.size()));
        HIPTRY(ctx, dupload(s.get(), &ip, m.indices, 3 * (size_t)m.num_triangles));
        HIPTRY(ctx, dupload(s.get(), &pp, m.positions, 3 * (size_t)m.num_vertices));
        np = nullptr;
        up = nullptr;
        if (!m.face_normals) HIPTRY(ctx, dupload(s.get(), &np, m.normals, 3 * (size_t)m.num_vertices));
        if (m.has_texcoords) HIPTRY(ctx, dupload(s.get(), &up, m.texcoords, 2 * (size_t)m.num_vertices));
        /* shading vertices by original triangle index: the same floats the kernel would gather
         * through the index array */
        std::vector<float4> sv(3 * (size_t)m.num_triangles);
        const float* va = m.face_normals ? m.positions : m.normals;
        for (int t = 0; t < m.num_triangles; t++)
            for (int k = 0; k < 3; k++) {
                const float* v = &va[3 * (size_t)m.indices[3 * t + k]];
                sv[3 * (size_t)t + k] = make_float4(v[0], v[1], v[2], 0.0f);
            }
        float4* svp;
        HIPTRY(ctx, dupload(s.get(), &svp, sv.data(), sv.size()));
        dm.nodes = bn; dm.qnodes = qn; dm.tris = tp; dm.idx = ip; dm.pos = pp; dm.nrm = np; dm.uv = up;
        dm.svert = svp;
        dm.root = bvh.root;
        dm.qroot = bvh.qroot;
        dm.face_normals = m.face_normals != 0;
        dm.has_texcoords = m.has_texcoords != 0;
        s->mesh_tris += m.num_triangles;
    }

    /* ---- instances and TLAS ---- */
    std::vector<DInst> dinst(d->num_instances);
    std::vector<Box> inst_boxes(d->num_instances);
    s->inst_mats.resize(48 * (size_t)d->num_instances);
    for (int i = 0; i < d->num_instances; i++) {
        hm::M4 mat, inv, tinv;
        hm::build_transform(d->instances[i].transform, mat, inv, tinv);
        DInst& di = dinst[i];
        std::memset(&di, 0, sizeof(di));
        m4_rows(inv, di.inv);
        m4_rows(tinv, di.tinv);
        di.mesh = d->instances[i].mesh;
        di.material = d->instances[i].material;
        di.nodes = dmeshes[di.mesh].nodes;
        di.qnodes = dmeshes[di.mesh].qnodes;
        di.tris = dmeshes[di.mesh].tris;
        di.root = dmeshes[di.mesh].root;
        di.qroot = dmeshes[di.mesh].qroot;
        di.svert = dmeshes[di.mesh].svert;
        di.uv = dmeshes[di.mesh].uv;
        di.idx = dmeshes[di.mesh].idx;
        di.pos = dmeshes[di.mesh].pos;
        di.flags = (d->meshes[di.mesh].face_normals ? ILM_INST_FACE_NORMALS : 0) |
                   (d->meshes[di.mesh].has_texcoords ? ILM_INST_TEXCOORDS : 0);
        for (int k = 0; k < 16; k++) {
            s->inst_mats[48 * i + k] = mat.m[k / 4][k % 4];
            s->inst_mats[48 * i + 16 + k] = inv.m[k / 4][k % 4];
            s->inst_mats[48 * i + 32 + k] = tinv.m[k / 4][k % 4];
        }
        /* conservative world box: 8 transformed corners in double, widened by a relative
         * margin that covers the float inverse used by the traversal */
        const Box& mb = mesh_bounds[di.mesh];
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        bool empty = d->meshes[di.mesh].num_triangles == 0;
        for (int c = 0; c < 8 && !empty; c++) {
            double p[3] = {(c & 1) ? mb.hi[0] : mb.lo[0], (c & 2) ? mb.hi[1] : mb.lo[1], (c & 4) ? mb.hi[2] : mb.lo[2]};
            for (int a = 0; a < 3; a++) {
                double w = (double)mat.m[a][0] * p[0] + (double)mat.m[a][1] * p[1] + (double)mat.m[a][2] * p[2] + (double)mat.m[a][3];
                lo[a] = std::min(lo[a], w);
                hi[a] = std::max(hi[a], w);
            }
        }
        Box& ib = inst_boxes[i];
        if (empty) { ib.reset(); continue; }
        double ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-30});
        for (int a = 0; a < 3; a++) {
            double mag = std::max(std::fabs(lo[a]), std::fabs(hi[a]));
            double pad = 1e-5 * ext + 1e-6 * mag;
            ib.lo[a] = round_down(lo[a] - pad);
            ib.hi[a] = round_up(hi[a] + pad);
        }
    }
    int tlas_depth = 0, tlas_qdepth = 0;
    hs.tlas_root = ILM_NO_CHILD;
    hs.qtlas_root = ILM_NO_CHILD;
    if (d->num_instances > 0) {
        BuildOpts to;
        to.max_leaf = 1;
        Tree tt = build_tree(inst_boxes, to);
        Layout2 T2 = make_layout2(tt);
        LayoutQ TQ = make_layout_q(tt);
        tlas_depth = tt.max_depth;
        tlas_qdepth = TQ.max_depth;
        s->quant_max_exp = std::max(s->quant_max_exp, qnodes_max_exponent(TQ.nodes));
        /* TLAS leaves hold one instance each: store the instance index itself as the leaf's
         * `first`, so that an entry reads the DInst record directly */
        auto direct = [&tt](int32_t c) {
            if (c == ILM_NO_CHILD || c >= 0) return c;
            return ilm_leaf_code((uint32_t)tt.order[ilm_leaf_first(c)], 1u);
        };
        for (BNode& n : T2.nodes) {
            n.d.x = direct(n.d.x);
            n.d.y = direct(n.d.y);
        }
        for (QNode& n : TQ.nodes)
            for (int k = 0; k < 4; k++) n.child[k] = direct(n.child[k]);
        BNode* bn; QNode* qn;
        HIPTRY(ctx, dupload(s.get(), &bn, T2.nodes.data(), T2.nodes.size()));
        HIPTRY(ctx, dupload(s.get(), &qn, TQ.nodes.data(), TQ.nodes.size()));
        hs.tlas = bn; hs.qtlas = qn;
        hs.tlas_root = direct(T2.root);
        hs.qtlas_root = direct(TQ.root);
    }
    /* device stack bound: pushes never exceed the depth of the current node per level */
    const int cap = ilm_k_lds_stack() + ilm_k_ovf_stack();
    s->stack_bound_fp32 = tlas_depth + 1 + blas_depth;
    s->stack_bound_q = 3 * (tlas_qdepth + 1) + 1 + 3 * (blas_qdepth + 1);
    if (s->stack_bound_fp32 > cap || (ctx->opts.bvh_layout == ILM_BVH_QUANT8 && s->stack_bound_q > cap))
        return fail(ctx, ILM_E_UNSUPPORTED, "BVH too deep for the traversal stack");
    if (ctx->opts.bvh_layout == ILM_BVH_QUANT8 && s->quant_max_exp > ILM_QUANT_MAX_EXP)
        return fail(ctx, ILM_E_UNSUPPORTED, "scene extent too large for 8-bit child boxes");

    /* ---- materials, mediums, spheres, environment, camera ---- */
    std::vector<DMat> dmats(d->num_materials);
    for (int i = 0; i < d->num_materials; i++) {
        const ilm_material_desc& m = d->materials[i];
        DMat& dm = dmats[i];
        std::memset(&dm, 0, sizeof(dm));
        if (m.type < ILM_MAT_EMISSIVE || m.type > ILM_MAT_PLASTIC) return fail(ctx, ILM_E_INVALID, "bad material type");
        dm.type = m.type;
        dm.medium = m.medium;
        dm.distribution = m.distribution;
        dm.ior = m.ior;
        for (int k = 0; k < 3; k++) { dm.eta[k] = m.eta[k]; dm.k[k] = m.k[k]; }
        if (m.type == ILM_MAT_PLASTIC) dm.diffuse_fresnel = hm::diffuse_fresnel(1.0f / m.ior, 1000000);  /* src/material.cu:176 */
        int rc = upload_texture(ctx, s.get(), m.albedo, dm.albedo);
        if (rc) return rc;
        rc = upload_texture(ctx, s.get(), m.roughness, dm.roughness);
        if (rc) return rc;
    }
    std::vector<DMedium> dmed(d->num_mediums);
    for (int i = 0; i < d->num_mediums; i++) {
        const ilm_medium_desc& m = d->mediums[i];
        DMedium& dm = dmed[i];
        std::memset(&dm, 0, sizeof(dm));
        for (int k = 0; k < 3; k++) dm.absorption[k] = m.sigma_a[k];
        dm.scattering = m.sigma_s;
        dm.g = m.g;
        float l2 = (m.sigma_a[0] * m.sigma_a[0]) + (m.sigma_a[1] * m.sigma_a[1]) + (m.sigma_a[2] * m.sigma_a[2]);
        dm.active = ((double)l2 > 1e-5 || m.sigma_s > 0);   /* src/medium.cu:12 */
    }
    std::vector<DSphere> dsph(d->num_spheres);
    for (int i = 0; i < d->num_spheres; i++) {
        std::memset(&dsph[i], 0, sizeof(DSphere));
        for (int k = 0; k < 3; k++) dsph[i].c[k] = d->spheres[i].center[k];
        dsph[i].r = d->spheres[i].radius;
        dsph[i].material = d->spheres[i].material;
    }
    if (d->has_environment) {
        int rc = upload_texture(ctx, s.get(), d->environment, hs.env);
        if (rc) return rc;
    } else {
        std::memset(&hs.env, 0, sizeof(hs.env));
        hs.env.type = ILM_TEX_CONSTANT;   /* black (src/texture.h:39-43) */
    }
    {
        hm::M4 mat, inv, tinv;
        hm::build_transform(d->camera.transform, mat, inv, tinv);
        m4_rows(mat, hs.cam);
        hs.dof = d->camera.dof;
        hs.aperture = d->camera.aperture;
        s->camera = d->camera;
    }
    DInst* di; DMesh* dmp; DSphere* dsp; DMat* dma; DMedium* dme;
    HIPTRY(ctx, dupload(s.get(), &di, dinst.data(), dinst.size()));
    HIPTRY(ctx, dupload(s.get(), &dmp, dmeshes.data(), dmeshes.size()));
    HIPTRY(ctx, dupload(s.get(), &dsp, dsph.data(), dsph.size()));
    HIPTRY(ctx, dupload(s.get(), &dma, dmats.data(), dmats.size()));
    HIPTRY(ctx, dupload(s.get(), &dme, dmed.data(), dmed.size()));
    hs.inst = di; hs.meshes = dmp; hs.spheres = dsp; hs.mats = dma; hs.mediums = dme;
    hs.n_inst = d->num_instances;
    hs.n_spheres = d->num_spheres;
    DScene* ds;
    HIPTRY(ctx, dupload(s.get(), &ds, &hs, 1));
    s->d_scene = ds;
    s->h_scene = hs;
    s->num_instances = d->num_instances;
    s->refs.store(1);
    {
        std::lock_guard<std::mutex> lk(ctx->sched_mu);
        ctx->in_use += s->bytes;
    }
    ctx->refs.fetch_add(1);
    *out = s.release();
    return ILM_OK;
}

void ctx_unref(ilm_ctx* ctx);

/* Drops one reference; the last frees the scene's device memory. sched_locked: the caller holds
 * the context's sched_mu (the scheduler's own release paths). */
static void scene_release(ilm_scene* s, bool sched_locked = false) {
    if (!s) return;
    if (s->refs.fetch_sub(1) != 1) return;
    ilm_ctx* ctx = s->ctx;
    (void)hipSetDevice(ctx->device);
    for (void* p : s->allocs) dfree(ctx->device, p);
    s->allocs.clear();
    (void)hipGetLastError();
    if (sched_locked) {
        ctx->in_use -= s->bytes;
    } else {
        std::lock_guard<std::mutex> lk(ctx->sched_mu);
        ctx->in_use -= s->bytes;
    }
    delete s;
    /* the scheduler thread never holds the last context reference (ilm_close joins it first) */
    ctx_unref(ctx);
}

void ilm_scene_free(ilm_scene* s) { scene_release(s); }
int64_t ilm_scene_device_bytes(const ilm_scene* s) { return s ? s->bytes : -1; }

int ilm_debug_instance_matrices(const ilm_scene* s, int32_t i, float* out48) {
    if (!s || !out48 || i < 0 || i >= s->num_instances) return fail(nullptr, ILM_E_INVALID, "bad instance");
    std::memcpy(out48, &s->inst_mats[48 * (size_t)i], 48 * sizeof(float));
    return ILM_OK;
}

/* ======================= context ======================= */

static void ctx_free_paths(ilm_ctx* ctx) {
    for (void* p : ctx->path_allocs) dfree(ctx->device, p);
    ctx->path_allocs.clear();
    ctx->own_bytes -= ctx->path_bytes;
    ctx->path_bytes = 0;
    ctx->A = DPaths{};
    ctx->B = DPaths{};
    ctx->hits = DHits{};
    ctx->contrib = nullptr;
    ctx->cap = 0;
}

static int ctx_alloc_paths(ilm_ctx* ctx, int64_t cap) {
    auto A = [&](void** p, size_t b) -> hipError_t {
        hipError_t e = hipMalloc(p, b);
        if (e == hipSuccess) { ctx->path_allocs.push_back(*p); ctx->own_bytes += (int64_t)b; ctx->path_bytes += (int64_t)b; }
        return e;
    };
    for (int k = 0; k < ctx->npipes; k++) {
        for (DPaths* P : {k ? &ctx->A2 : &ctx->A, k ? &ctx->B2 : &ctx->B}) {
            HIPTRY(ctx, A((void**)&P->o, cap * sizeof(float4)));
            HIPTRY(ctx, A((void**)&P->d, cap * sizeof(float4)));
            HIPTRY(ctx, A((void**)&P->c, cap * sizeof(float4)));
            HIPTRY(ctx, A((void**)&P->rng, cap * sizeof(uint2)));
        }
        DHits& h = k ? ctx->hits2 : ctx->hits;
        HIPTRY(ctx, A((void**)&h.h, cap * sizeof(float4)));
        HIPTRY(ctx, A((void**)&h.inst, cap * sizeof(int)));
        HIPTRY(ctx, A((void**)(k ? &ctx->contrib2 : &ctx->contrib), cap * sizeof(float4)));
    }
    ctx->cap = cap;
    return ILM_OK;
}

static bool opts_valid(const ilm_opts& o) {
    const bool sg_ok = o.sample_group >= 0 && o.sample_group <= 4096 && (o.sample_group & (o.sample_group - 1)) == 0;
    return !(o.bvh_layout < 0 || o.bvh_layout > ILM_BVH_QUANT8 || o.tri_test < 0 || o.tri_test > ILM_TRI_WATERTIGHT ||
             o.extend_mode < 0 || o.extend_mode > ILM_EXTEND_STEP_REFILL || !sg_ok || o.tile_rows < 0 || o.tile_rows > 256 ||
             o.refill_min < 0 || o.refill_min > 64 || o.ext_fetch < 0 || o.ext_fetch > 65536 || (o.ext_fetch & 63) != 0 ||
             o.leaf_wait_pct < 0 || o.leaf_wait_pct > 10000 || o.shade_sort < 0 || o.shade_sort > 2 ||
             o.sched_fair < 0 || o.sched_fair > 1 || o.camera_packets < 0 || o.camera_packets > 2 ||
             o.chunk_passes < 0 || o.paths_in_flight < 0 || o.hbm_budget_bytes < 0 || o.max_queued < 0 ||
             o.pipes < 0 || o.pipes > 2 || o.lone_split < 0 || o.lone_split > 2);
}

/* extend scheduling actually used (ILM_EXTEND_DEFAULT is the per-step refill) */
static int ext_mode_of(const ilm_opts& o) {
    return o.extend_mode == ILM_EXTEND_DEFAULT ? ILM_EXTEND_STEP_REFILL : o.extend_mode;
}

static int64_t cap_of(const ilm_opts& o) {
    /* 2^26 paths (~9.5 GB of path state) hold a whole 1024x1024x64 render in one chunk: every
     * depth is one launch over all live paths, so launch tails are amortised */
    int64_t cap = o.paths_in_flight > 0 ? o.paths_in_flight : (int64_t)1 << 26;
    cap = std::max<int64_t>(cap, 4096);
    return std::min<int64_t>(cap, (int64_t)1 << 30);
}

/* the traversal-stack spill area: one half per chunk pipeline */
static size_t ovf_pipe_ints(const ilm_ctx* ctx) {
    return (size_t)ctx->ovf_grid * ilm_k_ext_block() * ilm_k_ovf_stack();
}
static size_t ovf_bytes(const ilm_ctx* ctx) {
    return (size_t)ctx->npipes * ovf_pipe_ints(ctx) * sizeof(int);
}

int ilm_open(int device, const ilm_opts* opts, ilm_ctx** out) {
    if (!out) return fail(nullptr, ILM_E_INVALID, "null out");
    *out = nullptr;
    if (opts && !opts_valid(*opts)) return fail(nullptr, ILM_E_INVALID, "bad ilm_opts field");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) return fail(nullptr, ILM_E_DEVICE, "no HIP device available");
    if (device < 0 || device >= n) return fail(nullptr, ILM_E_INVALID, "device index out of range");
    std::unique_ptr<ilm_ctx> c(new ilm_ctx());
    ilm_ctx* ctx = c.get();
    ctx->device = device;
    ctx->t0 = std::chrono::steady_clock::now();
    std::memset(&ctx->opts, 0, sizeof(ctx->opts));
    ctx->opts.use_graphs = 1;
    if (opts) ctx->opts = *opts;
    if (ctx->opts.bvh_cache_dir) ctx->bvh_cache_dir = ctx->opts.bvh_cache_dir;
    ctx->opts.bvh_cache_dir = nullptr;   /* the caller's string is not kept */
    HIPTRY(ctx, hipSetDevice(device));
    hipDeviceProp_t prop;
    HIPTRY(ctx, hipGetDeviceProperties(&prop, device));
    ctx->n_cu = prop.multiProcessorCount;
    HIPTRY(ctx, hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    HIPTRY(ctx, hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
    ctx->npipes = ctx->opts.pipes == 1 ? 1 : 2;
    if (ctx->npipes == 2) HIPTRY(ctx, hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking));
    for (int i = 0; i < 2; i++) HIPTRY(ctx, hipEventCreateWithFlags(&ctx->acc_ev[i], hipEventDisableTiming));
    HIPTRY(ctx, hipEventCreateWithFlags(&ctx->fill_ev, hipEventDisableTiming));
    int occ = ilm_k_extend_occupancy(ctx->opts.tri_test, ctx->opts.bvh_layout, ext_mode_of(ctx->opts));
    ctx->ext_grid = ctx->n_cu * std::max(1, occ);
    ctx->gen_grid = ctx->n_cu * 8;
    /* 512-thread blocks, block-stride loop: one resident wave of blocks, no stragglers */
    ctx->shade_grid = ctx->n_cu * ilm_k_shade_occupancy(ctx->opts.shade_sort != 2);
    int rc = ctx_alloc_paths(ctx, cap_of(ctx->opts));
    if (rc) return rc;
    auto A = [&](void** p, size_t b) -> hipError_t {
        hipError_t e2 = hipMalloc(p, b);
        if (e2 == hipSuccess) { ctx->allocs.push_back(*p); ctx->own_bytes += (int64_t)b; }
        return e2;
    };
    ctx->ovf_grid = ctx->ext_grid;
    HIPTRY(ctx, hipMalloc((void**)&ctx->ovf, ovf_bytes(ctx)));
    ctx->own_bytes += (int64_t)ovf_bytes(ctx);
    HIPTRY(ctx, A((void**)&ctx->ctl, sizeof(DCtl)));
    HIPTRY(ctx, A((void**)&ctx->d_chunk, sizeof(DChunk)));
    if (ctx->npipes == 2) {
        HIPTRY(ctx, A((void**)&ctx->ctl2, sizeof(DCtl)));
        HIPTRY(ctx, A((void**)&ctx->d_chunk2, sizeof(DChunk)));
    }
    HIPTRY(ctx, A((void**)&ctx->d_seg_total, 2 * sizeof(unsigned long long)));
    HIPTRY(ctx, A((void**)&ctx->d_trav, 2 * sizeof(unsigned long long)));
    HIPTRY(ctx, hipMemsetAsync(ctx->d_seg_total, 0, 2 * sizeof(unsigned long long), ctx->stream));
    HIPTRY(ctx, hipMemsetAsync(ctx->d_trav, 0, 2 * sizeof(unsigned long long), ctx->stream));
    HIPTRY(ctx, hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < ILM_RING; i++) {
        HIPTRY(ctx, hipHostMalloc((void**)&ctx->h_chunk[i], sizeof(DChunk), hipHostMallocDefault));
        HIPTRY(ctx, hipEventCreateWithFlags(&ctx->ring_ev[i], hipEventDisableTiming));
    }
    HIPTRY(ctx, hipHostMalloc((void**)&ctx->h_u64, 4 * sizeof(unsigned long long), hipHostMallocDefault));
    HIPTRY(ctx, hipHostMalloc((void**)&ctx->h_seg_ring, ILM_RING * sizeof(unsigned long long), hipHostMallocDefault));
    for (int i = 0; i < 2; i++) HIPTRY(ctx, hipEventCreate(&ctx->na_ev[i]));
    HIPTRY(ctx, noalloc_scratch_of(device, &ctx->noalloc_scratch));
    size_t free_b = 0, total_b = 0;
    HIPTRY(ctx, hipMemGetInfo(&free_b, &total_b));
    ctx->budget = ctx->opts.hbm_budget_bytes > 0 ? ctx->opts.hbm_budget_bytes : (int64_t)(0.9 * (double)free_b);
    *out = c.release();
    return ILM_OK;
}

static void server_stop(ilm_ctx* ctx);

static void ctx_destroy(ilm_ctx* ctx) {
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
    for (auto& kv : ctx->graphs) (void)hipGraphExecDestroy(kv.second);
    ctx_free_paths(ctx);
    dfree(ctx->device, ctx->ovf);
    for (auto& kv : ctx->bufs.free) dfree(ctx->device, kv.second);
    ctx->bufs.free.clear();
    for (void* p : ctx->allocs) dfree(ctx->device, p);
    dfree(ctx->device, ctx->seg_count);
    for (int i = 0; i < ILM_RING; i++) {
        if (ctx->h_chunk[i]) (void)hipHostFree(ctx->h_chunk[i]);
        if (ctx->ring_ev[i]) (void)hipEventDestroy(ctx->ring_ev[i]);
    }
    if (ctx->h_u64) (void)hipHostFree(ctx->h_u64);
    if (ctx->h_seg_ring) (void)hipHostFree(ctx->h_seg_ring);
    for (int i = 0; i < 2; i++) {
        if (ctx->na_ev[i]) (void)hipEventDestroy(ctx->na_ev[i]);
        if (ctx->acc_ev[i]) (void)hipEventDestroy(ctx->acc_ev[i]);
    }
    if (ctx->fill_ev) (void)hipEventDestroy(ctx->fill_ev);
    if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
    if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
    (void)hipStreamDestroy(ctx->stream);
    (void)hipGetLastError();
    delete ctx;
}

void ctx_unref(ilm_ctx* ctx) {
    if (ctx->refs.fetch_sub(1) == 1) ctx_destroy(ctx);
}

/* Stops the scheduler (queued and running requests end cancelled) and drops the caller's
 * reference; scenes and request handles still alive keep the device resources until freed. */
void ilm_close(ilm_ctx* ctx) {
    if (!ctx) return;
    server_stop(ctx);
    {
        std::lock_guard<std::mutex> dev(ctx->dev_mu);
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        if (ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
    }
    ctx_unref(ctx);
}

/* Retune an idle context: bvh_layout, extend_mode, refill_min, ext_fetch, leaf_wait_pct, shade_sort,
 * sched_fair, camera_packets, sample_group, tile_rows, chunk_passes, paths_in_flight and lone_split take the values in *o (the
 * other fields are ignored). Every scene carries both node layouts, so a layout change needs no
 * upload. Path buffers are reallocated when the chunk size changes; captured graphs are dropped. */
int ilm_tune(ilm_ctx* ctx, const ilm_opts* o) {
    if (!ctx || !o) return fail(ctx, ILM_E_INVALID, "null argument");
    if (!opts_valid(*o)) return fail(ctx, ILM_E_INVALID, "bad ilm_opts field");
    std::lock_guard<std::mutex> sched(ctx->sched_mu);
    if (!ctx->pending.empty() || !ctx->active.empty()) return fail(ctx, ILM_E_BUSY, "requests in flight");
    std::lock_guard<std::mutex> dev(ctx->dev_mu);
    HIPTRY(ctx, hipSetDevice(ctx->device));
    HIPTRY(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->stream2) HIPTRY(ctx, hipStreamSynchronize(ctx->stream2));
    for (auto& kv : ctx->graphs) (void)hipGraphExecDestroy(kv.second);
    ctx->graphs.clear();
    if (ext_mode_of(*o) != ext_mode_of(ctx->opts) || o->bvh_layout != ctx->opts.bvh_layout) {
        const int occ = ilm_k_extend_occupancy(ctx->opts.tri_test, o->bvh_layout, ext_mode_of(*o));
        const int grid = ctx->n_cu * std::max(1, occ);
        if (grid > ctx->ovf_grid) {
            /* the spill area is sized for the grid: grow it before the grid grows (on failure
             * the context keeps its old mode, grid and buffer) */
            const size_t nb = (size_t)ctx->npipes * grid * ilm_k_ext_block() * ilm_k_ovf_stack() * sizeof(int);
            int* fresh = nullptr;
            HIPTRY(ctx, hipMalloc((void**)&fresh, nb));
            dfree(ctx->device, ctx->ovf);
            ctx->own_bytes += (int64_t)nb - (int64_t)ovf_bytes(ctx);
            ctx->ovf = fresh;
            ctx->ovf_grid = grid;
        }
        ctx->ext_grid = grid;
    }
    ctx->opts.bvh_layout = o->bvh_layout;
    ctx->opts.extend_mode = o->extend_mode;
    ctx->opts.sample_group = o->sample_group;
    ctx->opts.tile_rows = o->tile_rows;
    ctx->opts.refill_min = o->refill_min;
    ctx->opts.ext_fetch = o->ext_fetch;
    ctx->opts.leaf_wait_pct = o->leaf_wait_pct;
    ctx->opts.shade_sort = o->shade_sort;
    ctx->opts.sched_fair = o->sched_fair;
    ctx->opts.camera_packets = o->camera_packets;
    ctx->shade_grid = ctx->n_cu * ilm_k_shade_occupancy(ctx->opts.shade_sort != 2);
    ctx->opts.chunk_passes = o->chunk_passes;
    ctx->opts.lone_split = o->lone_split;
    const int64_t cap = cap_of(*o);
    if (cap != ctx->cap) {
        ctx_free_paths(ctx);
        int rc = ctx_alloc_paths(ctx, cap);
        if (rc) {
            ctx_free_paths(ctx);
            if (ctx_alloc_paths(ctx, 4096) != ILM_OK) return fail(ctx, ILM_E_NOMEM, "path buffers lost after a failed resize");
            return rc;
        }
    }
    ctx->opts.paths_in_flight = o->paths_in_flight;
    return ILM_OK;
}

const char* ilm_last_error(ilm_ctx* ctx) {
    if (!ctx) return g_err.c_str();
    std::lock_guard<std::mutex> lk(ctx->err_mu);
    ctx->err_copy = ctx->err;
    return ctx->err_copy.c_str();
}

/* ======================= chunk pipeline ======================= */

/* whether a scene needs the spilling traversal stack under this context's layout */
static bool scene_needs_ovf(const ilm_ctx* ctx, const ilm_scene* s) {
    const int bound = ctx->opts.bvh_layout == ILM_BVH_QUANT8 ? s->stack_bound_q : s->stack_bound_fp32;
    return bound > ilm_k_lds_stack();
}
static bool jobs_need_ovf(const ilm_ctx* ctx, const std::vector<Job*>& jobs) {
    for (const Job* j : jobs)
        if (scene_needs_ovf(ctx, j->scene)) return true;
    return false;
}

/* The context's own chunk pipeline k (0, or 1 when npipes == 2). */
static Pipe ctx_pipe(const ilm_ctx* ctx, int k = 0) {
    Pipe q;
    q.idx = k;
    q.stream = k ? ctx->stream2 : ctx->stream;
    q.A = k ? ctx->A2 : ctx->A;
    q.B = k ? ctx->B2 : ctx->B;
    q.hits = k ? ctx->hits2 : ctx->hits;
    q.contrib = k ? ctx->contrib2 : ctx->contrib;
    q.ctl = k ? ctx->ctl2 : ctx->ctl;
    q.chunk = k ? ctx->d_chunk2 : ctx->d_chunk;
    q.ovf = ctx->ovf + (k ? ovf_pipe_ints(ctx) : 0);
    q.ext_grid = ctx->ext_grid;
    q.seg_total = ctx->d_seg_total;
    return q;
}

static KParams base_params(ilm_ctx* ctx, const Pipe& q) {
    KParams P;
    std::memset(&P, 0, sizeof(P));
    P.chunk = q.chunk;
    P.ctl = q.ctl;
    P.in = q.A;
    P.out = q.B;
    P.hits = q.hits;
    P.contrib = q.contrib;
    P.bvh_layout = ctx->opts.bvh_layout;
    P.tri_test = ctx->opts.tri_test;
    P.extend_mode = ext_mode_of(ctx->opts);
    P.tile_rows = ctx->opts.tile_rows > 0 ? ctx->opts.tile_rows : 4;
    P.refill_min = ctx->opts.refill_min > 0 ? ctx->opts.refill_min : 24;
    P.ext_fetch = ctx->opts.ext_fetch > 0 ? ctx->opts.ext_fetch : 256;
    P.leaf_pct = ctx->opts.leaf_wait_pct > 0 ? ctx->opts.leaf_wait_pct : 60;
    P.shade_sort = ctx->opts.shade_sort == 2 ? 0 : 1;
    P.packet0 = ctx->opts.camera_packets == 2 ? 0 : 1;
    P.trav_stats = ctx->opts.count_tests ? ctx->d_trav : nullptr;
    return P;
}
static KParams base_params(ilm_ctx* ctx) { return base_params(ctx, ctx_pipe(ctx)); }

/* Stage a chunk descriptor: pinned ring slot -> the device DChunk dst (ordered on stream s). */
static int stage_chunk(ilm_ctx* ctx, const DChunk& hc, DChunk* dst, hipStream_t s) {
    int slot = ctx->ring_pos++ % ILM_RING;
    HIPTRY(ctx, hipEventSynchronize(ctx->ring_ev[slot]));
    std::memcpy(ctx->h_chunk[slot], &hc, sizeof(DChunk));
    HIPTRY(ctx, hipMemcpyAsync(dst, ctx->h_chunk[slot], sizeof(DChunk), hipMemcpyHostToDevice, s));
    HIPTRY(ctx, hipEventRecord(ctx->ring_ev[slot], s));
    return ILM_OK;
}

/* A chunk's first launch after k_ctl_init. With per-step refill the first depth's extend and
 * shade form the camera rays where they use them, so only the per-path results are cleared
 * (by the packet extend itself when the first depth runs as packets); the other extend modes
 * read the rays k_raygen writes out. */
static hipError_t start_paths(ilm_ctx* ctx, const Pipe& q, KParams& P, hipStream_t s) {
    P.in = q.A;
    P.cam0 = P.extend_mode == ILM_EXTEND_STEP_REFILL ? 1 : 0;
    if (P.cam0 && P.packet0) return hipSuccess;
    return P.cam0 ? ilm_k_path_clear(P, ctx->gen_grid, s) : ilm_k_raygen(P, ctx->gen_grid, s);
}

/* The launch sequence of one chunk with depth_steps depth iterations, in two parts: the depths
 * (ctl init, first rays, extend and shade per depth), then the accumulation into the requests'
 * Kahan sums and the segment count. */
static int enqueue_depths(ilm_ctx* ctx, const Pipe& q, const KParams& base, int depth_steps, ProfileSink* prof) {
    hipStream_t s = q.stream;
    KParams P = base;
    HIPTRY(ctx, ilm_k_ctl_init(P, s));
    HIPTRY(ctx, start_paths(ctx, q, P, s));
    for (int d = 0; d < depth_steps; d++) {
        P.in = (d & 1) ? q.B : q.A;
        P.out = (d & 1) ? q.A : q.B;
        if (prof) {
            hipEvent_t a, b;
            HIPTRY(ctx, hipEventCreate(&a));
            HIPTRY(ctx, hipEventCreate(&b));
            HIPTRY(ctx, hipEventRecord(a, s));
            HIPTRY(ctx, ilm_k_extend(P, d, q.ovf, q.ext_grid, s));
            HIPTRY(ctx, hipEventRecord(b, s));
            prof->ev.push_back({a, b});
        } else {
            HIPTRY(ctx, ilm_k_extend(P, d, q.ovf, q.ext_grid, s));
        }
        HIPTRY(ctx, ilm_k_shade(P, d, ctx->shade_grid, s));
    }
    return ILM_OK;
}
static int enqueue_finish(ilm_ctx* ctx, const Pipe& q, const KParams& base) {
    HIPTRY(ctx, ilm_k_accumulate(base, ctx->gen_grid, q.stream));
    HIPTRY(ctx, ilm_k_chunk_end(base, q.seg_total, q.stream));
    return ILM_OK;
}
/* both parts, on one stream (debug paths and the allocation-free render) */
static int enqueue_pipeline(ilm_ctx* ctx, const Pipe& q, const KParams& base, int depth_steps, ProfileSink* prof) {
    int rc = enqueue_depths(ctx, q, base, depth_steps, prof);
    return rc ? rc : enqueue_finish(ctx, q, base);
}

/* whether every request of the chunk renders the same scene (the extend kernel then reads it
 * once from the chunk descriptor, so a graph captured for it serves any later single-scene chunk) */
static bool chunk_uni_scene(const DChunk& hc) {
    bool uni = hc.n_work > 0;
    for (int k = 1; k < hc.n_work && uni; k++) uni = hc.reqs[hc.work[k].req].scene == hc.reqs[hc.work[0].req].scene;
    return uni && hc.work[0].req == 0;
}

/* Replay (capturing on first use) the graph of one part of a chunk's launch sequence. */
static int run_graph(ilm_ctx* ctx, int key, const Pipe& q, const KParams& P, int D, bool finish) {
    auto it = ctx->graphs.find(key);
    if (it == ctx->graphs.end()) {
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        {
            std::lock_guard<std::mutex> cap(capture_mu(ctx->device));
            HIPTRY(ctx, hipStreamBeginCapture(q.stream, hipStreamCaptureModeThreadLocal));
            int rc2 = finish ? enqueue_finish(ctx, q, P) : enqueue_depths(ctx, q, P, D, nullptr);
            hipError_t ec = hipStreamEndCapture(q.stream, &g);
            if (rc2) return rc2;
            if (ec != hipSuccess) return fail(ctx, ILM_E_DEVICE, std::string("graph capture: ") + hipGetErrorString(ec));
        }
        HIPTRY(ctx, hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
        (void)hipGraphDestroy(g);
        it = ctx->graphs.emplace(key, ge).first;
    }
    HIPTRY(ctx, hipGraphLaunch(it->second, q.stream));
    return ILM_OK;
}

/* Launch a chunk on the next pipeline. Its depths start at once, beside whatever the other
 * pipeline still runs; its accumulation waits for the previous chunk's (acc_ev of the other
 * pipeline) and for the latest clear of request buffers (fill_ev), so every pixel's sums grow
 * in pass order. Profiling and test counting keep one pipeline. Sets *pipe_used. */
static int launch_chunk(ilm_ctx* ctx, const DChunk& hc, bool ovf, ProfileSink* prof, int* pipe_used = nullptr) {
    const bool two = ctx->npipes == 2 && !prof && !ctx->opts.count_tests;
    const int k = two ? ctx->next_pipe : 0;
    if (two) ctx->next_pipe ^= 1;
    const Pipe q = ctx_pipe(ctx, k);
    int rc = stage_chunk(ctx, hc, q.chunk, q.stream);
    if (rc) return rc;
    KParams P = base_params(ctx, q);
    P.use_ovf = ovf ? 1 : 0;
    const bool uni = chunk_uni_scene(hc);
    P.uni_scene = uni ? 1 : 0;
    const int D = hc.depth_steps;
    const bool graphs = ctx->opts.use_graphs && !prof && !ctx->opts.count_tests;
    if (graphs) rc = run_graph(ctx, 2 * (8 * D + 4 * k + (uni ? 2 : 0) + (ovf ? 1 : 0)), q, P, D, false);
    else rc = enqueue_depths(ctx, q, P, D, prof);
    if (rc) return rc;
    if (ctx->npipes == 2) {
        HIPTRY(ctx, hipStreamWaitEvent(q.stream, ctx->acc_ev[k ^ 1], 0));
        if (k == 1) HIPTRY(ctx, hipStreamWaitEvent(q.stream, ctx->fill_ev, 0));
    }
    if (graphs) rc = run_graph(ctx, 2 * k + 1, q, P, D, true);
    else rc = enqueue_finish(ctx, q, P);
    if (rc) return rc;
    HIPTRY(ctx, hipEventRecord(ctx->acc_ev[k], q.stream));
    ctx->last_pipe = k;
    if (pipe_used) *pipe_used = k;
    return ILM_OK;
}

/* ======================= jobs ======================= */

/* The request's device descriptor, over its Kahan buffer d_accum (2*W*H float4: sums, then
 * compensations); the issue cursor starts at pass 0. */
static void job_fill_req(Job* j) {
    const ilm_render_params& p = j->p;
    const int64_t wh = (int64_t)p.width * p.height;
    std::memset(&j->dreq, 0, sizeof(j->dreq));
    j->dreq.scene = j->scene->d_scene;
    j->dreq.width = p.width;
    j->dreq.height = p.height;
    j->dreq.max_depth = p.max_depth;
    j->dreq.ray_bias = p.ray_bias;
    j->dreq.tonemap = p.tonemap;
    j->dreq.spp = p.spp;
    hm::render_info(j->scene->camera, p.width, p.height, j->dreq.image_width, j->dreq.pixel_size, j->dreq.left, j->dreq.top);
    j->dreq.accum = j->d_accum;
    j->dreq.comp = j->d_accum + wh;
    j->next_pass = 0;
    j->next_px = 0;
}

static int job_prepare(ilm_ctx* ctx, Job* j) {
    const ilm_render_params& p = j->p;
    const int64_t wh = (int64_t)p.width * p.height;
    const size_t ab = 2 * wh * sizeof(float4), pb = wh * sizeof(ilm_pixel);
    j->pool = &ctx->bufs;
    j->d_accum = (float4*)ctx->bufs.take(ab, &j->accum_cap);
    if (!j->d_accum) {
        HIPTRY(ctx, hipMalloc((void**)&j->d_accum, ab));
        j->accum_cap = ab;
    }
    j->d_px = (ilm_pixel*)ctx->bufs.take(pb, &j->px_cap);
    if (!j->d_px) {
        HIPTRY(ctx, hipMalloc((void**)&j->d_px, pb));
        j->px_cap = pb;
    }
    HIPTRY(ctx, ilm_k_fill_zero4(j->d_accum, 2 * wh, ctx->stream));
    HIPTRY(ctx, hipEventRecord(ctx->fill_ev, ctx->stream));
    job_fill_req(j);
    return ILM_OK;
}

/* Return the request's buffers to the pool (stream order keeps a later user of the same buffer
 * behind every kernel already queued on it). */
static void job_release_device(Job* j) {
    const int dev = j->scene ? j->scene->ctx->device : 0;
    if (j->d_accum && !(j->pool && j->pool->give(j->d_accum, j->accum_cap))) dfree(dev, j->d_accum);
    if (j->d_px && !(j->pool && j->pool->give(j->d_px, j->px_cap))) dfree(dev, j->d_px);
    j->d_accum = nullptr;
    j->d_px = nullptr;
}

static bool job_has_work(const Job* j) { return j->next_pass < j->p.spp; }

/* Fill a chunk from jobs in priority order. Each job contributes at most one slice per chunk
 * (all pixels x several passes, or a pixel range of one pass), so no pixel appears twice in a
 * chunk and every pixel's passes are accumulated in pass order. */
static int64_t job_demand(const Job* j) {
    const int64_t wh = (int64_t)j->p.width * j->p.height;
    return (int64_t)(j->p.spp - j->next_pass) * wh - j->next_px;
}

/* Fair-share quotas (water filling): jobs needing less than an equal share of the chunk get
 * what they need, the others split the rest equally. */
static void fair_quotas(const std::vector<Job*>& jobs, int64_t cap, std::vector<int64_t>& quota) {
    std::vector<std::pair<int64_t, size_t>> d;
    for (size_t k = 0; k < jobs.size(); k++) {
        quota[k] = 0;
        if (job_has_work(jobs[k]) && !jobs[k]->cancel.load()) d.push_back({job_demand(jobs[k]), k});
    }
    std::sort(d.begin(), d.end());
    int64_t left = cap;
    size_t n = d.size();
    for (auto& e : d) {
        const int64_t level = left / (int64_t)std::max<size_t>(n, 1);
        const int64_t q = std::min(e.first, std::max<int64_t>(level, 1));
        quota[e.second] = q;
        left -= q;
        n--;
    }
}

/* The divisors a request's pixel walk needs on the device (tile_pixel with th rows). */
static void req_divs(DReq& R, int th) {
    const uint32_t w = (uint32_t)R.width, h = (uint32_t)R.height, t = th > 1 ? (uint32_t)th : 1u;
    R.tile_th = t;
    R.tile_full = (h / t) * (t * w);
    R.w_div = udiv_make(w);
    R.band_div = udiv_make(t * w);
    R.th_div = udiv_make(t);
}

static void chunk_reset(DChunk& c) {
    c.total_paths = 0;
    c.total_px = 0;
    c.n_work = 0;
    c.depth_steps = 0;
}

/* Append job j's next slice, at most lim paths, as request nreq of chunk c: all pixels x as
 * many whole passes as fit, or a pixel range of one pass. Advances j's issue cursor. */
static void add_slice(const ilm_ctx* ctx, Job* j, int64_t lim, int nreq, DChunk& c) {
    const int64_t wh = (int64_t)j->p.width * j->p.height;
    DWork w;
    std::memset(&w, 0, sizeof(w));
    w.req = nreq;
    w.pass0 = j->next_pass;
    if (j->split_npx > 0) {
        /* a range of pixel slots with all of the request's passes (lim is not binding: the
         * caller chose split_npx so that one range fills at most a chunk) */
        const int64_t npx = std::min<int64_t>(wh - j->next_px, j->split_npx);
        w.npass = j->p.spp;
        w.px0 = (uint32_t)j->next_px;
        w.npx = (uint32_t)npx;
        j->next_px += npx;
        if (j->next_px >= wh) { j->next_px = 0; j->next_pass = j->p.spp; }
    } else if (j->next_px == 0 && wh <= lim) {
        int64_t np = std::min<int64_t>(j->p.spp - j->next_pass, lim / wh);
        if (ctx->opts.chunk_passes > 0) np = std::min<int64_t>(np, ctx->opts.chunk_passes);
        w.npass = (int)np;
        w.px0 = 0;
        w.npx = (uint32_t)wh;
        j->next_pass += (int)np;
    } else {
        int64_t npx = std::min<int64_t>(wh - j->next_px, lim);
        w.npass = 1;
        w.px0 = (uint32_t)j->next_px;
        w.npx = (uint32_t)npx;
        j->next_px += npx;
        if (j->next_px >= wh) { j->next_px = 0; j->next_pass += 1; }
    }
    w.sgroup = 1;
    {
        const uint32_t sg_max = ctx->opts.sample_group > 0 ? (uint32_t)ctx->opts.sample_group : 64u;
        while (w.sgroup * 2 <= sg_max && w.npass % (w.sgroup * 2) == 0) w.sgroup *= 2;
    }
    w.sg_log2 = (uint32_t)__builtin_ctz(w.sgroup);
    w.span_div = udiv_make(w.npx * w.sgroup);
    w.pid_base = c.total_paths;
    w.acc_base = c.total_px;
    c.total_paths += (uint32_t)((int64_t)w.npass * w.npx);
    c.total_px += w.npx;
    c.reqs[nreq] = j->dreq;
    req_divs(c.reqs[nreq], ctx->opts.tile_rows > 0 ? ctx->opts.tile_rows : 4);
    c.work[c.n_work++] = w;
    c.depth_steps = std::max(c.depth_steps, j->p.max_depth);
}

static void build_chunk(ilm_ctx* ctx, const std::vector<Job*>& jobs, DChunk& c, std::vector<Job*>& in_chunk) {
    chunk_reset(c);
    in_chunk.clear();
    int64_t left = ctx->cap;
    int nreq = 0;
    std::vector<int64_t> quota(jobs.size(), ctx->cap);
    if (ctx->opts.sched_fair && jobs.size() > 1) fair_quotas(jobs, ctx->cap, quota);
    /* First the requests that render the scene of the first request with work, so that a chunk
     * usually holds one scene (the extend kernel then reads that scene's values once per wave);
     * requests of other scenes join only a chunk that would otherwise stay under half full. */
    const ilm_scene* s0 = nullptr;
    for (Job* j : jobs)
        if (job_has_work(j) && !j->cancel.load()) { s0 = j->scene; break; }
    for (int round = 0; round < 2; round++) {
        if (round == 1 && (int64_t)c.total_paths * 2 >= ctx->cap) break;
        for (size_t jk = 0; jk < jobs.size(); jk++) {
            Job* j = jobs[jk];
            if (!job_has_work(j) || j->cancel.load()) continue;
            if ((j->scene == s0) != (round == 0)) continue;
            if (nreq >= ILM_MAX_JOBS || c.n_work >= ILM_MAX_WORK || left <= 0) break;
            const int64_t lim = std::min(left, std::max<int64_t>(quota[jk], 1));
            const uint32_t before = c.total_paths;
            add_slice(ctx, j, lim, nreq, c);
            left -= (int64_t)(c.total_paths - before);
            nreq++;
            in_chunk.push_back(j);
        }
    }
}

static int validate_params(ilm_ctx* ctx, const ilm_render_params* p) {
    if (!p) return fail(ctx, ILM_E_INVALID, "null params");
    if (p->width <= 0 || p->height <= 0 || p->spp <= 0 || p->max_depth <= 0)
        return fail(ctx, ILM_E_INVALID, "width, height, spp and max_depth must be positive");
    if (p->max_depth > ILM_MAX_DEPTH) return fail(ctx, ILM_E_INVALID, "max_depth above 256");
    if ((int64_t)p->width * p->height >= (1LL << 31)) return fail(ctx, ILM_E_INVALID, "image too large");
    if (p->tonemap < 0 || p->tonemap > ILM_TONEMAP_UNCHARTED) return fail(ctx, ILM_E_INVALID, "bad tonemap");
    if (p->first_sample < 0) return fail(ctx, ILM_E_INVALID, "negative first_sample");
    return ILM_OK;
}

/* A scene uploaded under one layout may not suit the other (ilm_tune can switch layouts after
 * the upload): a BLAS too deep for its stack, or 8-bit child boxes past the largest quantum. */
static int check_scene_stack(ilm_ctx* ctx, const ilm_scene* s) {
    const int cap = ilm_k_lds_stack() + ilm_k_ovf_stack();
    const bool q = ctx->opts.bvh_layout == ILM_BVH_QUANT8;
    if ((q ? s->stack_bound_q : s->stack_bound_fp32) > cap)
        return fail(ctx, ILM_E_UNSUPPORTED, "BVH too deep for the traversal stack");
    if (q && s->quant_max_exp > ILM_QUANT_MAX_EXP)
        return fail(ctx, ILM_E_UNSUPPORTED, "scene extent too large for 8-bit child boxes");
    return ILM_OK;
}

static void accum4_to_rgb(const std::vector<float4>& a4, float* rgb, int64_t n) {
    for (int64_t i = 0; i < n; i++) { rgb[3 * i] = a4[i].x; rgb[3 * i + 1] = a4[i].y; rgb[3 * i + 2] = a4[i].z; }
}

/* a lone request of at least this many paths is rendered as two chunks (run_jobs_sync) */
#define ILM_SPLIT_MIN_PATHS (1LL << 22)
#define ILM_SPLIT_MIN_DEPTH 24

/* Run jobs to completion on the calling thread (plain and batched forms). */
static int run_jobs_sync(ilm_ctx* ctx, std::vector<Job*>& jobs, ilm_render_info* info) {
    std::lock_guard<std::mutex> dev(ctx->dev_mu);
    HIPTRY(ctx, hipSetDevice(ctx->device));
    for (Job* j : jobs) {
        int rc = job_prepare(ctx, j);
        if (rc) return rc;
        /* pass i of the render seeds wang_hash(first_sample + i) */
        j->next_pass = 0;
    }
    /* A lone request that fits one chunk may run as two chunks of half its pixels (all passes
     * each) on the two pipelines, so that each chunk's latency-bound last launches run beside
     * the other's (lone_split). Splitting by pixels keeps a pixel's samples together in the path
     * order, which the first depth's camera packets rely on; splitting by passes would halve the
     * sample group. Each pixel's passes stay in one chunk, so the sums are unchanged. Deep
     * renders gain (960x720, 32 spp, depth 45 in a subsurface medium: +3 %); shallow ones lose
     * a little to the two chunks' full launches competing (1024x1024, 64 spp, depth 16: -1.6 %),
     * hence the default's depth test. */
    if (jobs.size() == 1 && ctx->npipes == 2 && !ctx->opts.profile && !ctx->opts.count_tests &&
        ctx->opts.chunk_passes == 0 && ctx->opts.lone_split != 1) {
        Job* j = jobs[0];
        const int64_t wh = (int64_t)j->p.width * j->p.height;
        const int64_t paths = wh * j->p.spp;
        const bool deep = ctx->opts.lone_split == 2 || j->p.max_depth >= ILM_SPLIT_MIN_DEPTH;
        if (deep && paths <= ctx->cap && paths >= ILM_SPLIT_MIN_PATHS)
            j->split_npx = (((wh + 1) / 2) + 63) & ~(int64_t)63;
    }
    ProfileSink prof;
    hipEvent_t e0, e1;
    HIPTRY(ctx, hipEventCreate(&e0));
    HIPTRY(ctx, hipEventCreate(&e1));
    /* the segment counter is cumulative; read it before and after */
    HIPTRY(ctx, hipMemcpyAsync(ctx->h_u64 + 3, ctx->d_seg_total, sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx->stream));
    HIPTRY(ctx, hipMemsetAsync(ctx->d_trav, 0, 2 * sizeof(unsigned long long), ctx->stream));
    HIPTRY(ctx, hipEventRecord(e0, ctx->stream));
    if (ctx->stream2) HIPTRY(ctx, hipStreamWaitEvent(ctx->stream2, e0, 0));
    DChunk* hc = new DChunk();
    std::unique_ptr<DChunk> hc_own(hc);
    std::vector<Job*> in_chunk;
    int chunks = 0;
    int64_t max_paths = 0;
    for (;;) {
        build_chunk(ctx, jobs, *hc, in_chunk);
        if (in_chunk.empty()) break;
        /* pass offsets: first_sample shifts every pass of the request */
        for (int w = 0; w < hc->n_work; w++) hc->work[w].pass0 += in_chunk[hc->work[w].req]->p.first_sample;
        int rc = launch_chunk(ctx, *hc, jobs_need_ovf(ctx, in_chunk), ctx->opts.profile ? &prof : nullptr);
        if (rc) return rc;
        chunks++;
        max_paths = std::max<int64_t>(max_paths, hc->total_paths);
    }
    /* the latest accumulation follows every earlier one, and each chunk's depths precede its
     * accumulation on its stream */
    if (chunks && ctx->last_pipe != 0) HIPTRY(ctx, hipStreamWaitEvent(ctx->stream, ctx->acc_ev[ctx->last_pipe], 0));
    for (Job* j : jobs) {
        const int64_t wh = (int64_t)j->p.width * j->p.height;
        HIPTRY(ctx, ilm_k_tonemap(j->d_accum, j->d_px, (int)wh, j->p.spp, j->p.tonemap, ctx->stream));
    }
    HIPTRY(ctx, hipEventRecord(e1, ctx->stream));   /* device_ms: every kernel of the render, tonemap included */
    HIPTRY(ctx, hipMemcpyAsync(ctx->h_u64, ctx->d_seg_total, sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx->stream));
    HIPTRY(ctx, hipMemcpyAsync(ctx->h_u64 + 1, ctx->d_trav, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx->stream));
    HIPTRY(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.0f;
    HIPTRY(ctx, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    double ext_ms = 0.0;
    for (auto& pr : prof.ev) {
        float m = 0.0f;
        (void)hipEventElapsedTime(&m, pr.first, pr.second);
        ext_ms += m;
        (void)hipEventDestroy(pr.first);
        (void)hipEventDestroy(pr.second);
    }
    const unsigned long long segs = ctx->h_u64[0] - ctx->h_u64[3];
    if (info) {
        info->segments = segs;
        info->device_ms = ms;
        info->extend_ms = ext_ms;
        info->extend_launches = (int32_t)prof.ev.size();
        info->chunks = chunks;
        info->paths_per_chunk = max_paths;
        info->box_tests = ctx->h_u64[1];
        info->tri_tests = ctx->h_u64[2];
    }
    {
        std::lock_guard<std::mutex> lk(ctx->stat_mu);   /* dev_mu is held: never sched_mu here */
        ctx->stat_busy_ms += ms;
        ctx->stat_segments += segs;
    }
    return ILM_OK;
}

static int copy_job_outputs(ilm_ctx* ctx, Job* j, float* accum_rgb, ilm_pixel* out) {
    const int64_t wh = (int64_t)j->p.width * j->p.height;
    if (accum_rgb) {
        std::vector<float4> a4(wh);
        HIPTRY(ctx, copy_sync(ctx, a4.data(), j->d_accum, wh * sizeof(float4), hipMemcpyDeviceToHost));
        accum4_to_rgb(a4, accum_rgb, wh);
    }
    if (out) HIPTRY(ctx, copy_sync(ctx, out, j->d_px, wh * sizeof(ilm_pixel), hipMemcpyDeviceToHost));
    return ILM_OK;
}

int ilm_render(ilm_ctx* ctx, ilm_scene* scene, const ilm_render_params* params, float* accum_rgb, ilm_pixel* out,
               ilm_render_info* info) {
    if (!ctx || !scene) return fail(ctx, ILM_E_INVALID, "null context or scene");
    int rc = validate_params(ctx, params);
    if (!rc) rc = check_scene_stack(ctx, scene);
    if (rc) return rc;
    Job j;
    j.scene = scene;
    j.p = *params;
    std::vector<Job*> jobs = {&j};
    if (info) std::memset(info, 0, sizeof(*info));
    rc = run_jobs_sync(ctx, jobs, info);
    if (!rc) rc = copy_job_outputs(ctx, &j, accum_rgb, out);
    job_release_device(&j);
    return rc;
}

int ilm_render_batch(ilm_ctx* ctx, int32_t n, ilm_scene* const* scenes, const ilm_render_params* params,
                     float* const* accum_rgb, ilm_pixel* const* out, ilm_render_info* info) {
    if (!ctx || n <= 0 || !scenes || !params) return fail(ctx, ILM_E_INVALID, "bad batch arguments");
    std::vector<Job> js(n);
    std::vector<Job*> jobs(n);
    for (int i = 0; i < n; i++) {
        int rc = validate_params(ctx, &params[i]);
        if (rc) return rc;
        if (!scenes[i]) return fail(ctx, ILM_E_INVALID, "null scene in batch");
        rc = check_scene_stack(ctx, scenes[i]);
        if (rc) return rc;
        js[i].scene = scenes[i];
        js[i].p = params[i];
        jobs[i] = &js[i];
    }
    if (info) std::memset(info, 0, sizeof(*info));
    int rc = ILM_OK;
    /* requests beyond one chunk's job table run in consecutive groups */
    for (int g = 0; g < n && !rc; g += ILM_MAX_JOBS) {
        std::vector<Job*> grp(jobs.begin() + g, jobs.begin() + std::min(n, g + ILM_MAX_JOBS));
        ilm_render_info gi;
        std::memset(&gi, 0, sizeof(gi));
        rc = run_jobs_sync(ctx, grp, &gi);
        if (info) {
            info->segments += gi.segments;
            info->device_ms += gi.device_ms;
            info->extend_ms += gi.extend_ms;
            info->extend_launches += gi.extend_launches;
            info->chunks += gi.chunks;
            info->paths_per_chunk = std::max(info->paths_per_chunk, gi.paths_per_chunk);
            info->box_tests += gi.box_tests;
            info->tri_tests += gi.tri_tests;
        }
        for (size_t k = 0; k < grp.size() && !rc; k++) {
            int i = g + (int)k;
            rc = copy_job_outputs(ctx, grp[k], accum_rgb ? accum_rgb[i] : nullptr, out ? out[i] : nullptr);
        }
        for (Job* j : grp) job_release_device(j);
    }
    return rc;
}

/* ======================= render without allocation ======================= */

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

/* bytes per path of the pipeline buffers: two copies of o, d, c (float4) and rng (uint2), the
 * hit record (float4 + int) and the per-path result (float4) */
#define ILM_PATH_BYTES (2 * (3 * sizeof(float4) + sizeof(uint2)) + sizeof(float4) + sizeof(int) + sizeof(float4))
/* spill area of a scene whose stack bound exceeds the LDS stack: at most this many extend
 * workgroups (one wave each) run when the scratch array holds it */
#define ILM_NOALLOC_SPILL_BLOCKS 1024

/* The scratch array's carving for one render: fixed blocks, the spill area (spill_blocks > 0
 * when the scene needs it), Kahan sums and compensations (2*W*H float4), pixels, then path
 * state for cap paths. Returns false when it does not fit with at least 4096 paths. */
struct NoallocLayout {
    size_t off_ctl, off_chunk, off_seg, off_ovf, off_accum, off_px, off_paths;
    int spill_blocks;
    int64_t cap;
};
static bool noalloc_layout(const ilm_ctx* ctx, int64_t wh, bool spill, NoallocLayout& L) {
    size_t o = 0;
    L.off_ctl = o;    o = align256(o + sizeof(DCtl));
    L.off_chunk = o;  o = align256(o + sizeof(DChunk));
    L.off_seg = o;    o = align256(o + sizeof(float4));
    L.off_ovf = o;
    L.spill_blocks = spill ? std::min(ctx->ext_grid, ILM_NOALLOC_SPILL_BLOCKS) : 0;
    o = align256(o + (size_t)L.spill_blocks * ilm_k_ext_block() * ilm_k_ovf_stack() * sizeof(int));
    L.off_accum = o;  o = align256(o + 2 * (size_t)wh * sizeof(float4));
    L.off_px = o;     o = align256(o + (size_t)wh * sizeof(ilm_pixel));
    L.off_paths = o;
    if (o >= (size_t)ILM_NOALLOC_SCRATCH_BYTES) return false;
    /* 9 arrays, each rounded up to 256 B */
    const size_t room = (size_t)ILM_NOALLOC_SCRATCH_BYTES - o;
    if (room < 9 * 256) return false;
    int64_t cap = (int64_t)((room - 9 * 256) / ILM_PATH_BYTES);
    cap = std::min<int64_t>(cap & ~(int64_t)63, cap_of(ctx->opts));
    L.cap = cap;
    return cap >= 4096;
}

int ilm_render_noalloc(ilm_ctx* ctx, ilm_scene* scene, const ilm_render_params* params, float* accum_rgb,
                       ilm_pixel* out, ilm_render_info* info) {
    if (!ctx || !scene) return fail(ctx, ILM_E_INVALID, "null context or scene");
    int rc = validate_params(ctx, params);
    if (!rc) rc = check_scene_stack(ctx, scene);
    if (rc) return rc;
    if (!ctx->noalloc_scratch) return fail(ctx, ILM_E_DEVICE, "no scratch address for this device");
    const int64_t wh = (int64_t)params->width * params->height;
    const bool spill = scene_needs_ovf(ctx, scene);
    NoallocLayout L;
    if (!noalloc_layout(ctx, wh, spill, L))
        return fail(ctx, ILM_E_UNSUPPORTED,
                    "image too large for ilm_render_noalloc: " + std::to_string(wh) + " pixels x 56 B plus 4096 paths x " +
                        std::to_string((int)ILM_PATH_BYTES) + " B" + (spill ? " and the traversal-stack spill area" : "") +
                        " exceed the " + std::to_string((long long)ILM_NOALLOC_SCRATCH_BYTES) + "-byte scratch array");
    std::lock_guard<std::mutex> na(g_noalloc_mu[ctx->device]);
    std::lock_guard<std::mutex> dev(ctx->dev_mu);
    HIPTRY(ctx, hipSetDevice(ctx->device));
    unsigned char* base = (unsigned char*)ctx->noalloc_scratch;
    Pipe q;
    q.stream = ctx->stream;
    q.ctl = (DCtl*)(base + L.off_ctl);
    q.chunk = (DChunk*)(base + L.off_chunk);
    q.seg_total = (unsigned long long*)(base + L.off_seg);
    q.ovf = spill ? (int*)(base + L.off_ovf) : nullptr;
    q.ext_grid = spill ? L.spill_blocks : ctx->ext_grid;
    {
        unsigned char* pp = base + L.off_paths;
        auto take = [&](size_t bytes) { void* r = pp; pp += align256(bytes); return r; };
        const size_t n = (size_t)L.cap;
        for (DPaths* D : {&q.A, &q.B}) {
            D->o = (float4*)take(n * sizeof(float4));
            D->d = (float4*)take(n * sizeof(float4));
            D->c = (float4*)take(n * sizeof(float4));
            D->rng = (uint2*)take(n * sizeof(uint2));
        }
        q.hits.h = (float4*)take(n * sizeof(float4));
        q.hits.inst = (int*)take(n * sizeof(int));
        q.contrib = (float4*)take(n * sizeof(float4));
        if (pp > base + (size_t)ILM_NOALLOC_SCRATCH_BYTES) return fail(ctx, ILM_E_DEVICE, "scratch carving overflow");
    }
    /* the array keeps the previous call's contents: the Kahan buffer and the segment counter
     * start at zero; the control block is cleared by every chunk's first kernel */
    float4* d_accum = (float4*)(base + L.off_accum);
    ilm_pixel* d_px = (ilm_pixel*)(base + L.off_px);
    HIPTRY(ctx, ilm_k_fill_zero4(d_accum, 2 * (size_t)wh, ctx->stream));
    HIPTRY(ctx, ilm_k_fill_zero4((float4*)q.seg_total, 1, ctx->stream));
    Job j;
    j.scene = scene;
    j.p = *params;
    j.d_accum = d_accum;
    job_fill_req(&j);
    HIPTRY(ctx, hipEventRecord(ctx->na_ev[0], ctx->stream));
    int chunks = 0;
    int64_t max_paths = 0;
    for (;;) {
        /* the descriptor is built in the pinned staging slot itself (no host allocation) */
        const int slot = ctx->ring_pos++ % ILM_RING;
        HIPTRY(ctx, hipEventSynchronize(ctx->ring_ev[slot]));
        DChunk& hc = *ctx->h_chunk[slot];
        chunk_reset(hc);
        if (!job_has_work(&j)) break;
        add_slice(ctx, &j, L.cap, 0, hc);
        hc.work[0].pass0 += params->first_sample;
        HIPTRY(ctx, hipMemcpyAsync(q.chunk, &hc, sizeof(DChunk), hipMemcpyHostToDevice, ctx->stream));
        HIPTRY(ctx, hipEventRecord(ctx->ring_ev[slot], ctx->stream));
        KParams P = base_params(ctx, q);
        P.trav_stats = nullptr;
        P.use_ovf = spill ? 1 : 0;
        P.uni_scene = 1;
        rc = enqueue_pipeline(ctx, q, P, hc.depth_steps, nullptr);
        if (rc) return rc;
        chunks++;
        max_paths = std::max<int64_t>(max_paths, hc.total_paths);
    }
    HIPTRY(ctx, ilm_k_tonemap(d_accum, d_px, (int)wh, params->spp, params->tonemap, ctx->stream));
    HIPTRY(ctx, hipEventRecord(ctx->na_ev[1], ctx->stream));
    /* packed RGB into the compensation half of the Kahan buffer, no longer needed */
    float* d_rgb = (float*)(d_accum + wh);
    if (accum_rgb) HIPTRY(ctx, ilm_k_sums_to_rgb(d_accum, d_rgb, (int)wh, ctx->stream));
    unsigned long long segs = 0;
    HIPTRY(ctx, hipMemcpyAsync(&segs, q.seg_total, sizeof(segs), hipMemcpyDeviceToHost, ctx->stream));
    if (accum_rgb) HIPTRY(ctx, hipMemcpyAsync(accum_rgb, d_rgb, (size_t)wh * 3 * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    if (out) HIPTRY(ctx, hipMemcpyAsync(out, d_px, (size_t)wh * sizeof(ilm_pixel), hipMemcpyDeviceToHost, ctx->stream));
    HIPTRY(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.0f;
    HIPTRY(ctx, hipEventElapsedTime(&ms, ctx->na_ev[0], ctx->na_ev[1]));
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->segments = segs;
        info->device_ms = ms;
        info->chunks = chunks;
        info->paths_per_chunk = max_paths;
    }
    {
        std::lock_guard<std::mutex> lk(ctx->stat_mu);
        ctx->stat_busy_ms += ms;
        ctx->stat_segments += segs;
    }
    return ILM_OK;
}

/* ======================= parity hooks ======================= */

int ilm_debug_closest_hit(ilm_ctx* ctx, ilm_scene* scene, const float* rays, int32_t n, int32_t* inst,
                          int32_t* prim, float* t) {
    if (!ctx || !scene || !rays || n < 0 || !inst || !prim || !t) return fail(ctx, ILM_E_INVALID, "bad arguments");
    if (int rc = check_scene_stack(ctx, scene)) return rc;
    std::lock_guard<std::mutex> dev(ctx->dev_mu);
    HIPTRY(ctx, hipSetDevice(ctx->device));
    std::unique_ptr<DChunk> hc(new DChunk());
    std::memset(hc.get(), 0, sizeof(DChunk));
    hc->reqs[0].scene = scene->d_scene;
    std::vector<float4> o4, d4, h4;
    std::vector<int> hi;
    for (int64_t base = 0; base < n; base += ctx->cap) {
        int64_t m = std::min<int64_t>(ctx->cap, n - base);
        o4.resize(m);
        d4.resize(m);
        for (int64_t i = 0; i < m; i++) {
            const float* r = rays + 6 * (base + i);
            o4[i] = make_float4(r[0], r[1], r[2], 0.0f);
            d4[i] = make_float4(r[3], r[4], r[5], 0.0f);
        }
        hc->total_paths = (uint32_t)m;
        hc->depth_steps = 1;
        int rc = stage_chunk(ctx, *hc, ctx->d_chunk, ctx->stream);
        if (rc) return rc;
        HIPTRY(ctx, hipMemcpyAsync(ctx->A.o, o4.data(), m * sizeof(float4), hipMemcpyHostToDevice, ctx->stream));
        HIPTRY(ctx, hipMemcpyAsync(ctx->A.d, d4.data(), m * sizeof(float4), hipMemcpyHostToDevice, ctx->stream));
        KParams P = base_params(ctx);
        P.trav_stats = nullptr;
        P.use_ovf = scene_needs_ovf(ctx, scene) ? 1 : 0;
        HIPTRY(ctx, ilm_k_ctl_init(P, ctx->stream));
        HIPTRY(ctx, ilm_k_extend(P, 0, ctx->ovf, ctx->ext_grid, ctx->stream));
        h4.resize(m);
        hi.resize(m);
        HIPTRY(ctx, hipMemcpyAsync(h4.data(), ctx->hits.h, m * sizeof(float4), hipMemcpyDeviceToHost, ctx->stream));
        HIPTRY(ctx, hipMemcpyAsync(hi.data(), ctx->hits.inst, m * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIPTRY(ctx, hipStreamSynchronize(ctx->stream));
        for (int64_t i = 0; i < m; i++) {
            int32_t pr;
            std::memcpy(&pr, &h4[i].w, 4);
            inst[base + i] = hi[i];
            prim[base + i] = hi[i] >= 0 ? pr : -1;
            t[base + i] = hi[i] == -1 ? 3.402823466e+38f : h4[i].x;
        }
    }
    return ILM_OK;
}

int ilm_debug_path_radiance(ilm_ctx* ctx, ilm_scene* scene, const ilm_render_params* params,
                            const uint32_t* pixel_pass, int32_t n, float* rgb, int32_t* segments) {
    if (!ctx || !scene || !pixel_pass || n < 0 || !rgb) return fail(ctx, ILM_E_INVALID, "bad arguments");
    int rc = validate_params(ctx, params);
    if (!rc) rc = check_scene_stack(ctx, scene);
    if (rc) return rc;
    std::lock_guard<std::mutex> dev(ctx->dev_mu);
    HIPTRY(ctx, hipSetDevice(ctx->device));
    if (!ctx->seg_count) HIPTRY(ctx, hipMalloc((void**)&ctx->seg_count, ctx->cap * sizeof(int)));
    uint32_t* d_list = nullptr;
    HIPTRY(ctx, hipMalloc((void**)&d_list, 2 * std::min<int64_t>(ctx->cap, n) * sizeof(uint32_t) + 8));
    std::unique_ptr<DChunk> hc(new DChunk());
    std::memset(hc.get(), 0, sizeof(DChunk));
    DReq& R = hc->reqs[0];
    R.scene = scene->d_scene;
    R.width = params->width;
    R.height = params->height;
    R.max_depth = params->max_depth;
    R.ray_bias = params->ray_bias;
    R.tonemap = params->tonemap;
    R.spp = params->spp;
    hm::render_info(scene->camera, params->width, params->height, R.image_width, R.pixel_size, R.left, R.top);
    req_divs(R, 1);
    std::vector<float4> c4;
    std::vector<int> sc;
    for (int64_t base = 0; base < n; base += ctx->cap) {
        int64_t m = std::min<int64_t>(ctx->cap, n - base);
        hc->total_paths = (uint32_t)m;
        hc->depth_steps = params->max_depth;
        rc = stage_chunk(ctx, *hc, ctx->d_chunk, ctx->stream);
        if (rc) break;
        if (hipMemcpyAsync(d_list, pixel_pass + 2 * base, 2 * m * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream) != hipSuccess) {
            rc = fail(ctx, ILM_E_DEVICE, "list upload");
            break;
        }
        KParams P = base_params(ctx);
        P.trav_stats = nullptr;
        P.use_ovf = scene_needs_ovf(ctx, scene) ? 1 : 0;
        P.list = d_list;
        P.seg_count = ctx->seg_count;
        if (ilm_k_ctl_init(P, ctx->stream) != hipSuccess || start_paths(ctx, ctx_pipe(ctx), P, ctx->stream) != hipSuccess) {
            rc = fail(ctx, ILM_E_DEVICE, "raygen launch");
            break;
        }
        for (int d = 0; d < params->max_depth; d++) {
            P.in = (d & 1) ? ctx->B : ctx->A;
            P.out = (d & 1) ? ctx->A : ctx->B;
            if (ilm_k_extend(P, d, ctx->ovf, ctx->ext_grid, ctx->stream) != hipSuccess ||
                ilm_k_shade(P, d, ctx->shade_grid, ctx->stream) != hipSuccess) {
                rc = fail(ctx, ILM_E_DEVICE, "depth launch");
                break;
            }
        }
        if (rc) break;
        c4.resize(m);
        sc.resize(m);
        if (hipMemcpyAsync(c4.data(), ctx->contrib, m * sizeof(float4), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipMemcpyAsync(sc.data(), ctx->seg_count, m * sizeof(int), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess) {
            rc = fail(ctx, ILM_E_DEVICE, "readback");
            break;
        }
        for (int64_t i = 0; i < m; i++) {
            rgb[3 * (base + i)] = c4[i].x;
            rgb[3 * (base + i) + 1] = c4[i].y;
            rgb[3 * (base + i) + 2] = c4[i].z;
            if (segments) segments[base + i] = sc[i];
        }
    }
    dfree(ctx->device, d_list);
    return rc;
}

/* ======================= served form ======================= */

static void admit_select_locked(ilm_ctx* ctx, std::vector<Job*>& picked);
static void admit_prepare(ilm_ctx* ctx, std::vector<Job*>& picked);

static void latency_record_locked(ilm_ctx* ctx, double ms) {
    if (ctx->latencies.size() < ILM_LATENCY_WINDOW) {
        ctx->latencies.push_back(ms);
    } else {
        ctx->latencies[ctx->lat_next] = ms;
        ctx->lat_next = (ctx->lat_next + 1) % ILM_LATENCY_WINDOW;
    }
}

/* Ends a job the scheduler holds (finished, failed or cancelled with nothing in flight). Its
 * device buffers go back first, outside sched_mu (a hipFree, when the pool is full, waits for the
 * device); then, under sched_mu, it leaves the active set, its state changes and the scheduler's
 * reference drops. The caller holds no lock. */
static void sched_finish(ilm_ctx* ctx, Job* j, int state, double t) {
    job_release_device(j);
    std::lock_guard<std::mutex> lk(ctx->sched_mu);
    ctx->in_use -= j->bytes;
    auto it = std::find(ctx->active.begin(), ctx->active.end(), j);
    if (it != ctx->active.end()) ctx->active.erase(it);
    j->end_ms = t;
    if (state == ILM_REQ_CANCELLED) {
        ctx->stat_cancelled++;
    } else {
        latency_record_locked(ctx, t - j->submit_ms);
        ctx->stat_done++;
        j->samples_done = j->p.spp;
    }
    j->state.store(state);
    j->cv.notify_all();
    if (--j->refs == 0) { scene_release(j->scene, true); delete j; }
}

static void server_loop(ilm_ctx* ctx) {
    (void)hipSetDevice(ctx->device);
    std::unique_ptr<DChunk> hc(new DChunk());
    std::vector<Job*> in_chunk;
    /* a chunk in flight: its completion event, the jobs it finishes and the pinned slot its
     * segment count is copied into */
    struct Inflight { hipEvent_t ev; std::vector<Job*> finishing; int seg_slot; };
    std::deque<Inflight> inflight;
    for (;;) {
        std::vector<Job*> active, picked;
        {
            std::unique_lock<std::mutex> lk(ctx->sched_mu);
            ctx->sched_cv.wait(lk, [&] {
                return ctx->stop || !ctx->pending.empty() || !ctx->active.empty() || !inflight.empty();
            });
            if (ctx->stop && ctx->pending.empty() && ctx->active.empty() && inflight.empty()) break;
            admit_select_locked(ctx, picked);
        }
        /* device buffers of newly admitted jobs: dev_mu only, never inside sched_mu */
        admit_prepare(ctx, picked);
        {
            std::lock_guard<std::mutex> lk(ctx->sched_mu);
            active = ctx->active;
        }
        bool launched = false;
        {
            std::lock_guard<std::mutex> dev(ctx->dev_mu);
            std::vector<Job*> issuable;
            for (Job* j : active)
                if (job_has_work(j) && !j->cancel.load()) issuable.push_back(j);
            if (!issuable.empty() && inflight.size() < 2) {
                build_chunk(ctx, issuable, *hc, in_chunk);
                for (int w = 0; w < hc->n_work; w++) hc->work[w].pass0 += in_chunk[hc->work[w].req]->p.first_sample;
                int pk = 0;
                int rc = launch_chunk(ctx, *hc, jobs_need_ovf(ctx, in_chunk), nullptr, &pk);
                const Pipe q = ctx_pipe(ctx, pk);
                Inflight f;
                f.seg_slot = ctx->seg_ring_pos++ % ILM_RING;   /* at most two chunks in flight */
                ctx->h_seg_ring[f.seg_slot] = 0;
                (void)hipMemcpyAsync(ctx->h_seg_ring + f.seg_slot, &q.ctl->segments, sizeof(unsigned long long),
                                     hipMemcpyDeviceToHost, q.stream);
                (void)hipEventCreateWithFlags(&f.ev, hipEventDisableTiming);
                for (Job* j : in_chunk) {
                    if (!job_has_work(j)) {
                        const int64_t wh = (int64_t)j->p.width * j->p.height;
                        if (ilm_k_tonemap(j->d_accum, j->d_px, (int)wh, j->p.spp, j->p.tonemap, q.stream) != hipSuccess)
                            j->failed = true;
                        (void)hipMemcpyAsync(j->h_accum, j->d_accum, wh * sizeof(float4), hipMemcpyDeviceToHost, q.stream);
                        (void)hipMemcpyAsync(j->h_px, j->d_px, wh * sizeof(ilm_pixel), hipMemcpyDeviceToHost, q.stream);
                        f.finishing.push_back(j);
                    }
                }
                (void)hipEventRecord(f.ev, q.stream);
                for (Job* j : in_chunk) j->last_ev_seq = ctx->chunk_seq;
                ctx->chunk_seq++;
                if (rc) {
                    for (Job* j : in_chunk) j->failed = true;
                }
                inflight.push_back(f);
                launched = true;
            }
        }
        /* retire finished chunks; block on the oldest when nothing new could be launched */
        while (!inflight.empty()) {
            Inflight& f = inflight.front();
            hipError_t q = launched ? hipEventQuery(f.ev) : hipEventSynchronize(f.ev);
            if (q == hipErrorNotReady) break;
            double t = now_ms(ctx);
            std::vector<Job*> done = f.finishing;
            {
                std::lock_guard<std::mutex> lk(ctx->stat_mu);
                ctx->stat_segments += ctx->h_seg_ring[f.seg_slot];
            }
            (void)hipEventDestroy(f.ev);
            inflight.pop_front();
            ctx->chunks_retired++;
            for (Job* j : done) sched_finish(ctx, j, j->failed ? ILM_REQ_FAILED : ILM_REQ_DONE, t);
            launched = false;
        }
        /* cancelled jobs: release once no chunk that used them is still in flight (stream order
         * keeps any later user of their pooled buffers behind the kernels already queued) */
        std::vector<Job*> cancelled;
        {
            std::lock_guard<std::mutex> lk(ctx->sched_mu);
            for (Job* j : ctx->active) {
                bool busy = j->last_ev_seq >= ctx->chunk_seq - (int64_t)inflight.size();
                if (j->cancel.load() && !busy) cancelled.push_back(j);
            }
        }
        for (Job* j : cancelled) sched_finish(ctx, j, ILM_REQ_CANCELLED, now_ms(ctx));
    }
}

/* Take queued jobs out of the queue while the HBM budget allows, reserving their bytes; the
 * caller holds sched_mu. Cancelled queued jobs end here. */
static void admit_select_locked(ilm_ctx* ctx, std::vector<Job*>& picked) {
    while (!ctx->pending.empty() && (int)(ctx->active.size() + picked.size()) < ILM_MAX_JOBS) {
        Job* j = ctx->pending.front();
        if (j->cancel.load()) {
            ctx->pending.pop_front();
            j->state.store(ILM_REQ_CANCELLED);
            j->end_ms = now_ms(ctx);
            ctx->stat_cancelled++;
            j->cv.notify_all();
            if (--j->refs == 0) { scene_release(j->scene, true); delete j; }
            continue;
        }
        if (ctx->in_use + j->bytes > ctx->budget && !(ctx->active.empty() && picked.empty())) break;
        ctx->pending.pop_front();
        ctx->in_use += j->bytes;
        picked.push_back(j);
    }
}

/* Allocate and clear the picked jobs' device buffers (under dev_mu), then make them active
 * (under sched_mu); the two locks are never held together. */
static void admit_prepare(ilm_ctx* ctx, std::vector<Job*>& picked) {
    for (Job* j : picked) {
        int rc;
        {
            std::lock_guard<std::mutex> dev(ctx->dev_mu);
            rc = job_prepare(ctx, j);
            if (rc != ILM_OK) job_release_device(j);
        }
        std::lock_guard<std::mutex> lk(ctx->sched_mu);
        if (rc != ILM_OK) {
            ctx->in_use -= j->bytes;
            j->end_ms = now_ms(ctx);
            j->state.store(ILM_REQ_FAILED);
            j->cv.notify_all();
            if (--j->refs == 0) { scene_release(j->scene, true); delete j; }
            continue;
        }
        if (ctx->stop) j->cancel.store(true);
        j->start_ms = now_ms(ctx);
        j->state.store(ILM_REQ_RUNNING);
        ctx->active.push_back(j);
    }
    picked.clear();
}

static void server_start(ilm_ctx* ctx) {
    if (ctx->server.joinable()) return;
    ctx->stop = false;
    ctx->server = std::thread(server_loop, ctx);
}

static void server_stop(ilm_ctx* ctx) {
    {
        std::lock_guard<std::mutex> lk(ctx->sched_mu);
        ctx->stop = true;
        for (Job* j : ctx->pending) j->cancel.store(true);
        for (Job* j : ctx->active) j->cancel.store(true);
    }
    ctx->sched_cv.notify_all();
    if (ctx->server.joinable()) ctx->server.join();
}

int ilm_submit(ilm_ctx* ctx, ilm_scene* scene, const ilm_render_params* params, ilm_req** out) {
    if (!ctx || !scene || !out) return fail(ctx, ILM_E_INVALID, "null argument");
    *out = nullptr;
    int rc = validate_params(ctx, params);
    if (!rc) rc = check_scene_stack(ctx, scene);
    if (rc) return rc;
    const int64_t wh = (int64_t)params->width * params->height;
    Job* j = new Job();
    j->scene = scene;
    j->p = *params;
    j->bytes = wh * (2 * (int64_t)sizeof(float4) + (int64_t)sizeof(ilm_pixel));
    j->device = ctx->device;
    bool pinned;
    {
        std::lock_guard<std::mutex> cap(capture_mu(ctx->device));
        pinned = hipHostMalloc((void**)&j->h_accum, wh * sizeof(float4), hipHostMallocDefault) == hipSuccess &&
                 hipHostMalloc((void**)&j->h_px, wh * sizeof(ilm_pixel), hipHostMallocDefault) == hipSuccess;
    }
    if (!pinned) {
        delete j;   /* ~Job frees whichever buffer was allocated */
        return fail(ctx, ILM_E_NOMEM, "pinned result buffers");
    }
    {
        std::lock_guard<std::mutex> lk(ctx->sched_mu);
        int maxq = ctx->opts.max_queued > 0 ? ctx->opts.max_queued : 4096;
        if ((int)ctx->pending.size() >= maxq) {
            ctx->stat_rejected++;
            delete j;
            return fail(ctx, ILM_E_BUSY, "admission queue full");
        }
        if (j->bytes > ctx->budget) {
            ctx->stat_rejected++;
            delete j;
            return fail(ctx, ILM_E_NOMEM, "request exceeds the HBM budget");
        }
        scene->refs.fetch_add(1);
        j->refs = 2;   /* scheduler + caller handle */
        j->submit_ms = now_ms(ctx);
        j->state.store(ILM_REQ_QUEUED);
        ctx->pending.push_back(j);
        server_start(ctx);
    }
    ctx->sched_cv.notify_all();
    ilm_req* r = new ilm_req();
    r->job = j;
    r->ctx = ctx;
    ctx->refs.fetch_add(1);
    *out = r;
    return ILM_OK;
}

int ilm_poll(ilm_req* r, ilm_req_status* st) {
    if (!r || !st) return fail(nullptr, ILM_E_INVALID, "null argument");
    Job* j = r->job;
    std::lock_guard<std::mutex> lk(r->ctx->sched_mu);
    st->state = j->state.load();
    st->samples_done = j->state.load() == ILM_REQ_DONE ? j->p.spp : j->next_pass;
    st->segments = 0;
    st->submit_ms = j->submit_ms;
    st->start_ms = j->start_ms;
    st->end_ms = j->end_ms;
    return ILM_OK;
}

int ilm_wait(ilm_req* r, int32_t timeout_ms) {
    if (!r) return fail(nullptr, ILM_E_INVALID, "null request");
    Job* j = r->job;
    std::unique_lock<std::mutex> lk(r->ctx->sched_mu);
    auto done = [&] { int s = j->state.load(); return s == ILM_REQ_DONE || s == ILM_REQ_CANCELLED || s == ILM_REQ_FAILED; };
    if (timeout_ms < 0) j->cv.wait(lk, done);
    else if (!j->cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), done)) return fail(r->ctx, ILM_E_TIMEOUT, "wait timed out");
    int s = j->state.load();
    if (s == ILM_REQ_CANCELLED) return fail(r->ctx, ILM_E_CANCELLED, "request cancelled");
    if (s == ILM_REQ_FAILED) return fail(r->ctx, ILM_E_DEVICE, "request failed");
    return ILM_OK;
}

int ilm_result(ilm_req* r, float* accum_rgb, ilm_pixel* out) {
    if (!r) return fail(nullptr, ILM_E_INVALID, "null request");
    Job* j = r->job;
    if (j->state.load() != ILM_REQ_DONE) return fail(r->ctx, ILM_E_STATE, "request not done");
    const int64_t wh = (int64_t)j->p.width * j->p.height;
    if (accum_rgb)
        for (int64_t i = 0; i < wh; i++) {
            accum_rgb[3 * i] = j->h_accum[i].x;
            accum_rgb[3 * i + 1] = j->h_accum[i].y;
            accum_rgb[3 * i + 2] = j->h_accum[i].z;
        }
    if (out) std::memcpy(out, j->h_px, wh * sizeof(ilm_pixel));
    return ILM_OK;
}

int ilm_cancel(ilm_req* r) {
    if (!r) return fail(nullptr, ILM_E_INVALID, "null request");
    r->job->cancel.store(true);
    r->ctx->sched_cv.notify_all();
    return ILM_OK;
}

void ilm_req_free(ilm_req* r) {
    if (!r) return;
    Job* j = r->job;
    ilm_ctx* ctx = r->ctx;
    bool del = false;
    {
        std::lock_guard<std::mutex> lk(ctx->sched_mu);
        j->cancel.store(true);
        del = (--j->refs == 0);
    }
    ctx->sched_cv.notify_all();
    if (del) {
        scene_release(j->scene);
        delete j;
    }
    delete r;
    ctx_unref(ctx);
}

Job::~Job() {
    if (!h_accum && !h_px) return;
    std::lock_guard<std::mutex> lk(capture_mu(device));
    if (h_accum) (void)hipHostFree(h_accum);
    if (h_px) (void)hipHostFree(h_px);
}

/* Reads host-side counters only (no device call, so a poll never waits behind a chunk in
 * flight): segments as of the last completed chunk or render, p50/p99 over the last
 * ILM_LATENCY_WINDOW completed requests. */
int ilm_stats_get(ilm_ctx* ctx, ilm_stats* st) {
    if (!ctx || !st) return fail(ctx, ILM_E_INVALID, "null argument");
    double lat[ILM_LATENCY_WINDOW];
    size_t n = 0;
    std::lock_guard<std::mutex> lk(ctx->sched_mu);
    std::memset(st, 0, sizeof(*st));
    st->requests_done = ctx->stat_done;
    st->requests_rejected = ctx->stat_rejected;
    st->requests_cancelled = ctx->stat_cancelled;
    {
        std::lock_guard<std::mutex> sl(ctx->stat_mu);
        st->segments = ctx->stat_segments;
        st->busy_ms = ctx->stat_busy_ms;
        st->bvh_cache_hits = ctx->stat_bvh_hits;
        st->bvh_cache_misses = ctx->stat_bvh_misses;
    }
    n = ctx->latencies.size();
    std::copy(ctx->latencies.begin(), ctx->latencies.end(), lat);
    st->hbm_budget_bytes = ctx->budget;
    st->hbm_in_use_bytes = ctx->in_use;
    if (n) {
        std::sort(lat, lat + n);
        auto q = [&](double f) {
            double pos = f * (double)(n - 1);
            size_t i = (size_t)pos;
            double fr = pos - (double)i;
            return i + 1 < n ? lat[i] * (1 - fr) + lat[i + 1] * fr : lat[i];
        };
        st->p50_ms = q(0.50);
        st->p99_ms = q(0.99);
    }
    return ILM_OK;
}
