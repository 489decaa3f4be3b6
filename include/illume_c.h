/*
 * illume_c.h - C ABI of the MI355X path-tracer hot path.
 *
 * Conventions
 *   - Every int-returning entry point returns ILM_OK (0) or a negative ILM_E_* code.
 *     Nothing calls exit(); the text of the last error is ilm_last_error(ctx)
 *     (or ilm_global_error() for calls without a context).
 *   - The caller owns every host buffer it passes in or out; the library owns device memory.
 *   - Descriptors are read during the call only; the library copies what it keeps.
 *   - ilm_scene handles stay device-resident until ilm_scene_free; requests in flight
 *     keep their scene alive (reference counted).
 *   - ilm_submit / ilm_poll / ilm_wait / ilm_cancel are thread-safe. One scheduler thread
 *     per context drives the device.
 *
 * Reference interface each group replaces (matthewshang/illume):
 *   ilm_scene_upload / ilm_scene_free    SceneRef(HostScene&) / ~SceneRef   src/scene/sceneref.cpp:9-148
 *   ilm_desc_load_json                   HostScene(rapidjson::Document&)    src/scene/scene.cpp:11-99
 *                                        Renderer ctor render_settings       src/renderer.cu:39-59
 *   ilm_render                           Renderer::render_to_bitmap(Bitmap*) src/renderer.cu:498-597
 *   ilm_pixel                            Pixel {size_t r,g,b}                src/bitmap.h:6-12
 *   ilm_bvh_*                            bvh_create / GPUNode / BVH          src/accel/bvh.h:11-42
 *   ilm_debug_closest_hit                get_min_hit                         src/renderer.cu:112-137
 *   ilm_debug_path_radiance              init_rays + pathtrace_kernel        src/renderer.cu:81-110,139-388
 *   ilm_submit .. ilm_stats, ilm_render_batch   no reference counterpart (served and batched forms)
 *   ilm_render_noalloc                   no reference counterpart (ilm_render for callers that
 *                                        cannot allocate device memory)
 */
#ifndef ILLUME_C_H
#define ILLUME_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ILM_ABI_VERSION 1

enum {
    ILM_OK = 0,
    ILM_E_INVALID = -1,     /* bad argument or descriptor */
    ILM_E_NOMEM = -2,       /* host or device allocation failed, or over the HBM budget */
    ILM_E_DEVICE = -3,      /* HIP runtime error / no usable GPU */
    ILM_E_IO = -4,          /* file could not be read */
    ILM_E_PARSE = -5,       /* malformed scene JSON or OBJ */
    ILM_E_BUSY = -6,        /* admission queue full */
    ILM_E_CANCELLED = -7,   /* request was cancelled */
    ILM_E_TIMEOUT = -8,     /* ilm_wait timed out */
    ILM_E_UNSUPPORTED = -9,
    ILM_E_STATE = -10       /* call not valid in the handle's current state */
};

/* Texture types (src/texture.h:8-11) */
enum { ILM_TEX_CONSTANT = 0, ILM_TEX_CHECKERBOARD = 1, ILM_TEX_BITMAP = 2 };

/* Material types, same numbering as src/material.h:12-20 */
enum {
    ILM_MAT_EMISSIVE = 0,
    ILM_MAT_DIFFUSE = 1,
    ILM_MAT_SPECULAR = 2,
    ILM_MAT_REFRACTIVE = 3,
    ILM_MAT_ROUGHREFLECTIVE = 4,
    ILM_MAT_ROUGHREFRACTIVE = 5,
    ILM_MAT_CONDUCTOR = 6,
    ILM_MAT_ROUGHCONDUCTOR = 7,
    ILM_MAT_PLASTIC = 8
};

/* Microfacet distribution. Beckmann is the reference's only model (src/microfacet.cu:7-52);
 * GGX is an additive extension selected by the "distribution":"ggx" material key. */
enum { ILM_DIST_BECKMANN = 0, ILM_DIST_GGX = 1 };

/* Tonemap operators (src/tonemapper.h:11-14) */
enum { ILM_TONEMAP_LINEAR = 0, ILM_TONEMAP_REINHARD = 1, ILM_TONEMAP_FILMIC = 2, ILM_TONEMAP_UNCHARTED = 3 };

/* Node layouts of the device BVH */
enum { ILM_BVH_FP32 = 0, ILM_BVH_QUANT8 = 1 };

/* Ray/triangle test: Moller-Trumbore exactly as src/primitives/mesh.cu:446-476, or watertight
 * (Woop, Benthin, Wald, JCGT 2013). */
enum { ILM_TRI_MOLLER_TRUMBORE = 0, ILM_TRI_WATERTIGHT = 1 };

typedef struct ilm_texture_desc {
    int32_t type;               /* ILM_TEX_* */
    float color[3];             /* constant */
    float on[3];                /* checkerboard on colour */
    float off[3];               /* checkerboard off colour */
    float scale[2];             /* checkerboard and bitmap uv scale (default 1,1) */
    int32_t width, height;      /* bitmap */
    int32_t is_hdr;             /* 1: float RGBA texels, bilinear; 0: 8-bit RGBA texels, point */
    const float* texels_f32;    /* width*height*4 when is_hdr */
    const uint8_t* texels_u8;   /* width*height*4 when !is_hdr */
} ilm_texture_desc;

typedef struct ilm_medium_desc {
    float sigma_a[3];           /* absorption coefficient = 1 - JSON "absorption" (src/medium.cu:9) */
    float sigma_s;              /* scattering coefficient */
    float g;                    /* Henyey-Greenstein asymmetry */
} ilm_medium_desc;

typedef struct ilm_material_desc {
    int32_t type;               /* ILM_MAT_* */
    ilm_texture_desc albedo;    /* default constant 0 */
    ilm_texture_desc roughness; /* default constant 1e-4 (src/material.cu:33) */
    float ior;
    int32_t medium;             /* index into the scene's mediums, -1 = air */
    float eta[3], k[3];         /* conductor complex IOR */
    int32_t distribution;       /* ILM_DIST_* (rough materials) */
} ilm_material_desc;

typedef struct ilm_mesh_desc {
    int32_t num_vertices;
    const float* positions;     /* num_vertices*3 */
    const float* normals;       /* num_vertices*3, required when !face_normals */
    const float* texcoords;     /* num_vertices*2, required when has_texcoords */
    int32_t num_triangles;
    const int32_t* indices;     /* num_triangles*3 */
    int32_t face_normals;       /* reference default 1 (src/primitives/mesh.cu:440) */
    int32_t has_texcoords;
    int32_t leaf_size;          /* reference "bvh_items_per_node" (default 4) */
} ilm_mesh_desc;

/* Object-to-world transform. With has_matrix=0 the library builds
 * translate * (Rx*Ry*Rz) * scale from degrees (src/math/transform.cu:28-44), or a look-at
 * frame when has_lookat=1; inverse and inverse-transpose follow src/math/transform.cu:6-16. */
typedef struct ilm_transform_desc {
    float translation[3];
    float scale[3];
    float rotation[3];          /* degrees */
    int32_t has_lookat;
    float lookat[3];
    int32_t has_matrix;
    float matrix[16];           /* row-major */
} ilm_transform_desc;

typedef struct ilm_instance_desc {
    int32_t mesh;
    int32_t material;
    ilm_transform_desc transform;
} ilm_instance_desc;

typedef struct ilm_sphere_desc {
    float center[3];
    float radius;
    int32_t material;
} ilm_sphere_desc;

typedef struct ilm_camera_desc {
    float fov;                  /* horizontal, degrees */
    float dof;                  /* image-plane distance ("depth_of_field") */
    float aperture;             /* thin-lens radius; 0 = pinhole */
    ilm_transform_desc transform;
} ilm_camera_desc;

typedef struct ilm_scene_desc {
    int32_t num_meshes;        const ilm_mesh_desc* meshes;
    int32_t num_instances;     const ilm_instance_desc* instances;
    int32_t num_spheres;       const ilm_sphere_desc* spheres;
    int32_t num_materials;     const ilm_material_desc* materials;
    int32_t num_mediums;       const ilm_medium_desc* mediums;
    int32_t has_environment;   ilm_texture_desc environment;
    ilm_camera_desc camera;
} ilm_scene_desc;

typedef struct ilm_render_params {
    int32_t width, height;
    int32_t spp;
    int32_t max_depth;
    float ray_bias;
    int32_t tonemap;            /* ILM_TONEMAP_* */
    int32_t first_sample;       /* pass index of the first sample; pass i seeds wang_hash(i) */
} ilm_render_params;

/* Same layout as the reference's Pixel {size_t red, green, blue} on LP64 */
typedef struct ilm_pixel {
    uint64_t r, g, b;
} ilm_pixel;

/* Extend-kernel scheduling (results are identical for every mode). BATCH: each wave traces 64
 * rays to completion; LEAF_REFILL: idle lanes take new rays when the wave reaches BLAS leaves;
 * STEP_REFILL: idle lanes take new rays after every traversal step (the default). */
enum { ILM_EXTEND_DEFAULT = 0, ILM_EXTEND_BATCH = 1, ILM_EXTEND_LEAF_REFILL = 2, ILM_EXTEND_STEP_REFILL = 3 };

typedef struct ilm_opts {
    int32_t bvh_layout;         /* ILM_BVH_* (default FP32) */
    int32_t tri_test;           /* ILM_TRI_* (default Moller-Trumbore) */
    int64_t paths_in_flight;    /* 0 = automatic */
    int64_t hbm_budget_bytes;   /* serving admission budget; 0 = 90% of free device memory */
    int32_t use_graphs;         /* capture the per-chunk pipeline as a hipGraph */
    int32_t max_queued;         /* serving: admission queue bound (0 = 4096) */
    int32_t extend_mode;        /* ILM_EXTEND_*: how extend waves schedule their rays */
    int32_t profile;            /* time every extend launch with HIP events (disables graphs) */
    int32_t count_tests;        /* count box/triangle tests of the extend kernel (slower) */
    int32_t verbose;
    /* Path order inside a chunk (results are identical for every choice; only coherence
     * changes): sample_group consecutive path ids take consecutive samples of one pixel (a
     * power of two; 1 = pixel-major; 0 = 64), tile_rows > 1 walks the image in bands of that many
     * rows, column by column, so a wave covers a compact block of pixels (1 = row order;
     * 0 = 4). */
    int32_t sample_group;
    int32_t tile_rows;
    /* at most this many passes of one request per chunk (0 = as many as paths_in_flight
     * holds); 1 gives the reference's loop shape, one pass after another (src/renderer.cu:531) */
    int32_t chunk_passes;
    /* directory of the per-mesh BVH cache (versioned, hashed files "<key>.ilmbvh" holding both
     * device layouts; replaces the reference's unversioned .bvh files, src/accel/bvh.cpp:399-439),
     * or NULL for none. The string is copied at ilm_open. */
    const char* bvh_cache_dir;
    /* refill modes: number of idle lanes of a wave that triggers a refill (0 = 24) */
    int32_t refill_min;
    /* rays an extend wave reserves from the queue at a time (multiple of 64; 0 = 256): its first
     * reservation by its index, later ones from the shared cursor. In the per-step and leaf
     * refill modes a queue shorter than the grid's first reservations is split evenly over the
     * waves instead; the batch mode and the first depth's camera packets shrink their later
     * reservations as the queue drains (down to 64 rays) */
    int32_t ext_fetch;
    /* per-step refill: a wave runs its BLAS-leaf triangle tests only on trips where the lanes
     * waiting at leaves are at least this percentage of its other busy lanes (0 = 60) */
    int32_t leaf_wait_pct;
    /* shade the paths of each 512-path block grouped by medium / miss / sphere / instance
     * (0 = default: on, 1 = on, 2 = off); the first depth, whose paths are in pixel-sample
     * order, is shaded in that order either way */
    int32_t shade_sort;
    /* chunk packing across requests (served and batched forms): 0 = in arrival order, each
     * request taking all it can; 1 = fair share (requests needing less than an equal share get
     * it all, the rest split the chunk), so short requests finish without waiting behind long
     * ones. Results are identical either way. */
    int32_t sched_fair;
    /* first depth with per-step refill: trace each wave's 64 camera rays as one packet (one
     * traversal for the wave, every lane testing its own ray; 0 = default: on, 1 = on, 2 = off,
     * each lane walking alone). Results are identical either way. */
    int32_t camera_packets;
    /* chunk pipelines (1 or 2; 0 = 2): consecutive chunks alternate between two sets of path
     * buffers on two streams, so a chunk's last, short launches run beside the next chunk's
     * first ones (paths_in_flight paths each). Results are identical either way. */
    int32_t pipes;
    /* a lone request (plain form) of at least 2^22 paths that fits one chunk, with two
     * pipelines: 0 = default (split when its max_depth is at least 24), 1 = never, 2 = always
     * split into two chunks of half its pixels (all passes each) that run side by side, so
     * each chunk's latency-bound last launches overlap the other's. Results are identical. */
    int32_t lone_split;
} ilm_opts;

typedef struct ilm_render_info {
    uint64_t segments;          /* closest-hit queries, same count as src/renderer.cu:554 */
    double device_ms;           /* device time of the render loop (HIP events) */
    double extend_ms;           /* device time inside the extend kernel (ilm_opts.profile) */
    int32_t extend_launches;    /* extend launches timed (ilm_opts.profile) */
    int32_t chunks;             /* chunks the render was split into */
    int64_t paths_per_chunk;
    uint64_t box_tests;         /* child-box tests (ilm_opts.count_tests) */
    uint64_t tri_tests;         /* triangle tests (ilm_opts.count_tests) */
} ilm_render_info;

typedef struct ilm_stats {
    uint64_t segments;
    uint64_t requests_done;
    uint64_t requests_rejected;
    uint64_t requests_cancelled;
    double busy_ms;
    double p50_ms, p99_ms;      /* request latency: completion - submission */
    int64_t hbm_budget_bytes;
    int64_t hbm_in_use_bytes;
    uint64_t bvh_cache_hits;    /* meshes whose BVH came from the cache */
    uint64_t bvh_cache_misses;  /* meshes built (and stored) with a cache directory set */
} ilm_stats;

enum { ILM_REQ_QUEUED = 0, ILM_REQ_RUNNING = 1, ILM_REQ_DONE = 2, ILM_REQ_CANCELLED = 3, ILM_REQ_FAILED = 4 };

typedef struct ilm_req_status {
    int32_t state;              /* ILM_REQ_* */
    int32_t samples_done;
    uint64_t segments;
    double submit_ms, start_ms, end_ms;   /* steady clock, ms since the context opened */
} ilm_req_status;

typedef struct ilm_bvh_info {
    int32_t num_nodes;          /* inner + leaf nodes */
    int32_t num_leaves;
    int32_t max_depth;
    int32_t max_leaf_size;
    double sah_cost;            /* sum over nodes of SA(node)/SA(root) * (inner: 1, leaf: count) */
} ilm_bvh_info;

typedef struct ilm_ctx ilm_ctx;
typedef struct ilm_scene ilm_scene;
typedef struct ilm_req ilm_req;
typedef struct ilm_desc ilm_desc;
typedef struct ilm_bvh ilm_bvh;

int ilm_version(void);
const char* ilm_global_error(void);

/* ---- context ---- */
int ilm_open(int device, const ilm_opts* opts, ilm_ctx** out);
void ilm_close(ilm_ctx* ctx);
const char* ilm_last_error(ilm_ctx* ctx);
int ilm_stats_get(ilm_ctx* ctx, ilm_stats* out);
/* Retune an idle context (no counterpart in the reference, whose launch shape is fixed):
 * bvh_layout, extend_mode, refill_min, ext_fetch, leaf_wait_pct, shade_sort, sched_fair,
 * camera_packets, sample_group, tile_rows, chunk_passes, paths_in_flight and lone_split
 * take the values in *opts; other fields are ignored. ILM_E_BUSY while served requests are queued or running.
 * The autotuner (illume_amd/autotune.py) drives this between short trial renders. */
int ilm_tune(ilm_ctx* ctx, const ilm_opts* opts);

/* ---- scenes ---- */
int ilm_scene_upload(ilm_ctx* ctx, const ilm_scene_desc* desc, ilm_scene** out);
void ilm_scene_free(ilm_scene* scene);
int64_t ilm_scene_device_bytes(const ilm_scene* scene);

/* Reference-schema scene JSON (+ OBJ meshes) to a host descriptor; host only. Replaces
 * HostScene's JSON constructor (src/scene/scene.cpp:11-99). Bitmap textures and environment maps
 * are read from PNG or binary PPM/PGM files (8-bit texels) and, for paths ending in "hdr", from
 * Radiance RGBE files (float texels), as the reference's stb_image path gives them
 * (src/imageio.cpp:6-21, src/texture.cu:49-63); PFM files give float texels as well. A texture
 * file that cannot be read becomes black, with a message on stderr; an OBJ file that cannot be
 * read fails the load with ILM_E_IO. */
int ilm_desc_load_json(const char* path, ilm_desc** out);
const ilm_scene_desc* ilm_desc_scene(const ilm_desc* d);
int ilm_desc_settings(const ilm_desc* d, ilm_render_params* out);
void ilm_desc_free(ilm_desc* d);
/* Wall time the load spent reading OBJ files (parse) and merging their vertices, summed over
 * the scene's meshes; the reference's reader is src/primitives/mesh.cu:159-340. */
int ilm_desc_load_times(const ilm_desc* d, double* parse_s, double* merge_s);
const char* ilm_loader_error(void);

/* ---- plain render (one request, synchronous) ----
 * accum_rgb: W*H*3 floats, Kahan-compensated sum of the spp sample radiances (nullable).
 * out: W*H pixels after tonemapping, as src/renderer.cu:426-441 (nullable). */
int ilm_render(ilm_ctx* ctx, ilm_scene* scene, const ilm_render_params* params,
               float* accum_rgb, ilm_pixel* out, ilm_render_info* info);

/* ---- plain render without allocation ----
 * No counterpart in the reference. For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the same render and
 * outputs as ilm_render (bit-identical accum_rgb and pixels, same segment count), but every
 * device buffer of the call - path state, hits, per-path results, control block, chunk
 * descriptor, traversal-stack spill area, Kahan sums and pixels - is carved from one static
 * device array of ILM_NOALLOC_SCRATCH_BYTES per GPU, which each GPU the library runs on holds
 * until the process ends. The call makes no hipMalloc/hipFree and no host allocation, launches
 * kernels directly (no graph) and synchronises before it returns.
 * Limit: W*H*56 bytes of pixel buffers plus at least 4096 paths (148 bytes each), and for a
 * scene whose BVH needs the spilling traversal stack a spill area of up to 28 MB, must fit the
 * array (up to about 2.38 M pixels, 1.87 M with the spill area); larger images return
 * ILM_E_UNSUPPORTED. The array keeps the last call's contents (each call clears what it needs);
 * calls on one GPU share it, so they would race: the library serialises them with a per-device
 * lock held until the call returns. Paths per chunk are what the array holds (at most the
 * context's paths_in_flight), so large images run as more chunks than under ilm_render.
 * info: segments, device_ms, chunks, paths_per_chunk (test counts and extend timing are 0). */
#define ILM_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int ilm_render_noalloc(ilm_ctx* ctx, ilm_scene* scene, const ilm_render_params* params,
                       float* accum_rgb, ilm_pixel* out, ilm_render_info* info);

/* ---- batched render: n requests' paths share every launch of one pipeline ---- */
int ilm_render_batch(ilm_ctx* ctx, int32_t n, ilm_scene* const* scenes, const ilm_render_params* params,
                     float* const* accum_rgb, ilm_pixel* const* out, ilm_render_info* info);

/* ---- served form: continuous batching ---- */
int ilm_submit(ilm_ctx* ctx, ilm_scene* scene, const ilm_render_params* params, ilm_req** out);
int ilm_poll(ilm_req* req, ilm_req_status* out);
int ilm_wait(ilm_req* req, int32_t timeout_ms);
int ilm_result(ilm_req* req, float* accum_rgb, ilm_pixel* out);
int ilm_cancel(ilm_req* req);
void ilm_req_free(ilm_req* req);

/* ---- BVH build (host only; the builder used by ilm_scene_upload) ---- */
int ilm_bvh_build(const ilm_mesh_desc* mesh, ilm_bvh** out);
int ilm_bvh_get_info(const ilm_bvh* b, ilm_bvh_info* out);
/* boxes: num_nodes*6 (lo.xyz, hi.xyz); children: num_nodes*2 (inner: child node ids;
 * leaf: -1, -1); leaf_first/leaf_count: num_nodes (first slot into tri_order, count; 0 for inner);
 * tri_order: num_triangles */
int ilm_bvh_export(const ilm_bvh* b, float* boxes, int32_t* children, int32_t* leaf_first,
                   int32_t* leaf_count, int32_t* tri_order);
/* 8-bit quantised BVH4 layout built from the same tree: node count, root child code
 * (>= 0 node, < 0 leaf code) and depth; per node and child k < 4: the dequantised box as the
 * kernel computes it (qboxes, 24 floats per node), the exact FP32 box it encodes (exact) and
 * the child code (child: >= 0 node, leaf code with bit 31 set, 0x7fffffff unused). */
int ilm_bvh_quantized_size(const ilm_bvh* b, int32_t* num_qnodes, int32_t* root, int32_t* max_depth);
int ilm_bvh_export_quantized(const ilm_bvh* b, float* qboxes, float* exact, int32_t* child);
void ilm_bvh_free(ilm_bvh* b);
/* Device layouts (FP32 BVH2 nodes, 8-bit BVH4 nodes, leaf triangle order) exactly as
 * ilm_scene_upload builds them, through the BVH cache in cache_dir (NULL: none). Replaces the
 * reference's .bvh cache read/write (src/accel/bvh.cpp:399-439, src/primitives/mesh.cu:407-414).
 * *nodes_bytes / *qnodes_bytes: capacity in, size out; NULL buffers only query.
 * info4: root, qroot, depth, qdepth. *from_cache = 1 when the cache supplied them. */
int ilm_bvh_device_layouts(const ilm_mesh_desc* m, const char* cache_dir, int32_t* from_cache, void* nodes,
                           int64_t* nodes_bytes, void* qnodes, int64_t* qnodes_bytes, int32_t* order,
                           int32_t* info4);

/* ---- parity hooks ---- */
/* rays: n*6 floats (o.xyz, d.xyz) in world space. inst: instance index, or -2-s for sphere s,
 * or -1 for a miss; prim: triangle index within the instance's mesh (-1 otherwise); t: hit distance. */
int ilm_debug_closest_hit(ilm_ctx* ctx, ilm_scene* scene, const float* rays, int32_t n,
                          int32_t* inst, int32_t* prim, float* t);
/* Radiance of single paths (pixel, pass) with the render's RNG stream; rgb: n*3;
 * segments: n (nullable) - number of closest-hit queries each path made. */
int ilm_debug_path_radiance(ilm_ctx* ctx, ilm_scene* scene, const ilm_render_params* params,
                            const uint32_t* pixel_pass, int32_t n, float* rgb, int32_t* segments);
/* Instance i's matrices as the device uses them: out[0..15] object-to-world, [16..31] inverse,
 * [32..47] inverse-transpose (row-major). */
int ilm_debug_instance_matrices(const ilm_scene* scene, int32_t i, float* out48);

#ifdef __cplusplus
}
#endif

#endif /* ILLUME_C_H */
