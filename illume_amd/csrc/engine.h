/* engine.h - host-side handle types of the C ABI */
#pragma once
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "common.h"

#define ILM_RING 8

struct ilm_scene {
    ~ilm_scene() {   /* device memory of an upload that failed part-way */
        for (void* p : allocs) (void)hipFree(p);
    }
    ilm_ctx* ctx = nullptr;
    std::atomic<int> refs{0};
    std::vector<void*> allocs;
    int64_t bytes = 0;
    DScene* d_scene = nullptr;
    DScene h_scene;
    ilm_camera_desc camera;
    std::vector<float> inst_mats;   /* 48 per instance: mat, inv, inverse-transpose */
    int num_instances = 0;
    int64_t mesh_tris = 0;
    int stack_bound_fp32 = 0;
    int stack_bound_q = 0;
    int quant_max_exp = -127;   /* largest child-box quantum exponent of the 8-bit trees */
};

/* One render request. */
/* Request buffers (Kahan sums, pixels) are recycled instead of freed: hipFree waits for the
 * whole device, which would stall the scheduler behind the chunk in flight. Holds at most
 * `cap` bytes; a buffer is reused for a request needing between half and all of it. */
struct BufPool {
    std::mutex mu;                          /* leaf lock */
    std::multimap<size_t, void*> free;
    size_t bytes = 0, cap = (size_t)2 << 30;
    void* take(size_t n, size_t* got) {
        std::lock_guard<std::mutex> lk(mu);
        auto it = free.lower_bound(n);
        if (it == free.end() || it->first > 2 * n) return nullptr;
        void* p = it->second;
        *got = it->first;
        bytes -= it->first;
        free.erase(it);
        return p;
    }
    bool give(void* p, size_t n) {          /* false: caller frees */
        std::lock_guard<std::mutex> lk(mu);
        if (bytes + n > cap) return false;
        free.emplace(n, p);
        bytes += n;
        return true;
    }
};

struct Job {
    ilm_scene* scene = nullptr;
    ilm_render_params p;
    DReq dreq;
    float4* d_accum = nullptr;      /* 2*W*H: Kahan sums then compensations */
    ilm_pixel* d_px = nullptr;
    size_t accum_cap = 0, px_cap = 0;   /* allocation sizes (pool buffers may be larger) */
    BufPool* pool = nullptr;
    int64_t bytes = 0;              /* device bytes while admitted */
    int next_pass = 0;              /* issue cursor: pass, pixel */
    int64_t next_px = 0;
    int64_t split_npx = 0;          /* > 0: issued as pixel ranges of this many slots, all passes each */
    /* served form */
    float4* h_accum = nullptr;      /* pinned results */
    ilm_pixel* h_px = nullptr;
    std::atomic<int> state{0};
    std::atomic<bool> cancel{false};
    bool failed = false;
    double submit_ms = 0, start_ms = 0, end_ms = 0;
    int samples_done = 0;
    int64_t last_ev_seq = -1;
    int refs = 1;
    int device = 0;                 /* of the context (served form), for freeing the pinned results */
    std::condition_variable cv;
    ~Job();
};

struct ilm_req {
    Job* job = nullptr;
    ilm_ctx* ctx = nullptr;
};

struct ProfileSink {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
};

/* Device buffers one chunk pipeline runs on: the context's own (allocated at ilm_open), or those
 * of ilm_render_noalloc, carved from its static device array. */
struct Pipe {
    hipStream_t stream = nullptr;
    int idx = 0;                    /* 0 or 1 */
    DPaths A, B;                    /* path state, ping-ponged across depths */
    DHits hits;
    float4* contrib = nullptr;      /* per-path radiance */
    DCtl* ctl = nullptr;
    DChunk* chunk = nullptr;        /* device copy of the chunk descriptor */
    int* ovf = nullptr;             /* traversal-stack spill area, sized for ext_grid */
    int ext_grid = 0;
    unsigned long long* seg_total = nullptr;
};

/* Latencies of the most recent completed requests (p50/p99 are taken over this window). */
#define ILM_LATENCY_WINDOW 4096

struct ilm_ctx {
    /* held by the caller's handle and by every live scene and request handle: the context's
     * device resources are released when the last of them goes */
    std::atomic<int> refs{1};
    int device = 0;
    hipStream_t stream = nullptr;
    /* host<->device copies made by caller threads (scene uploads, result readback). A blocking
     * copy on the legacy stream while the scheduler thread captures a graph fails with
     * hipErrorStreamCaptureImplicit, so those copies go through this non-blocking stream. */
    hipStream_t copy_stream = nullptr;
    ilm_opts opts;
    int n_cu = 0;
    int ext_grid = 0, shade_grid = 0, gen_grid = 0;
    int ovf_grid = 0;               /* extend grid the spill area (ovf) is sized for, >= ext_grid */
    std::chrono::steady_clock::time_point t0;

    /* device buffers */
    int64_t cap = 0;                /* paths per chunk */
    DPaths A, B;
    DHits hits;
    float4* contrib = nullptr;
    int* seg_count = nullptr;
    int* ovf = nullptr;
    DCtl* ctl = nullptr;
    DChunk* d_chunk = nullptr;
    /* the second chunk pipeline (opts.pipes == 2): consecutive chunks alternate between the
     * two, each on its own stream, so one chunk's last short launches run beside the next
     * chunk's first ones; accumulations stay in chunk order (acc_ev) */
    int npipes = 1;
    int next_pipe = 0;
    hipStream_t stream2 = nullptr;
    DPaths A2, B2;
    DHits hits2;
    float4* contrib2 = nullptr;
    DCtl* ctl2 = nullptr;
    DChunk* d_chunk2 = nullptr;
    hipEvent_t acc_ev[2] = {};      /* after each pipeline's latest accumulation */
    hipEvent_t fill_ev = nullptr;   /* after the latest clear of request buffers (stream) */
    int last_pipe = 0;              /* pipeline of the latest chunk */
    unsigned long long* d_seg_total = nullptr;
    unsigned long long* d_trav = nullptr;
    std::vector<void*> allocs;
    BufPool bufs;
    std::vector<void*> path_allocs;  /* path state, hits, contributions (resized by ilm_tune) */
    int64_t own_bytes = 0, path_bytes = 0;

    /* chunk descriptor staging */
    DChunk* h_chunk[ILM_RING] = {};
    hipEvent_t ring_ev[ILM_RING] = {};
    int ring_pos = 0;
    unsigned long long* h_u64 = nullptr;

    std::map<int, hipGraphExec_t> graphs;
    std::mutex dev_mu;              /* one user of the device pipeline at a time */

    /* serving */
    std::mutex sched_mu;
    std::condition_variable sched_cv;
    std::deque<Job*> pending;
    std::vector<Job*> active;
    std::thread server;
    bool stop = false;
    int64_t budget = 0;
    int64_t in_use = 0;
    int64_t chunk_seq = 0;
    int64_t chunks_retired = 0;
    std::vector<double> latencies;  /* ring of ILM_LATENCY_WINDOW entries */
    size_t lat_next = 0;
    uint64_t stat_done = 0, stat_rejected = 0, stat_cancelled = 0;
    /* segment counter as of the last completed chunk or render (stat_mu); read by ilm_stats_get
     * without touching the device. h_seg_ring: pinned slots the device counter is copied into
     * behind each served chunk, read when that chunk's event completes. */
    uint64_t stat_segments = 0;
    unsigned long long* h_seg_ring = nullptr;   /* ILM_RING slots */
    int seg_ring_pos = 0;
    /* ilm_render_noalloc: this device's address of the static scratch array (from
     * hipGetSymbolAddress at ilm_open) and the timing events of its renders */
    void* noalloc_scratch = nullptr;
    hipEvent_t na_ev[2] = {};
    uint64_t stat_bvh_hits = 0, stat_bvh_misses = 0;
    std::string bvh_cache_dir;      /* empty: no BVH cache */
    double stat_busy_ms = 0.0;

    /* Lock order: sched_mu before dev_mu; stat_mu is a leaf (no other lock is taken while it
     * is held). Nothing takes sched_mu while holding dev_mu. */
    std::mutex stat_mu;             /* stat_busy_ms, stat_bvh_*, stat_segments */
    std::mutex err_mu;
    std::string err, err_copy;
};
