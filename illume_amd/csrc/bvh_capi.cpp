/*
 * bvh_capi.cpp - host-only C ABI over the BVH builder, so the builder the scene upload uses can
 * be inspected (invariants, SAH cost, quantisation) without a GPU.
 * Reference counterpart: bvh_create / GPUNode (src/accel/bvh.h:11-42, src/accel/bvh.cpp:399-439).
 */
#include <cstring>
#include <string>

#include "bvh.h"
#include "bvh_cache.h"

using namespace ilm;

struct ilm_bvh {
    Tree tree;
    LayoutQ q;
};

extern "C" int ilm_bvh_build(const ilm_mesh_desc* m, ilm_bvh** out) {
    if (!m || !out || m->num_triangles < 0 || (m->num_triangles > 0 && (!m->positions || !m->indices))) return ILM_E_INVALID;
    for (int64_t k = 0; k < 3LL * m->num_triangles; k++)
        if (m->indices[k] < 0 || m->indices[k] >= m->num_vertices) return ILM_E_INVALID;
    ilm_bvh* b = new ilm_bvh();
    BuildOpts o;
    b->tree = build_tree(triangle_boxes(m->positions, m->indices, m->num_triangles), o);
    b->q = make_layout_q(b->tree);
    *out = b;
    return ILM_OK;
}

extern "C" int ilm_bvh_get_info(const ilm_bvh* b, ilm_bvh_info* info) {
    if (!b || !info) return ILM_E_INVALID;
    info->num_nodes = (int32_t)b->tree.nodes.size();
    info->num_leaves = b->tree.num_leaves;
    info->max_depth = b->tree.max_depth;
    info->max_leaf_size = b->tree.max_leaf;
    info->sah_cost = b->tree.sah;
    return ILM_OK;
}

extern "C" int ilm_bvh_export(const ilm_bvh* b, float* boxes, int32_t* children, int32_t* leaf_first,
                              int32_t* leaf_count, int32_t* tri_order) {
    if (!b) return ILM_E_INVALID;
    const Tree& t = b->tree;
    for (size_t i = 0; i < t.nodes.size(); i++) {
        const Node2& n = t.nodes[i];
        if (boxes) for (int a = 0; a < 3; a++) { boxes[6 * i + a] = n.box.lo[a]; boxes[6 * i + 3 + a] = n.box.hi[a]; }
        if (children) { children[2 * i] = n.left; children[2 * i + 1] = n.right; }
        if (leaf_first) leaf_first[i] = n.leaf() ? n.first : 0;
        if (leaf_count) leaf_count[i] = n.leaf() ? n.count : 0;
    }
    if (tri_order) std::memcpy(tri_order, t.order.data(), t.order.size() * sizeof(int32_t));
    return ILM_OK;
}

/* Quantised layout: number of QNodes; per QNode 4 children: dequantised box (6), exact box (6),
 * child code. */
extern "C" int ilm_bvh_quantized_size(const ilm_bvh* b, int32_t* num_qnodes, int32_t* root, int32_t* max_depth) {
    if (!b) return ILM_E_INVALID;
    if (num_qnodes) *num_qnodes = (int32_t)b->q.nodes.size();
    if (root) *root = b->q.root;
    if (max_depth) *max_depth = b->q.max_depth;
    return ILM_OK;
}

extern "C" int ilm_bvh_export_quantized(const ilm_bvh* b, float* qboxes, float* exact, int32_t* child) {
    if (!b) return ILM_E_INVALID;
    for (size_t i = 0; i < b->q.nodes.size(); i++) {
        const QNode& n = b->q.nodes[i];
        for (int k = 0; k < 4; k++) {
            float lo[3], hi[3];
            qnode_child_box(n, k, lo, hi);
            for (int a = 0; a < 3; a++) {
                if (qboxes) { qboxes[24 * i + 6 * k + a] = lo[a]; qboxes[24 * i + 6 * k + 3 + a] = hi[a]; }
                if (exact) {
                    exact[24 * i + 6 * k + a] = b->q.exact[4 * i + k].lo[a];
                    exact[24 * i + 6 * k + 3 + a] = b->q.exact[4 * i + k].hi[a];
                }
            }
            if (child) child[4 * i + k] = n.child[k];
        }
    }
    return ILM_OK;
}

extern "C" void ilm_bvh_free(ilm_bvh* b) { delete b; }

/* The device layouts ilm_scene_upload would use for this mesh, built or loaded from cache_dir
 * (NULL: no cache). *nodes_bytes / *qnodes_bytes hold the buffers' capacities on entry and the
 * layout sizes on return (ILM_E_INVALID when a buffer is too small; NULL buffers only query).
 * order: num_triangles ints. info4: root, qroot, depth, qdepth. */
extern "C" int ilm_bvh_device_layouts(const ilm_mesh_desc* m, const char* cache_dir, int32_t* from_cache, void* nodes,
                                      int64_t* nodes_bytes, void* qnodes, int64_t* qnodes_bytes, int32_t* order,
                                      int32_t* info4) {
    if (!m || !nodes_bytes || !qnodes_bytes || m->num_triangles < 0 ||
        (m->num_triangles > 0 && (!m->positions || !m->indices)))
        return ILM_E_INVALID;
    for (int64_t k = 0; k < 3LL * m->num_triangles; k++)
        if (m->indices[k] < 0 || m->indices[k] >= m->num_vertices) return ILM_E_INVALID;
    BuildOpts o;
    MeshBVH b;
    bool hit = false;
    uint64_t key = 0;
    if (cache_dir) {
        key = mesh_key(m->positions, m->num_vertices, m->indices, m->num_triangles, o);
        hit = cache_load(cache_dir, key, m->num_triangles, b);
    }
    if (!hit) {
        b = build_mesh_bvh(m->positions, m->indices, m->num_triangles, o);
        if (cache_dir) (void)cache_store(cache_dir, key, b);
    }
    if (from_cache) *from_cache = hit ? 1 : 0;
    const int64_t nb = (int64_t)(b.nodes.size() * sizeof(BNode)), qb = (int64_t)(b.qnodes.size() * sizeof(QNode));
    const bool fits = (!nodes || *nodes_bytes >= nb) && (!qnodes || *qnodes_bytes >= qb);
    *nodes_bytes = nb;
    *qnodes_bytes = qb;
    if (!fits) return ILM_E_INVALID;
    if (nodes) std::memcpy(nodes, b.nodes.data(), nb);
    if (qnodes) std::memcpy(qnodes, b.qnodes.data(), qb);
    if (order) std::memcpy(order, b.order.data(), b.order.size() * sizeof(int32_t));
    if (info4) { info4[0] = b.root; info4[1] = b.qroot; info4[2] = b.depth; info4[3] = b.qdepth; }
    return ILM_OK;
}
