/*
 * bvh_cache.cpp - per-mesh BVH build with an on-disk cache of the device layouts.
 *
 * The reference caches its sweep-built BVH in an unversioned "<mesh>.bvh" file of 32-byte
 * GPUNodes with no check that the file matches the mesh (src/accel/bvh.cpp:399-439,
 * src/primitives/mesh.cu:407-414). Here a cache entry is keyed by a hash of the mesh's
 * positions, indices and build options, carries a format version and a hash of its payload,
 * and holds both device layouts (FP32 BVH2 and 8-bit BVH4) plus the leaf triangle order, so a
 * resident-scene server can cold-start without rebuilding. A file that fails any check is
 * ignored and rebuilt; writes go to a temporary name and are renamed into place.
 */
#include "bvh_cache.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <unistd.h>

namespace ilm {

namespace {

const char kMagic[8] = {'I', 'L', 'M', 'B', 'V', 'H', 0, 1};
const uint32_t kVersion = 2;   /* 2: padded triangle boxes and 8-bit child boxes (bvh.cpp) */

struct CacheHeader {
    char magic[8];
    uint32_t version;
    uint32_t header_bytes;
    uint64_t key;
    uint64_t payload_hash;
    int32_t root, qroot, depth, qdepth;
    uint64_t n_nodes, n_qnodes, n_order;
    uint32_t sizeof_bnode, sizeof_qnode;
};

/* 64-bit hash over a byte range, eight bytes at a time (multiply-xorshift mixing). */
struct Hasher {
    uint64_t h = 0x243f6a8885a308d3ull;
    void mix(uint64_t w) {
        h ^= w + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
        h *= 0xff51afd7ed558ccdull;
        h ^= h >> 33;
    }
    void bytes(const void* p, size_t n) {
        const unsigned char* c = (const unsigned char*)p;
        size_t i = 0;
        for (; i + 8 <= n; i += 8) {
            uint64_t w;
            std::memcpy(&w, c + i, 8);
            mix(w);
        }
        uint64_t tail = 0;
        std::memcpy(&tail, c + i, n - i);
        mix(tail ^ ((uint64_t)(n - i) << 56));
        mix((uint64_t)n);
    }
};

uint64_t payload_hash(const MeshBVH& b) {
    Hasher H;
    H.bytes(b.nodes.data(), b.nodes.size() * sizeof(BNode));
    H.bytes(b.qnodes.data(), b.qnodes.size() * sizeof(QNode));
    H.bytes(b.order.data(), b.order.size() * sizeof(int));
    return H.h;
}

/* A child code of either layout: an inner node index below n_nodes, or a leaf of 1..16
 * triangles inside [0, n_order) whose slots are marked in cover (each slot once). */
bool leaf_ok(int32_t c, size_t n_order, std::vector<unsigned char>& cover) {
    const uint64_t first = ilm_leaf_first(c), count = ilm_leaf_count(c);
    if (first + count > n_order) return false;
    for (uint64_t k = first; k < first + count; k++) {
        if (cover[k]) return false;
        cover[k] = 1;
    }
    return true;
}

/* Walks both layouts from their roots (the file's own child indices are never trusted): every
 * inner index in range and reached exactly once (a tree: no cycle, no sharing, no orphan), every
 * leaf range inside the triangle order and the leaves of each layout covering every slot once.
 * Sets depth and qdepth to what the walk finds: the FP32 tree's depth counts the leaf level
 * (root 0), the 8-bit tree's the depth of its deepest node. */
bool layouts_valid(MeshBVH& b) {
    const size_t n_order = b.order.size();
    struct Item { int32_t node; int depth; };
    std::vector<Item> st;
    /* FP32 BVH2 */
    {
        std::vector<unsigned char> seen(b.nodes.size(), 0), cover(n_order, 0);
        int depth = 0;
        size_t visited = 0;
        if (b.root == ILM_NO_CHILD) {
            if (!b.nodes.empty() || n_order != 0) return false;
        } else if (b.root < 0) {
            if (!b.nodes.empty() || !leaf_ok(b.root, n_order, cover)) return false;
        } else {
            if (b.root != 0 || b.nodes.empty()) return false;
            st.push_back({0, 0});
        }
        while (!st.empty()) {
            const Item it = st.back();
            st.pop_back();
            if (it.node < 0 || (size_t)it.node >= b.nodes.size() || seen[it.node]) return false;
            seen[it.node] = 1;
            visited++;
            const int32_t kids[2] = {b.nodes[it.node].d.x, b.nodes[it.node].d.y};
            for (int32_t c : kids) {
                if (c == ILM_NO_CHILD) return false;
                if (c >= 0) {
                    st.push_back({c, it.depth + 1});
                } else {
                    if (!leaf_ok(c, n_order, cover)) return false;
                    depth = std::max(depth, it.depth + 1);
                }
            }
        }
        if (visited != b.nodes.size()) return false;
        for (unsigned char v : cover)
            if (!v) return false;
        b.depth = depth;
    }
    /* 8-bit BVH4 */
    {
        std::vector<unsigned char> seen(b.qnodes.size(), 0), cover(n_order, 0);
        int depth = 0;
        size_t visited = 0;
        if (b.qroot == ILM_NO_CHILD) {
            if (!b.qnodes.empty() || n_order != 0) return false;
        } else if (b.qroot < 0) {
            if (!b.qnodes.empty() || !leaf_ok(b.qroot, n_order, cover)) return false;
        } else {
            if (b.qroot != 0 || b.qnodes.empty()) return false;
            st.push_back({0, 0});
        }
        while (!st.empty()) {
            const Item it = st.back();
            st.pop_back();
            if (it.node < 0 || (size_t)it.node >= b.qnodes.size() || seen[it.node]) return false;
            seen[it.node] = 1;
            visited++;
            depth = std::max(depth, it.depth);
            const QNode& q = b.qnodes[it.node];
            if (q.nchild < 2 || q.nchild > 4) return false;   /* the 8-bit box test assumes two or more */
            for (int k = 0; k < 4; k++) {
                const int32_t c = q.child[k];
                if (k >= q.nchild) {
                    if (c != ILM_NO_CHILD) return false;
                } else if (c == ILM_NO_CHILD) {
                    return false;
                } else if (c >= 0) {
                    st.push_back({c, it.depth + 1});
                } else if (!leaf_ok(c, n_order, cover)) {
                    return false;
                }
            }
        }
        if (visited != b.qnodes.size()) return false;
        for (unsigned char v : cover)
            if (!v) return false;
        b.qdepth = depth;
    }
    return true;
}

std::string entry_path(const std::string& dir, uint64_t key) {
    char name[32];
    std::snprintf(name, sizeof(name), "%016llx.ilmbvh", (unsigned long long)key);
    return dir + "/" + name;
}

}  // namespace

uint64_t mesh_key(const float* pos, int num_vertices, const int* idx, int num_triangles, const BuildOpts& o) {
    Hasher H;
    H.mix(kVersion);
    H.mix((uint64_t)num_vertices);
    H.mix((uint64_t)num_triangles);
    H.mix(((uint64_t)o.max_leaf << 40) ^ ((uint64_t)o.sah_depth_cap << 20) ^ (uint64_t)o.bins);
    H.mix(sizeof(BNode) | (sizeof(QNode) << 16));
    H.bytes(pos, 3 * sizeof(float) * (size_t)num_vertices);
    H.bytes(idx, 3 * sizeof(int) * (size_t)num_triangles);
    return H.h;
}

MeshBVH build_mesh_bvh(const float* pos, const int* idx, int num_triangles, const BuildOpts& o) {
    Tree tree = build_tree(triangle_boxes(pos, idx, num_triangles), o);
    Layout2 L2 = make_layout2(tree);
    LayoutQ LQ = make_layout_q(tree);
    MeshBVH b;
    b.nodes = std::move(L2.nodes);
    b.qnodes = std::move(LQ.nodes);
    b.order = std::move(tree.order);
    b.root = L2.root;
    b.qroot = LQ.root;
    b.depth = tree.max_depth;
    b.qdepth = LQ.max_depth;
    return b;
}

bool cache_load(const std::string& dir, uint64_t key, int num_triangles, MeshBVH& out) {
    FILE* f = std::fopen(entry_path(dir, key).c_str(), "rb");
    if (!f) return false;
    CacheHeader h;
    bool ok = std::fread(&h, sizeof(h), 1, f) == 1 && std::memcmp(h.magic, kMagic, 8) == 0 &&
              h.version == kVersion && h.header_bytes == sizeof(CacheHeader) && h.key == key &&
              h.sizeof_bnode == sizeof(BNode) && h.sizeof_qnode == sizeof(QNode) &&
              h.n_order == (uint64_t)num_triangles && h.n_nodes < (1ull << 31) && h.n_qnodes < (1ull << 31);
    if (ok) {
        MeshBVH b;
        b.nodes.resize(h.n_nodes);
        b.qnodes.resize(h.n_qnodes);
        b.order.resize(h.n_order);
        ok = std::fread(b.nodes.data(), sizeof(BNode), h.n_nodes, f) == h.n_nodes &&
             std::fread(b.qnodes.data(), sizeof(QNode), h.n_qnodes, f) == h.n_qnodes &&
             std::fread(b.order.data(), sizeof(int), h.n_order, f) == h.n_order;
        if (ok) {
            b.root = h.root;
            b.qroot = h.qroot;
            b.depth = h.depth;
            b.qdepth = h.qdepth;
            ok = payload_hash(b) == h.payload_hash;
            /* leaf order must be a permutation of the triangles */
            if (ok) {
                std::vector<unsigned char> seen(num_triangles, 0);
                for (int t : b.order) {
                    if (t < 0 || t >= num_triangles || seen[t]) { ok = false; break; }
                    seen[t] = 1;
                }
            }
            /* the tree itself: indices in range, a proper tree, depths recomputed (a stale or
             * crafted entry whose header understates them would overflow the traversal stack) */
            if (ok) ok = layouts_valid(b) && b.depth == h.depth && b.qdepth == h.qdepth;
            if (ok) out = std::move(b);
        }
    }
    std::fclose(f);
    return ok;
}

bool cache_store(const std::string& dir, uint64_t key, const MeshBVH& b) {
    CacheHeader h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, kMagic, 8);
    h.version = kVersion;
    h.header_bytes = sizeof(CacheHeader);
    h.key = key;
    h.payload_hash = payload_hash(b);
    h.root = b.root;
    h.qroot = b.qroot;
    h.depth = b.depth;
    h.qdepth = b.qdepth;
    h.n_nodes = b.nodes.size();
    h.n_qnodes = b.qnodes.size();
    h.n_order = b.order.size();
    h.sizeof_bnode = sizeof(BNode);
    h.sizeof_qnode = sizeof(QNode);
    const std::string path = entry_path(dir, key);
    const std::string tmp = path + ".tmp" + std::to_string((long long)getpid());
    FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) return false;
    bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1 &&
              std::fwrite(b.nodes.data(), sizeof(BNode), b.nodes.size(), f) == b.nodes.size() &&
              std::fwrite(b.qnodes.data(), sizeof(QNode), b.qnodes.size(), f) == b.qnodes.size() &&
              std::fwrite(b.order.data(), sizeof(int), b.order.size(), f) == b.order.size();
    ok = (std::fclose(f) == 0) && ok;
    if (ok) ok = std::rename(tmp.c_str(), path.c_str()) == 0;
    if (!ok) std::remove(tmp.c_str());
    return ok;
}

}  // namespace ilm
