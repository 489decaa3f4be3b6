/* bvh_cache.h - per-mesh BVH build (both device layouts) and its versioned on-disk cache */
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "bvh.h"

namespace ilm {

struct MeshBVH {
    std::vector<BNode> nodes;   /* FP32 BVH2 layout */
    std::vector<QNode> qnodes;  /* 8-bit BVH4 layout */
    std::vector<int> order;     /* triangle ids in leaf order */
    int root = ILM_NO_CHILD, qroot = ILM_NO_CHILD;
    int depth = 0, qdepth = 0;  /* tree depths (traversal stack bounds) */
};

/* Cache key: hash of the format version, build options, positions and indices. */
uint64_t mesh_key(const float* pos, int num_vertices, const int* idx, int num_triangles, const BuildOpts& o);

MeshBVH build_mesh_bvh(const float* pos, const int* idx, int num_triangles, const BuildOpts& o);

/* Load dir/<key>.ilmbvh; false (and `out` untouched) unless every check passes. */
bool cache_load(const std::string& dir, uint64_t key, int num_triangles, MeshBVH& out);
bool cache_store(const std::string& dir, uint64_t key, const MeshBVH& b);

}  // namespace ilm
