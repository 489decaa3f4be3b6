/* bvh.h - host BVH builder (binned SAH) and its device layouts */
#pragma once
#include <cstdint>
#include <vector>

#include "common.h"

namespace ilm {

struct Box {
    float lo[3], hi[3];
    void reset() { for (int i = 0; i < 3; i++) { lo[i] = 3.402823466e+38f; hi[i] = -3.402823466e+38f; } }
    void grow(const Box& b) {
        for (int i = 0; i < 3; i++) { lo[i] = lo[i] < b.lo[i] ? lo[i] : b.lo[i]; hi[i] = hi[i] > b.hi[i] ? hi[i] : b.hi[i]; }
    }
    void grow(const float* p) {
        for (int i = 0; i < 3; i++) { lo[i] = lo[i] < p[i] ? lo[i] : p[i]; hi[i] = hi[i] > p[i] ? hi[i] : p[i]; }
    }
    double half_area() const {
        double dx = (double)hi[0] - lo[0], dy = (double)hi[1] - lo[1], dz = (double)hi[2] - lo[2];
        if (dx < 0 || dy < 0 || dz < 0) return 0.0;
        return dx * dy + dy * dz + dz * dx;
    }
};

/* Binary tree produced by the builder. Leaves own [first, first+count) of `order`. */
struct Node2 {
    Box box;
    int left = -1, right = -1;
    int first = 0, count = 0;
    int depth = 0;
    bool leaf() const { return left < 0; }
};

struct Tree {
    std::vector<Node2> nodes;      /* node 0 = root */
    std::vector<int> order;        /* primitive ids in leaf order */
    int max_depth = 0;
    int max_leaf = 0;
    int num_leaves = 0;
    double sah = 0.0;
};

struct BuildOpts {
    int max_leaf = 4;        /* SAH leaves hold at most this many primitives */
    int sah_depth_cap = 40;  /* below this depth only median splits */
    int bins = 32;
};

/* prim_boxes: one box per primitive */
Tree build_tree(const std::vector<Box>& prim_boxes, const BuildOpts& opts);

/* One box per triangle for build_tree: the vertices' bounds padded outward by 2^-20 of the
 * mesh's largest extent (see bvh.cpp). */
std::vector<Box> triangle_boxes(const float* pos, const int* idx, int num_triangles);

/* Device layouts */
struct Layout2 {
    std::vector<BNode> nodes;
    int root = ILM_NO_CHILD;       /* child code of the root */
};
Layout2 make_layout2(const Tree& t);

struct LayoutQ {
    std::vector<QNode> nodes;
    std::vector<Box> exact;        /* 4 exact child boxes per QNode, for tests */
    int root = ILM_NO_CHILD;
    int max_depth = 0;
};
LayoutQ make_layout_q(const Tree& t);

/* Largest child-box quantum of the 8-bit layout the traversal accepts (2^24): with
 * |1/d| <= 1e30 the kernel's per-node products quantum * (1/d) stay finite. */
#define ILM_QUANT_MAX_EXP 24
/* Largest quantum exponent over a set of QNodes (-127 when empty). */
int qnodes_max_exponent(const std::vector<QNode>& nodes);

/* Dequantise a QNode's child k box exactly as the kernel does. */
void qnode_child_box(const QNode& n, int k, float lo[3], float hi[3]);

}  // namespace ilm
