/*
 * bvh.cpp - binned-SAH BVH builder and its two device layouts.
 *
 * Replaces the reference's sweep builder (src/accel/bvh.cpp:65-249: ~1024/(depth+1) planes per
 * axis, each costed by a full pass) with a binned SAH build (32 centroid bins per axis,
 * O(N) per level), a depth cap after which only object-median splits are made (so the device
 * traversal stack has a proven bound), and two flat layouts:
 *   - FP32 BVH2, 64 B nodes holding both children's boxes (BNode)
 *   - 8-bit quantised BVH4, 64 B nodes (QNode): child boxes relative to the node origin on a
 *     power-of-two grid, lo rounded down and hi rounded up, each verified in float so the
 *     dequantised box the kernel computes always contains the exact box.
 */
#include "bvh.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace ilm {

namespace {

struct Builder {
    const std::vector<Box>& pb;
    const BuildOpts& o;
    std::vector<float> cen;      /* 3 per prim */
    Tree& t;
    double root_area = 1.0;

    Builder(const std::vector<Box>& p, const BuildOpts& op, Tree& tr) : pb(p), o(op), t(tr) {}

    int make_leaf(int me, int first, int count) {
        Node2& n = t.nodes[me];
        n.first = first;
        n.count = count;
        n.left = n.right = -1;
        t.num_leaves++;
        t.max_leaf = std::max(t.max_leaf, count);
        t.sah += n.box.half_area() / root_area * (double)count;
        return me;
    }

    int median_split(int first, int count, int axis) {
        int* b = t.order.data() + first;
        int mid = count / 2;
        std::nth_element(b, b + mid, b + count, [&](int x, int y) {
            float cx = cen[3 * x + axis], cy = cen[3 * y + axis];
            return cx < cy || (cx == cy && x < y);
        });
        return mid;
    }

    int build(int first, int count, int depth) {
        int me = (int)t.nodes.size();
        t.nodes.emplace_back();
        Box bounds, cb;
        bounds.reset();
        cb.reset();
        for (int i = first; i < first + count; i++) {
            int p = t.order[i];
            bounds.grow(pb[p]);
            cb.grow(&cen[3 * p]);
        }
        t.nodes[me].box = bounds;
        t.nodes[me].depth = depth;
        t.max_depth = std::max(t.max_depth, depth);
        if (me == 0) root_area = std::max(bounds.half_area(), 1e-30);
        if (count <= 1) return make_leaf(me, first, count);

        int ext_axis = 0;
        float ext[3];
        for (int a = 0; a < 3; a++) ext[a] = cb.hi[a] - cb.lo[a];
        if (ext[1] > ext[ext_axis]) ext_axis = 1;
        if (ext[2] > ext[ext_axis]) ext_axis = 2;

        int nl = -1;
        if (depth >= o.sah_depth_cap) {
            if (count <= ILM_LEAF_MAX) return make_leaf(me, first, count);
            nl = median_split(first, count, ext_axis);
        } else {
            const int B = o.bins;
            double best = 1e300;
            int best_axis = -1, best_bin = -1;
            std::vector<Box> bb(B);
            std::vector<int> bc(B);
            std::vector<double> right_cost(B);
            for (int a = 0; a < 3; a++) {
                if (!(ext[a] > 0.0f)) continue;
                for (int k = 0; k < B; k++) { bb[k].reset(); bc[k] = 0; }
                const double scale = (double)B / (double)ext[a];
                for (int i = first; i < first + count; i++) {
                    int p = t.order[i];
                    int k = (int)(((double)cen[3 * p + a] - cb.lo[a]) * scale);
                    k = k < 0 ? 0 : (k >= B ? B - 1 : k);
                    bb[k].grow(pb[p]);
                    bc[k]++;
                }
                Box acc;
                acc.reset();
                int n = 0;
                for (int k = B - 1; k >= 1; k--) {
                    acc.grow(bb[k]);
                    n += bc[k];
                    right_cost[k] = n ? acc.half_area() * n : 0.0;
                }
                acc.reset();
                n = 0;
                for (int k = 0; k < B - 1; k++) {
                    acc.grow(bb[k]);
                    n += bc[k];
                    if (n == 0 || n == count) continue;
                    double c = acc.half_area() * n + right_cost[k + 1];
                    if (c < best) { best = c; best_axis = a; best_bin = k; }
                }
            }
            const double area = std::max(bounds.half_area(), 1e-30);
            const double split_cost = 1.0 + best / area;
            const double leaf_cost = (double)count;
            if (best_axis >= 0 && !(count <= o.max_leaf && leaf_cost <= split_cost)) {
                const double scale = (double)B / (double)ext[best_axis];
                const float lo = cb.lo[best_axis];
                int* mid = std::partition(t.order.data() + first, t.order.data() + first + count, [&](int p) {
                    int k = (int)(((double)cen[3 * p + best_axis] - lo) * scale);
                    k = k < 0 ? 0 : (k >= B ? B - 1 : k);
                    return k <= best_bin;
                });
                nl = (int)(mid - (t.order.data() + first));
                if (nl == 0 || nl == count) nl = -1;
            } else if (count <= o.max_leaf) {
                return make_leaf(me, first, count);
            }
            if (nl < 0) {
                if (count <= o.max_leaf) return make_leaf(me, first, count);
                nl = median_split(first, count, ext_axis);
            }
        }
        t.sah += bounds.half_area() / root_area;
        int l = build(first, nl, depth + 1);
        int r = build(first + nl, count - nl, depth + 1);
        t.nodes[me].left = l;
        t.nodes[me].right = r;
        return me;
    }
};

}  // namespace

Tree build_tree(const std::vector<Box>& prim_boxes, const BuildOpts& opts) {
    Tree t;
    const int n = (int)prim_boxes.size();
    t.order.resize(n);
    std::iota(t.order.begin(), t.order.end(), 0);
    if (n == 0) return t;
    Builder b(prim_boxes, opts, t);
    b.cen.resize(3 * (size_t)n);
    for (int i = 0; i < n; i++)
        for (int a = 0; a < 3; a++) b.cen[3 * i + a] = 0.5f * (prim_boxes[i].lo[a] + prim_boxes[i].hi[a]);
    t.nodes.reserve(2 * (size_t)n / std::max(1, opts.max_leaf) + 16);
    b.build(0, n, 0);
    return t;
}

/* Moller-Trumbore is not watertight: near an edge it can report a hit slightly outside the
 * triangle (its barycentrics carry rounding of about 1e-7 of the distance from the ray origin to
 * the triangle). With boxes tight on the vertices, such a hit is found or not depending on which
 * boxes a walk enters, so two node layouts, or a packet walk and a lone ray, could settle on
 * different triangles at nearly the same distance. Padding every triangle box by 2^-20 of the
 * mesh extent keeps those hits inside every box on their way for rays that start within a few
 * mesh extents of the triangle, so every walk over either layout finds them. */
std::vector<Box> triangle_boxes(const float* pos, const int* idx, int num_triangles) {
    std::vector<Box> pb(num_triangles > 0 ? num_triangles : 0);
    Box all;
    all.reset();
    for (int t = 0; t < num_triangles; t++) {
        pb[t].reset();
        for (int j = 0; j < 3; j++) pb[t].grow(&pos[3 * idx[3 * t + j]]);
        all.grow(pb[t]);
    }
    if (num_triangles == 0) return pb;
    double ext = 0.0;
    for (int a = 0; a < 3; a++) ext = std::max(ext, (double)all.hi[a] - (double)all.lo[a]);
    const double pad = ext / 1048576.0;
    if (!(pad > 0.0)) return pb;
    for (Box& b : pb)
        for (int a = 0; a < 3; a++) {
            float lo = (float)((double)b.lo[a] - pad), hi = (float)((double)b.hi[a] + pad);
            if ((double)lo > (double)b.lo[a] - pad) lo = std::nextafter(lo, -INFINITY);
            if ((double)hi < (double)b.hi[a] + pad) hi = std::nextafter(hi, INFINITY);
            b.lo[a] = lo;
            b.hi[a] = hi;
        }
    return pb;
}

static int child_code(const Tree& t, const std::vector<int>& inner_id, int c) {
    const Node2& n = t.nodes[c];
    if (n.leaf()) return ilm_leaf_code((uint32_t)n.first, (uint32_t)std::max(1, n.count));
    return inner_id[c];
}

Layout2 make_layout2(const Tree& t) {
    Layout2 L;
    if (t.nodes.empty()) return L;
    std::vector<int> inner_id(t.nodes.size(), -1);
    int k = 0;
    for (size_t i = 0; i < t.nodes.size(); i++)
        if (!t.nodes[i].leaf()) inner_id[i] = k++;
    L.nodes.resize(k);
    for (size_t i = 0; i < t.nodes.size(); i++) {
        const Node2& n = t.nodes[i];
        if (n.leaf()) continue;
        const Box& b0 = t.nodes[n.left].box;
        const Box& b1 = t.nodes[n.right].box;
        BNode& o = L.nodes[inner_id[i]];
        o.a = make_float4(b0.lo[0], b0.hi[0], b0.lo[1], b0.hi[1]);
        o.b = make_float4(b0.lo[2], b0.hi[2], b1.lo[0], b1.hi[0]);
        o.c = make_float4(b1.lo[1], b1.hi[1], b1.lo[2], b1.hi[2]);
        o.d = make_int4(child_code(t, inner_id, n.left), child_code(t, inner_id, n.right), 0, 0);
    }
    L.root = t.nodes[0].leaf() ? child_code(t, inner_id, 0) : 0;
    if (t.nodes[0].leaf() && t.nodes[0].count == 0) L.root = ILM_NO_CHILD;
    return L;
}

static inline float q_deq(float o, int q, float s) { return o + (float)q * s; }

void qnode_child_box(const QNode& n, int k, float lo[3], float hi[3]) {
    const float org[3] = {n.ox, n.oy, n.oz};
    const uint8_t ex[3] = {n.ex, n.ey, n.ez};
    for (int a = 0; a < 3; a++) {
        uint32_t bits = (uint32_t)ex[a] << 23;
        float s;
        std::memcpy(&s, &bits, 4);
        lo[a] = q_deq(org[a], n.qlo[a][k], s);
        hi[a] = q_deq(org[a], n.qhi[a][k], s);
    }
}

int qnodes_max_exponent(const std::vector<QNode>& nodes) {
    int e = 0;
    for (const QNode& n : nodes) e = std::max(e, (int)std::max(n.ex, std::max(n.ey, n.ez)));
    return e - 127;
}

/* Outward pad of every quantised child box, as a fraction of its node's extent per axis. The
 * traversal forms a plane's slab distance from B = (o_n - o)/d, whose rounding leaves an
 * absolute error of at most 2^-23 * 255 * s / |d| (s the axis quantum, 255 s <= 2.03 x the
 * node extent); moving every plane out by 2^-20 x the node extent widens each slab interval by
 * four times that, so the rounded test still contains the exact box whatever the ray origin. */
#define QPAD_REL (1.0 / 1048576.0)

static void quantize_node(QNode& q, const Box& nb, const Box* cb, int nc) {
    /* per axis: the pad, the padded child bounds (double) and an origin below the padded node */
    double pad[3], plo[4][3], phi[4][3];
    float org[3];
    for (int a = 0; a < 3; a++) {
        const double ext = (double)nb.hi[a] - (double)nb.lo[a];
        pad[a] = ext > 0.0 ? ext * QPAD_REL : 0.0;
        for (int k = 0; k < nc; k++) {
            plo[k][a] = (double)cb[k].lo[a] - pad[a];
            phi[k][a] = (double)cb[k].hi[a] + pad[a];
        }
        float o = (float)((double)nb.lo[a] - pad[a]);
        if ((double)o > (double)nb.lo[a] - pad[a]) o = std::nextafter(o, -INFINITY);
        org[a] = o;
    }
    q.ox = org[0]; q.oy = org[1]; q.oz = org[2];
    uint8_t* ex[3] = {&q.ex, &q.ey, &q.ez};
    for (int a = 0; a < 3; a++) {
        double ext = (double)nb.hi[a] + pad[a] - (double)org[a];
        int e;
        if (!(ext > 0.0)) e = -126;
        else {
            std::frexp(ext / 252.0, &e);   /* ext/252 < 2^e */
            if (e < -126) e = -126;
            if (e > 127) e = 127;
        }
        float s = std::ldexp(1.0f, e);
        for (;;) {
            bool ok = true;
            for (int k = 0; k < nc && ok; k++) {
                double fl = std::floor((plo[k][a] - org[a]) / s);
                double ch = std::ceil((phi[k][a] - org[a]) / s);
                int ql = (int)std::max(0.0, std::min(255.0, fl));
                int qh = (int)std::max(0.0, std::min(255.0, ch));
                /* the float plane the kernel forms (org + q*s, q*s exact) must lie outside the
                 * padded bound */
                while (ql > 0 && (double)q_deq(org[a], ql, s) > plo[k][a]) ql--;
                while (qh < 255 && (double)q_deq(org[a], qh, s) < phi[k][a]) qh++;
                if ((double)q_deq(org[a], ql, s) > plo[k][a] || (double)q_deq(org[a], qh, s) < phi[k][a]) { ok = false; break; }
                q.qlo[a][k] = (uint8_t)ql;
                q.qhi[a][k] = (uint8_t)qh;
            }
            if (ok) break;
            e++;
            s = std::ldexp(1.0f, e);
        }
        *ex[a] = (uint8_t)(e + 127);
        for (int k = nc; k < 4; k++) { q.qlo[a][k] = 0; q.qhi[a][k] = 0; }
    }
}

LayoutQ make_layout_q(const Tree& t) {
    LayoutQ L;
    if (t.nodes.empty()) return L;
    const Node2& r = t.nodes[0];
    if (r.leaf()) {
        L.root = r.count ? ilm_leaf_code((uint32_t)r.first, (uint32_t)r.count) : ILM_NO_CHILD;
        return L;
    }
    struct Item { int node; int depth; };
    /* DFS preorder assignment */
    std::vector<int> stack;
    std::vector<int> qid(t.nodes.size(), -1);
    std::vector<std::vector<int>> kids;
    std::vector<int> qsrc;
    std::vector<int> qdepth;
    auto gather = [&](int n) {
        std::vector<int> ch = {t.nodes[n].left, t.nodes[n].right};
        while (ch.size() < 4) {
            int best = -1;
            double ba = -1.0;
            for (size_t i = 0; i < ch.size(); i++) {
                const Node2& c = t.nodes[ch[i]];
                if (!c.leaf() && c.box.half_area() > ba) { ba = c.box.half_area(); best = (int)i; }
            }
            if (best < 0) break;
            int c = ch[best];
            ch.erase(ch.begin() + best);
            ch.insert(ch.begin() + best, t.nodes[c].right);
            ch.insert(ch.begin() + best, t.nodes[c].left);
        }
        return ch;
    };
    /* iterative preorder over the collapsed tree */
    std::vector<std::pair<int, int>> st = {{0, 0}};
    while (!st.empty()) {
        auto [n, d] = st.back();
        st.pop_back();
        int id = (int)qsrc.size();
        qid[n] = id;
        qsrc.push_back(n);
        qdepth.push_back(d);
        L.max_depth = std::max(L.max_depth, d);
        kids.push_back(gather(n));
        const std::vector<int>& ch = kids.back();
        for (int i = (int)ch.size() - 1; i >= 0; i--)
            if (!t.nodes[ch[i]].leaf()) st.push_back({ch[i], d + 1});
    }
    L.nodes.resize(qsrc.size());
    L.exact.resize(qsrc.size() * 4);
    for (size_t i = 0; i < qsrc.size(); i++) {
        const std::vector<int>& ch = kids[i];
        QNode& q = L.nodes[i];
        std::memset(&q, 0, sizeof(q));
        Box cb[4];
        for (size_t k = 0; k < ch.size(); k++) {
            const Node2& c = t.nodes[ch[k]];
            cb[k] = c.box;
            q.child[k] = c.leaf() ? ilm_leaf_code((uint32_t)c.first, (uint32_t)std::max(1, c.count)) : qid[ch[k]];
            L.exact[4 * i + k] = c.box;
        }
        for (size_t k = ch.size(); k < 4; k++) {
            q.child[k] = ILM_NO_CHILD;
            L.exact[4 * i + k].reset();
        }
        q.nchild = (uint8_t)ch.size();
        quantize_node(q, t.nodes[qsrc[i]].box, cb, (int)ch.size());
    }
    L.root = 0;
    return L;
}

}  // namespace ilm
