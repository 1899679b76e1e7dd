// kdbuild.cpp — host k-d build with the reference algorithm, emitted directly in the
// flat device layout (common/layout.hpp).
//
// KdTree::new / build_recursive (src/mesh.rs:75-197): leaf when depth >= 16 or <= 10
// triangles; axis = depth % 3; stable sort by centroid[axis] (centroid = ((v0+v1)+v2)/3);
// split = centroid of element n/2; a triangle goes left when min[axis] <= split and
// right when max[axis] > split (in sorted order); an empty side takes the other side's
// last triangle; children recurse left then right. Nodes are written in that same
// depth-first pre-order, so the left child of node i is i + 1.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <future>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <thread>

#include "../common/layout.hpp"
#include "../common/vec.hpp"
#include "kdbuild.hpp"

namespace rt {

namespace {

// One subtree's arrays in pre-order, with node and record indices local to it.
struct Part {
    std::vector<KdNode> nodes;
    std::vector<float> recs;
    std::vector<double> nb64;
    int leaves = 0;
    int max_leaf = 0;
};

// A node of the top of the tree (above kParDepth), or a subtree built on its own thread.
struct Top {
    KdNode node{};
    double nb[6] = {0, 0, 0, 0, 0, 0};
    std::unique_ptr<Top> left, right;
    std::future<Part> sub;  // valid for a subtree root
};

constexpr int kParDepth = 4;          // subtrees below this depth build concurrently
constexpr size_t kParMinTris = 20000; // smaller meshes build on the calling thread

struct Builder {
    const double* v;  // 9 per triangle
    double cx, cy, cz;
    Part* out;

    double centroid(int t, int axis) const {
        const double* p = v + (size_t)t * 9;
        double s = p[axis] + p[3 + axis];
        s = s + p[6 + axis];
        return s / 3.0;
    }
    double tmin(int t, int axis) const {
        const double* p = v + (size_t)t * 9;
        return fmin_r(fmin_r(p[axis], p[3 + axis]), p[6 + axis]);
    }
    double tmax(int t, int axis) const {
        const double* p = v + (size_t)t * 9;
        return fmax_r(fmax_r(p[axis], p[3 + axis]), p[6 + axis]);
    }
    // tight bounds of a triangle list (compute_bounds, src/mesh.rs:200-215), stored in
    // f32 relative to the centre and rounded outward so the f32 box contains the f64 one
    void set_bounds(KdNode& n, const std::vector<int>& tris) const {
        double nb[6];
        bounds(n, tris, nb);
        out->nb64.insert(out->nb64.end(), nb, nb + 6);
    }
    void bounds(KdNode& n, const std::vector<int>& tris, double* nb) const {
        double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int t : tris)
            for (int k = 0; k < 3; ++k) {
                mn[k] = fmin_r(mn[k], tmin(t, k));
                mx[k] = fmax_r(mx[k], tmax(t, k));
            }
        const double c[3] = {cx, cy, cz};
        if (tris.empty())
            for (int k = 0; k < 3; ++k) mn[k] = mx[k] = 0.0;  // compute_bounds of an empty list
        for (int k = 0; k < 3; ++k) {  // a zero bound is stored without sign (min / max of +-0 is order-dependent)
            if (mn[k] == 0.0) mn[k] = 0.0;
            if (mx[k] == 0.0) mx[k] = 0.0;
        }
        for (int k = 0; k < 3; ++k) {
            nb[k] = mn[k];
            nb[3 + k] = mx[k];
        }
        for (int k = 0; k < 3; ++k) {
            double lo = mn[k] - c[k], hi = mx[k] - c[k];
            float flo = (float)lo, fhi = (float)hi;
            if ((double)flo > lo) flo = std::nextafter(flo, -INFINITY);
            if ((double)fhi < hi) fhi = std::nextafter(fhi, INFINITY);
            n.bmin[k] = flo;
            n.bmax[k] = fhi;
        }
    }
    void emit_leaf(const std::vector<int>& tris) {
        KdNode n{};
        set_bounds(n, tris);
        n.split = 0.0;
        n.a = (uint32_t)(out->recs.size() / 12);
        n.b = ((uint32_t)tris.size() << 2) | kAxisLeaf;
        out->nodes.push_back(n);
        for (int t : tris) push_rec(t);
        out->max_leaf = std::max<int>(out->max_leaf, (int)tris.size());
        out->leaves++;
    }
    void push_rec(int t) {
        const double c[3] = {cx, cy, cz};
        float r[12];
        tri_rec(v + (size_t)t * 9, (uint32_t)t, c, r);
        out->recs.insert(out->recs.end(), r, r + 12);
    }
    // build_recursive's split of an inner node: the axis, the split value, the two
    // triangle lists and the empty-side flag.
    void split_node(const std::vector<int>& tris, int depth, double& split, std::vector<int>& L, std::vector<int>& R,
                    uint32_t& flags) const {
        int axis = depth % 3;
        std::vector<std::pair<double, int>> keyed(tris.size());
        for (size_t i = 0; i < tris.size(); ++i) {
            double c = centroid(tris[i], axis);
            if (std::isnan(c)) throw std::runtime_error("k-d build: NaN centroid (the reference panics here)");
            keyed[i] = {c, tris[i]};
        }
        std::stable_sort(keyed.begin(), keyed.end(),
                         [](const std::pair<double, int>& a, const std::pair<double, int>& b) { return a.first < b.first; });
        split = keyed[keyed.size() / 2].first;
        L.reserve(keyed.size());
        R.reserve(keyed.size());
        for (auto& k : keyed) {
            if (tmin(k.second, axis) <= split) L.push_back(k.second);
            if (tmax(k.second, axis) > split) R.push_back(k.second);
        }
        flags = 0;
        if (L.empty()) {
            L.push_back(R.back());
            R.pop_back();
            flags = kFlagMisplaced;
        } else if (R.empty()) {
            R.push_back(L.back());
            L.pop_back();
            flags = kFlagMisplaced;
        }
    }
    static bool is_leaf(const std::vector<int>& tris, int depth) { return depth >= kMaxKdDepth || tris.size() <= 10; }
    void build(std::vector<int> tris, int depth) {
        if (is_leaf(tris, depth)) {
            emit_leaf(tris);
            return;
        }
        double split;
        std::vector<int> L, R;
        uint32_t flags;
        split_node(tris, depth, split, L, R, flags);
        size_t me = out->nodes.size();
        KdNode n{};
        set_bounds(n, tris);
        n.split = split;
        n.b = (uint32_t)(depth % 3) | flags;
        out->nodes.push_back(n);
        tris = std::vector<int>();  // free before recursing
        build(std::move(L), depth + 1);
        out->nodes[me].a = (uint32_t)out->nodes.size();
        build(std::move(R), depth + 1);
    }
};

Part build_part(const double* v, const double* c, std::vector<int> tris, int depth) {
    Part p;
    Builder b{v, c[0], c[1], c[2], &p};
    b.build(std::move(tris), depth);
    return p;
}

// The top kParDepth levels on this thread; each subtree below them on its own task.
std::unique_ptr<Top> plan_top(const double* v, const double* c, std::vector<int> tris, int depth) {
    auto t = std::make_unique<Top>();
    if (depth >= kParDepth || Builder::is_leaf(tris, depth)) {
        t->sub = std::async(std::launch::async, build_part, v, c, std::move(tris), depth);
        return t;
    }
    Part scratch;
    Builder b{v, c[0], c[1], c[2], &scratch};
    double split;
    std::vector<int> L, R;
    uint32_t flags;
    b.split_node(tris, depth, split, L, R, flags);
    b.bounds(t->node, tris, t->nb);
    t->node.split = split;
    t->node.b = (uint32_t)(depth % 3) | flags;
    tris = std::vector<int>();
    t->left = plan_top(v, c, std::move(L), depth + 1);
    t->right = plan_top(v, c, std::move(R), depth + 1);
    return t;
}

// Appends a subtree's arrays, rebasing its local node and record indices.
void append_part(Part& dst, const Part& p) {
    const uint32_t node0 = (uint32_t)dst.nodes.size(), rec0 = (uint32_t)(dst.recs.size() / 12);
    for (KdNode n : p.nodes) {
        n.a += (n.b & 3u) == kAxisLeaf ? rec0 : node0;
        dst.nodes.push_back(n);
    }
    dst.recs.insert(dst.recs.end(), p.recs.begin(), p.recs.end());
    dst.nb64.insert(dst.nb64.end(), p.nb64.begin(), p.nb64.end());
    dst.leaves += p.leaves;
    dst.max_leaf = std::max(dst.max_leaf, p.max_leaf);
}

// Writes the planned tree in the same depth-first pre-order the serial build produces.
void merge_top(Part& dst, Top& t) {
    if (t.sub.valid()) {
        append_part(dst, t.sub.get());
        return;
    }
    const size_t me = dst.nodes.size();
    dst.nodes.push_back(t.node);
    dst.nb64.insert(dst.nb64.end(), t.nb, t.nb + 6);
    merge_top(dst, *t.left);
    dst.nodes[me].a = (uint32_t)dst.nodes.size();
    merge_top(dst, *t.right);
}

}  // namespace

void build_flat_kd(const std::vector<double>& verts, FlatKd& out, bool parallel) {
    size_t n = verts.size() / 9;
    out = FlatKd{};
    // bounds (Mesh::compute_bounds, src/mesh.rs:1017-1033) and the centre used for
    // the f32 copies
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = 0; i < n * 3; ++i)
        for (int k = 0; k < 3; ++k) {
            mn[k] = fmin_r(mn[k], verts[i * 3 + k]);
            mx[k] = fmax_r(mx[k], verts[i * 3 + k]);
        }
    if (n == 0) {
        for (int k = 0; k < 3; ++k) mn[k] = mx[k] = 0.0;
    }
    for (int k = 0; k < 3; ++k) {
        out.bmin[k] = mn[k];
        out.bmax[k] = mx[k];
        out.center[k] = 0.5 * (mn[k] + mx[k]);
    }
    Part tree;
    Builder b{verts.data(), out.center[0], out.center[1], out.center[2], &tree};
    if (n > 0) {
        std::vector<int> all(n);
        std::iota(all.begin(), all.end(), 0);
        if (parallel && n >= kParMinTris && std::thread::hardware_concurrency() > 1) {
            // same tree, same order: the subtrees are independent and splice back in pre-order
            std::unique_ptr<Top> top = plan_top(verts.data(), out.center, std::move(all), 0);
            merge_top(tree, *top);
        } else {
            b.build(std::move(all), 0);
        }
    }
    // brute-force order: one leaf over every triangle in file order (src/ray.rs:650-671)
    out.root_bf = (uint32_t)tree.nodes.size();
    KdNode bf{};
    {
        std::vector<int> all(n);
        std::iota(all.begin(), all.end(), 0);
        b.set_bounds(bf, all);
    }
    bf.a = (uint32_t)(tree.recs.size() / 12);
    bf.b = ((uint32_t)n << 2) | kAxisLeaf;
    tree.nodes.push_back(bf);
    for (size_t t = 0; t < n; ++t) b.push_rec((int)t);
    out.nodes = std::move(tree.nodes);
    out.recs = std::move(tree.recs);
    out.nb64 = std::move(tree.nb64);
    out.leaves = tree.leaves;
    out.max_leaf = tree.max_leaf;
    if (n == 0) {
        // an empty mesh has no k-d root (src/mesh.rs:82): point it at the empty leaf
        out.root_kd = out.root_bf;
    }
}

}  // namespace rt
