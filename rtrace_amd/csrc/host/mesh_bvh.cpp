// mesh_bvh.cpp — binned surface-area-heuristic BVH over a mesh's triangles, the
// traversal structure of the closest-hit and shadow walks (common/layout.hpp).
//
// Each node splits its triangles in two by centroid: the best of 16 bins per axis by
// the SAH cost kTraverse + (A_l n_l + A_r n_r) / A, a leaf when that is not below n (and
// n <= kMeshBvhLeafMax). The depth is capped at kMeshBvhDepth: a node at depth d holds
// at most kMeshBvhLeafMax << (kMeshBvhDepth - d) triangles, and splits that would give a
// child more than its share are not taken (an object-median split is used instead).
#include "mesh_bvh.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

#include "kdbuild.hpp"

namespace rt {
namespace {

constexpr int kBins = 16;
// SAH cost of a node visit, in triangle tests (a node tests two boxes; 3 measured best
// of 1, 2 and 3 on the 1080p icosphere scene, the 64-scene batch and the denoiser scenes)
constexpr double kTraverse = 3.0;

struct Box {
    double lo[3] = {INFINITY, INFINITY, INFINITY};
    double hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const Box& b) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    double area() const {
        const double x = hi[0] - lo[0], y = hi[1] - lo[1], z = hi[2] - lo[2];
        return x < 0 ? 0.0 : 2.0 * (x * y + y * z + z * x);
    }
};

float down(double x) {
    float f = (float)x;
    if ((double)f > x) f = std::nextafter(f, -INFINITY);
    return f;
}
float up(double x) {
    float f = (float)x;
    if ((double)f < x) f = std::nextafter(f, INFINITY);
    return f;
}

struct Builder {
    const double* v;
    double c[3];
    std::vector<Box> tb;       // triangle boxes
    std::vector<double> cen;   // centroids, 3 per triangle
    std::vector<int> idx;
    MeshBvh* out;

    void store(MeshBvhNode& n, int side, const Box& b, uint32_t link) {
        float* lo = side ? n.lo1 : n.lo0;
        float* hi = side ? n.hi1 : n.hi0;
        for (int k = 0; k < 3; ++k) {
            lo[k] = down(b.lo[k] - c[k]);
            hi[k] = up(b.hi[k] - c[k]);
        }
        n.link[side] = link;
    }
    uint32_t leaf(int b, int e, int depth) {
        const uint32_t first = (uint32_t)(out->recs.size() / 12);
        for (int i = b; i < e; ++i) {
            float r[12];
            tri_rec(v + (size_t)idx[i] * 9, (uint32_t)idx[i], c, r);
            out->recs.insert(out->recs.end(), r, r + 12);
        }
        out->leaves++;
        out->depth = std::max(out->depth, depth);
        return kMeshLeaf | first << 6 | (uint32_t)(e - b);
    }
    // builds triangles idx[b, e) at `depth`; returns the link and their box
    uint32_t build(int b, int e, int depth, Box& box) {
        box = Box{};
        Box cb;
        for (int i = b; i < e; ++i) {
            box.grow(tb[idx[i]]);
            for (int k = 0; k < 3; ++k) {
                const double x = cen[(size_t)idx[i] * 3 + k];
                cb.lo[k] = std::min(cb.lo[k], x);
                cb.hi[k] = std::max(cb.hi[k], x);
            }
        }
        const int n = e - b;
        if (n <= 2 || depth >= kMeshBvhDepth) return leaf(b, e, depth);
        const long long cap = (long long)kMeshBvhLeafMax << (kMeshBvhDepth - depth - 1);  // per child
        double best = INFINITY;
        int best_axis = -1, best_bin = 0;
        const double inv_area = 1.0 / std::max(box.area(), 1e-300);
        for (int a = 0; a < 3; ++a) {
            const double ext = cb.hi[a] - cb.lo[a];
            if (!(ext > 0)) continue;
            int cnt[kBins] = {};
            Box bb[kBins];
            for (int i = b; i < e; ++i) {
                const int t = idx[i];
                int k = (int)((cen[(size_t)t * 3 + a] - cb.lo[a]) / ext * kBins);
                k = std::min(std::max(k, 0), kBins - 1);
                cnt[k]++;
                bb[k].grow(tb[t]);
            }
            double ra[kBins];
            int rn[kBins];
            Box acc;
            int an = 0;
            for (int k = kBins - 1; k > 0; --k) {
                acc.grow(bb[k]);
                an += cnt[k];
                ra[k] = acc.area();
                rn[k] = an;
            }
            acc = Box{};
            an = 0;
            for (int k = 1; k < kBins; ++k) {
                acc.grow(bb[k - 1]);
                an += cnt[k - 1];
                const int nl = an, nr = rn[k];
                if (nl == 0 || nr == 0 || nl > cap || nr > cap) continue;
                const double cost = kTraverse + (acc.area() * nl + ra[k] * nr) * inv_area;
                if (cost < best) {
                    best = cost;
                    best_axis = a;
                    best_bin = k;
                }
            }
        }
        if (n <= kMeshBvhLeafMax && !(best < (double)n)) return leaf(b, e, depth);
        int mid;
        if (best_axis >= 0) {
            const int a = best_axis;
            const double lo = cb.lo[a], ext = cb.hi[a] - cb.lo[a];
            mid = (int)(std::partition(idx.begin() + b, idx.begin() + e,
                                       [&](int t) {
                                           int k = (int)((cen[(size_t)t * 3 + a] - lo) / ext * kBins);
                                           k = std::min(std::max(k, 0), kBins - 1);
                                           return k < best_bin;
                                       }) -
                        idx.begin());
        } else {
            // object median on the widest centroid axis (any halves when all centroids coincide)
            int a = 0;
            for (int k = 1; k < 3; ++k)
                if (cb.hi[k] - cb.lo[k] > cb.hi[a] - cb.lo[a]) a = k;
            mid = b + n / 2;
            std::nth_element(idx.begin() + b, idx.begin() + mid, idx.begin() + e,
                             [&](int x, int y) { return cen[(size_t)x * 3 + a] < cen[(size_t)y * 3 + a]; });
        }
        const uint32_t me = (uint32_t)out->nodes.size();
        out->nodes.push_back(MeshBvhNode{});
        Box bl, br;
        const uint32_t l = build(b, mid, depth + 1, bl);
        const uint32_t r = build(mid, e, depth + 1, br);
        store(out->nodes[me], 0, bl, l);
        store(out->nodes[me], 1, br, r);
        return me;
    }
};

// 4-wide node of binary node b: its children, each replaced by its own two children
// when it is an inner node
uint32_t collapse(const std::vector<MeshBvhNode>& bin, uint32_t b, std::vector<MeshBvh4Node>& out) {
    const uint32_t me = (uint32_t)out.size();
    out.push_back(MeshBvh4Node{});
    struct Slot {
        uint32_t link;
        const float *lo, *hi;
    };
    Slot slots[4];
    int n = 0;
    for (int side = 0; side < 2; ++side) {
        const MeshBvhNode& nb = bin[b];
        const uint32_t l = nb.link[side];
        if (l & kMeshLeaf) {
            slots[n++] = Slot{l, side ? nb.lo1 : nb.lo0, side ? nb.hi1 : nb.hi0};
        } else {
            const MeshBvhNode& nc = bin[l];
            slots[n++] = Slot{nc.link[0], nc.lo0, nc.hi0};
            slots[n++] = Slot{nc.link[1], nc.lo1, nc.hi1};
        }
    }
    uint32_t links[4] = {kMeshNone, kMeshNone, kMeshNone, kMeshNone};
    for (int i = 0; i < n; ++i) links[i] = slots[i].link & kMeshLeaf ? slots[i].link : collapse(bin, slots[i].link, out);
    MeshBvh4Node& q = out[me];
    for (int i = 0; i < 4; ++i) {
        const bool used = i < n;
        q.lox[i] = used ? slots[i].lo[0] : 0.0f;
        q.loy[i] = used ? slots[i].lo[1] : 0.0f;
        q.loz[i] = used ? slots[i].lo[2] : 0.0f;
        q.hix[i] = used ? slots[i].hi[0] : 0.0f;
        q.hiy[i] = used ? slots[i].hi[1] : 0.0f;
        q.hiz[i] = used ? slots[i].hi[2] : 0.0f;
        q.link[i] = links[i];
    }
    return me;
}

}  // namespace

bool build_mesh_bvh(const std::vector<double>& verts, const double centre[3], MeshBvh& out) {
    out = MeshBvh{};
    const size_t n = verts.size() / 9;
    if (n == 0 || n > ((size_t)kMeshBvhLeafMax << kMeshBvhDepth)) return false;
    Builder B;
    B.v = verts.data();
    for (int k = 0; k < 3; ++k) B.c[k] = centre[k];
    B.tb.resize(n);
    B.cen.resize(n * 3);
    B.idx.resize(n);
    B.out = &out;
    for (size_t t = 0; t < n; ++t) {
        const double* p = B.v + t * 9;
        Box& b = B.tb[t];
        for (int k = 0; k < 3; ++k) {
            b.lo[k] = std::min(std::min(p[k], p[3 + k]), p[6 + k]);
            b.hi[k] = std::max(std::max(p[k], p[3 + k]), p[6 + k]);
            B.cen[t * 3 + k] = (p[k] + p[3 + k] + p[6 + k]) / 3.0;
        }
        B.idx[t] = (int)t;
    }
    // the root is node 0; a mesh that is one leaf gets a root whose second child is an
    // empty leaf
    Box box;
    const uint32_t root = B.build(0, (int)n, 0, box);
    if (root & kMeshLeaf) {
        out.nodes.insert(out.nodes.begin(), MeshBvhNode{});
        B.store(out.nodes[0], 0, box, root);
        B.store(out.nodes[0], 1, box, kMeshLeaf);
    }
    collapse(out.nodes, 0, out.nodes4);
    return true;
}

}  // namespace rt
