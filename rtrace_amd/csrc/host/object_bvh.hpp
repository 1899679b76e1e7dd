// object_bvh.hpp — host build of the object BVH over a scene's spheres (BvhNode,
// common/layout.hpp), walked by obvh_visit with its per-lane LDS stack.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "../common/layout.hpp"

namespace rt {

// A sphere's box (centre +- radius, f32 rounded outward) and f64 centre.
struct SphereBox {
    float lo[3], hi[3];
    double c[3];
    int32_t obj;
};

// obvh_visit pushes at most one stack entry per inner node on the path to a leaf, and its
// stack holds kMaxKdDepth entries: no leaf of the object BVH may sit deeper than that.
constexpr int kObvhMaxDepth = kMaxKdDepth;

// Leaf capacity for n spheres: 4, or more once n exceeds 4 << 14, so that median splits
// reach it within 14 levels (a node at depth d holds at most ceil(n / 2^d) spheres).
size_t obvh_leaf_max(size_t n);

// Median split of the centroids on the widest axis down to leaves of obvh_leaf_max(n)
// spheres; node 0 is the root, `order` receives the sphere object indices in leaf order.
// Returns the depth of the deepest leaf (the root alone is depth 0).
int build_obvh(std::vector<SphereBox>& items, std::vector<BvhNode>& nodes, std::vector<int32_t>& order);

}  // namespace rt
