// object_bvh.cpp — object BVH over a scene's spheres (object_bvh.hpp).
#include "object_bvh.hpp"

#include <algorithm>
#include <cmath>

namespace rt {
namespace {

uint32_t build(std::vector<SphereBox>& items, size_t begin, size_t end, size_t leaf_max, int depth, int& deepest,
               std::vector<BvhNode>& nodes, std::vector<int32_t>& order) {
    const uint32_t me = (uint32_t)nodes.size();
    nodes.push_back(BvhNode{});
    BvhNode nd{};
    for (int k = 0; k < 3; ++k) {
        nd.bmin[k] = INFINITY;
        nd.bmax[k] = -INFINITY;
    }
    double cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = begin; i < end; ++i)
        for (int k = 0; k < 3; ++k) {
            nd.bmin[k] = std::min(nd.bmin[k], items[i].lo[k]);
            nd.bmax[k] = std::max(nd.bmax[k], items[i].hi[k]);
            cmin[k] = std::min(cmin[k], items[i].c[k]);
            cmax[k] = std::max(cmax[k], items[i].c[k]);
        }
    if (end - begin <= leaf_max || depth >= kObvhMaxDepth) {
        nd.a = (uint32_t)order.size();
        nd.b = kBvhLeaf | (uint32_t)(end - begin);
        for (size_t i = begin; i < end; ++i) order.push_back(items[i].obj);
        nodes[me] = nd;
        deepest = std::max(deepest, depth);
        return me;
    }
    int axis = 0;
    for (int k = 1; k < 3; ++k)
        if (cmax[k] - cmin[k] > cmax[axis] - cmin[axis]) axis = k;
    const size_t mid = (begin + end) / 2;
    std::nth_element(items.begin() + begin, items.begin() + mid, items.begin() + end,
                     [axis](const SphereBox& x, const SphereBox& y) { return x.c[axis] < y.c[axis]; });
    nd.a = build(items, begin, mid, leaf_max, depth + 1, deepest, nodes, order);
    nd.b = build(items, mid, end, leaf_max, depth + 1, deepest, nodes, order);
    nodes[me] = nd;
    return me;
}

}  // namespace

size_t obvh_leaf_max(size_t n) {
    const size_t per_leaf = (n + (1u << 14) - 1) >> 14;  // ceil(n / 2^14)
    return std::max<size_t>(4, per_leaf);
}

int build_obvh(std::vector<SphereBox>& items, std::vector<BvhNode>& nodes, std::vector<int32_t>& order) {
    nodes.clear();
    order.clear();
    int deepest = 0;
    if (items.empty()) return 0;
    build(items, 0, items.size(), obvh_leaf_max(items.size()), 0, deepest, nodes, order);
    return deepest;
}

}  // namespace rt
