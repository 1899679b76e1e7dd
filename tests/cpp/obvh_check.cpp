// Builds the object BVH over seeded sphere sets (1 to 1,048,576 spheres, including ones
// with every centre equal) and checks: every sphere in exactly one leaf; leaves hold at
// most obvh_leaf_max(n) spheres and sit at depth <= kObvhMaxDepth (the traversal stack's
// size), as build_obvh reports; every node box holds the boxes below it. Prints
// "ok <sets>" on success.
#include <cstdio>
#include <random>
#include <vector>

#include "../../rtrace_amd/csrc/host/object_bvh.hpp"

static std::vector<rt::BvhNode> N;
static std::vector<int32_t> O;
static std::vector<rt::SphereBox> items_by_obj;
static std::vector<int> seen;
static size_t leaf_max;

static bool fail(const char* what) {
    std::printf("FAIL %s\n", what);
    return false;
}

static bool inside(const float* lo, const float* hi, const float* plo, const float* phi) {
    for (int k = 0; k < 3; ++k)
        if (lo[k] < plo[k] || hi[k] > phi[k]) return false;
    return true;
}

static bool walk(uint32_t node, int depth, int& deepest) {
    if (depth > rt::kObvhMaxDepth) return fail("depth");
    const rt::BvhNode& nd = N.at(node);
    if (nd.b & rt::kBvhLeaf) {
        const uint32_t cnt = nd.b & ~rt::kBvhLeaf;
        if (cnt == 0 || cnt > leaf_max) return fail("leaf size");
        deepest = depth > deepest ? depth : deepest;
        for (uint32_t i = 0; i < cnt; ++i) {
            const int32_t obj = O.at(nd.a + i);
            seen.at(obj)++;
            const rt::SphereBox& b = items_by_obj.at(obj);
            if (!inside(b.lo, b.hi, nd.bmin, nd.bmax)) return fail("leaf box");
        }
        return true;
    }
    for (uint32_t c : {nd.a, nd.b})
        if (!inside(N.at(c).bmin, N.at(c).bmax, nd.bmin, nd.bmax)) return fail("child box");
    return walk(nd.a, depth + 1, deepest) && walk(nd.b, depth + 1, deepest);
}

int main() {
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> U(-20.0, 20.0), R(0.1, 0.6);
    const size_t sizes[] = {1, 4, 5, 16, 1000, 65536, 65537, 300000, 1048576};
    int sets = 0;
    for (int equal = 0; equal < 2; ++equal)
        for (size_t n : sizes) {
            if (equal && n > 70000) continue;
            std::vector<rt::SphereBox> items(n);
            for (size_t i = 0; i < n; ++i) {
                const double c[3] = {equal ? 1.0 : U(rng), equal ? 2.0 : U(rng), equal ? 3.0 : U(rng)};
                const double r = R(rng);
                for (int k = 0; k < 3; ++k) {
                    items[i].c[k] = c[k];
                    items[i].lo[k] = (float)(c[k] - r) - 1e-5f;
                    items[i].hi[k] = (float)(c[k] + r) + 1e-5f;
                }
                items[i].obj = (int32_t)i;
            }
            items_by_obj = items;
            leaf_max = rt::obvh_leaf_max(n);
            const int reported = rt::build_obvh(items, N, O);
            seen.assign(n, 0);
            int deepest = 0;
            if (!walk(0, 0, deepest)) return 1;
            for (size_t i = 0; i < n; ++i)
                if (seen[i] != 1) return fail("sphere not in exactly one leaf"), 1;
            if (deepest != reported || reported > rt::kObvhMaxDepth) return fail("reported depth"), 1;
            if (n == 300000 && deepest > 16) return fail("300k depth"), 1;
            ++sets;
        }
    std::printf("ok %d\n", sets);
    return 0;
}
