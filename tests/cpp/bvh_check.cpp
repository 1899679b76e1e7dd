// Builds the mesh BVH of seeded meshes (random, flat with a degenerate triangle, tiny,
// and one with every centroid equal) and checks its invariants: each triangle in exactly
// one leaf, with the record tri_rec writes; leaves of at most kMeshBvhLeafMax triangles
// at depth at most kMeshBvhDepth; every child box (f32, rounded outward) holds the f64
// box of every triangle below it. Prints "ok <meshes>" on success.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "../../rtrace_amd/csrc/host/kdbuild.hpp"
#include "../../rtrace_amd/csrc/host/mesh_bvh.hpp"

static const double* V;
static double C[3];
static const rt::MeshBvh* B;
static std::vector<int> seen;

static bool fail(const char* what) {
    std::printf("FAIL %s\n", what);
    return false;
}

// checks the subtree under `link` against the box (lo, hi); returns false on a violation
static bool walk(uint32_t link, const float* lo, const float* hi, int depth) {
    if (depth > rt::kMeshBvhDepth) return fail("depth");
    if (link & rt::kMeshLeaf) {
        const uint32_t first = (link & ~rt::kMeshLeaf) >> 6, cnt = link & 63u;
        if (cnt > (uint32_t)rt::kMeshBvhLeafMax) return fail("leaf size");
        for (uint32_t i = 0; i < cnt; ++i) {
            const float* r = B->recs.data() + (size_t)(first + i) * 12;
            uint32_t t;
            std::memcpy(&t, r + 9, 4);
            seen.at(t)++;
            float want[12];
            rt::tri_rec(V + (size_t)t * 9, t, C, want);
            if (std::memcmp(r, want, sizeof(want)) != 0) return fail("record");
            for (int v = 0; v < 3; ++v)
                for (int k = 0; k < 3; ++k) {
                    const double x = V[(size_t)t * 9 + v * 3 + k] - C[k];
                    if (!((double)lo[k] <= x && x <= (double)hi[k])) return fail("box");
                }
        }
        return true;
    }
    if (link >= B->nodes.size()) return fail("link");
    const rt::MeshBvhNode& n = B->nodes[link];
    for (int s = 0; s < 2; ++s) {
        const float* clo = s ? n.lo1 : n.lo0;
        const float* chi = s ? n.hi1 : n.hi0;
        if (!walk(n.link[s], clo, chi, depth + 1)) return false;
    }
    return true;
}

static std::vector<int> seen4;

// the 4-wide form: the same leaves, child boxes holding their triangles, depth <= 8
static bool walk4(uint32_t link, const float* lo, const float* hi, int depth) {
    if (link & rt::kMeshLeaf) {
        const uint32_t first = (link & ~rt::kMeshLeaf) >> 6, cnt = link & 63u;
        for (uint32_t i = 0; i < cnt; ++i) {
            uint32_t t;
            std::memcpy(&t, B->recs.data() + (size_t)(first + i) * 12 + 9, 4);
            seen4.at(t)++;
            for (int v = 0; v < 3; ++v)
                for (int k = 0; k < 3; ++k) {
                    const double x = V[(size_t)t * 9 + v * 3 + k] - C[k];
                    if (!((double)lo[k] <= x && x <= (double)hi[k])) return fail("box4");
                }
        }
        return true;
    }
    if (depth >= rt::kMeshBvhDepth / 2) return fail("depth4");
    if (link >= B->nodes4.size()) return fail("link4");
    const rt::MeshBvh4Node& n = B->nodes4[link];
    int used = 0;
    for (int i = 0; i < 4; ++i) {
        if (n.link[i] == rt::kMeshNone) continue;
        ++used;
        const float clo[3] = {n.lox[i], n.loy[i], n.loz[i]}, chi[3] = {n.hix[i], n.hiy[i], n.hiz[i]};
        if (!walk4(n.link[i], clo, chi, depth + 1)) return false;
    }
    return used >= 2 ? true : fail("slots");
}

int main() {
    std::mt19937_64 rng(777);
    std::uniform_real_distribution<double> u(-1.0, 1.0), small(-0.05, 0.05);
    const int sizes[] = {1, 2, 3, 9, 1000, 81920, 200000};
    int checked = 0;
    for (int n : sizes)
        for (int variant = 0; variant < 3; ++variant) {
            std::vector<double> v((size_t)n * 9);
            for (int t = 0; t < n; ++t) {
                double c[3] = {u(rng) * 10, u(rng) * 10, u(rng) * (variant == 1 ? 0.0 : 10)};  // 1: flat in z
                if (variant == 2) c[0] = c[1] = c[2] = 0.5;                                    // 2: equal centroids
                for (int k = 0; k < 9; ++k) v[(size_t)t * 9 + k] = c[k % 3] + (variant == 2 ? 0.0 : small(rng));
                if (variant == 2) {
                    v[(size_t)t * 9 + 0] -= 0.01 * (t % 7);
                    v[(size_t)t * 9 + 3] += 0.01 * (t % 7);
                }
            }
            if (variant == 1) v[0] = v[3] = v[6] = 1.0;  // a degenerate (zero-area) triangle
            const double centre[3] = {0.25, -0.5, 0.125};
            rt::MeshBvh bvh;
            if (!rt::build_mesh_bvh(v, centre, bvh)) return fail("build"), 1;
            V = v.data();
            std::memcpy(C, centre, sizeof(C));
            B = &bvh;
            seen.assign(n, 0);
            if (bvh.nodes.empty()) return fail("root"), 1;
            const rt::MeshBvhNode& r = bvh.nodes[0];
            if (!walk(r.link[0], r.lo0, r.hi0, 1) || !walk(r.link[1], r.lo1, r.hi1, 1)) return 1;
            for (int t = 0; t < n; ++t)
                if (seen[t] != 1) return fail("coverage"), 1;
            seen4.assign(n, 0);
            const float inf[3] = {-INFINITY, -INFINITY, -INFINITY}, sup[3] = {INFINITY, INFINITY, INFINITY};
            if (bvh.nodes4.empty() || !walk4(0, inf, sup, 0)) return 1;
            for (int t = 0; t < n; ++t)
                if (seen4[t] != 1) return fail("coverage4"), 1;
            if (bvh.recs.size() != (size_t)n * 12) return fail("record count"), 1;
            ++checked;
        }
    std::vector<double> none;
    rt::MeshBvh e;
    const double c0[3] = {0, 0, 0};
    if (rt::build_mesh_bvh(none, c0, e) || !e.nodes.empty()) return fail("empty mesh"), 1;
    std::printf("ok %d\n", checked);
    return 0;
}
