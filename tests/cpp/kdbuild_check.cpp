// Builds the k-d tree of seeded random meshes serially and with the concurrent subtree
// build, and checks that every array is identical. Prints "ok <meshes>" on success.
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "../../rtrace_amd/csrc/host/kdbuild.hpp"

template <class T>
static bool same(const std::vector<T>& a, const std::vector<T>& b) {
    return a.size() == b.size() && (a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(T)) == 0);
}

int main() {
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> u(-1.0, 1.0), small(-0.05, 0.05);
    const int sizes[] = {20000, 33333, 120000};
    int checked = 0;
    for (int n : sizes) {
        for (int variant = 0; variant < 2; ++variant) {
            std::vector<double> v((size_t)n * 9);
            for (int t = 0; t < n; ++t) {
                double c[3] = {u(rng) * 10, u(rng) * 10, u(rng) * (variant ? 0.0 : 10)};  // variant 1: flat in z
                for (int k = 0; k < 9; ++k) v[(size_t)t * 9 + k] = c[k % 3] + small(rng);
            }
            if (variant) v[0] = v[3] = v[6] = 1.0;  // a degenerate (zero-area) triangle
            rt::FlatKd a, b;
            rt::build_flat_kd(v, a, false);
            rt::build_flat_kd(v, b, true);
            if (!same(a.nodes, b.nodes) || !same(a.recs, b.recs) || !same(a.nb64, b.nb64) || a.root_kd != b.root_kd ||
                a.root_bf != b.root_bf || a.leaves != b.leaves || a.max_leaf != b.max_leaf) {
                std::printf("mismatch n=%d variant=%d nodes %zu/%zu recs %zu/%zu\n", n, variant, a.nodes.size(),
                            b.nodes.size(), a.recs.size(), b.recs.size());
                return 1;
            }
            ++checked;
        }
    }
    std::printf("ok %d\n", checked);
    return 0;
}
