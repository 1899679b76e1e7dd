// kdbuild.hpp — host k-d build (reference algorithm) into the device layout.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../common/layout.hpp"

namespace rt {

struct FlatKd {
    std::vector<KdNode> nodes;
    std::vector<float> recs;  // 12 floats per TriRec
    std::vector<double> nb64; // 6 doubles per node: f64 tight bounds
    uint32_t root_kd = 0;
    uint32_t root_bf = 0;
    double center[3] = {0, 0, 0};
    double bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};
    int leaves = 0;
    int max_leaf = 0;
};

// TriRec of triangle t (vertices p[0..8], f64) relative to centre c: v0 - c, e1, e2 in
// f32, the index, and E = max(|e1|_1, |e2|_1), the scale of the f32 edge rounding.
inline void tri_rec(const double* p, uint32_t t, const double c[3], float r[12]) {
    r[0] = (float)(p[0] - c[0]);
    r[1] = (float)(p[1] - c[1]);
    r[2] = (float)(p[2] - c[2]);
    r[3] = (float)(p[3] - p[0]);
    r[4] = (float)(p[4] - p[1]);
    r[5] = (float)(p[5] - p[2]);
    r[6] = (float)(p[6] - p[0]);
    r[7] = (float)(p[7] - p[1]);
    r[8] = (float)(p[8] - p[2]);
    std::memcpy(&r[9], &t, 4);
    const double n1 = std::fabs(p[3] - p[0]) + std::fabs(p[4] - p[1]) + std::fabs(p[5] - p[2]);
    const double n2 = std::fabs(p[6] - p[0]) + std::fabs(p[7] - p[1]) + std::fabs(p[8] - p[2]);
    r[10] = (float)(std::max(n1, n2) * (1.0 + 1e-6));
    r[11] = 0.0f;
}

// parallel: build the subtrees below the top levels on their own threads (meshes of
// 20k triangles or more); the result is the same as the serial build, array for array.
void build_flat_kd(const std::vector<double>& verts, FlatKd& out, bool parallel = true);


}  // namespace rt
