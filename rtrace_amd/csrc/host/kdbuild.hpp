// kdbuild.hpp — host k-d build (reference algorithm) into the device layout.
#pragma once

#include <cstdint>
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

// parallel: build the subtrees below the top levels on their own threads (meshes of
// 20k triangles or more); the result is the same as the serial build, array for array.
void build_flat_kd(const std::vector<double>& verts, FlatKd& out, bool parallel = true);


}  // namespace rt
