// mesh_bvh.hpp — host build of the mesh BVH (MeshBvhNode, common/layout.hpp).
#pragma once

#include <cstdint>
#include <vector>

#include "../common/layout.hpp"

namespace rt {

struct MeshBvh {
    std::vector<MeshBvhNode> nodes;    // binary tree, node 0 is the root
    std::vector<MeshBvh4Node> nodes4;  // the same tree with 4-wide nodes (walked form), node 0 is the root
    std::vector<float> recs;           // 12 floats per TriRec, in leaf order
    int depth = 0;                     // deepest leaf of the binary tree (root's children at depth 1)
    int leaves = 0;
};

// Binned-SAH BVH over the triangles of `verts` (9 f64 per triangle), boxes relative to
// `centre`. Returns false, with `out` empty, for an empty mesh or one larger than
// kMeshBvhLeafMax << kMeshBvhDepth triangles.
bool build_mesh_bvh(const std::vector<double>& verts, const double centre[3], MeshBvh& out);

}  // namespace rt
