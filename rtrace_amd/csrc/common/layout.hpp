// layout.hpp — device-resident scene layout (HBM) shared by the host flattener and
// the kernels.
//
// Per mesh:
//   nodes   : KdNode[48 B], depth-first pre-order (left child = self + 1). Each node
//             carries the tight AABB of its triangles (the `bounds` every KdNode of
//             the reference holds, src/mesh.rs:41-55) in f32 relative to the mesh
//             centre, rounded outward, and inner nodes keep the reference's f64 split
//             so that every traversal decision (origin side, far-child visit,
//             src/mesh.rs:422-462) is taken in f64 exactly as the reference takes it.
//   recs    : TriRec[48 B] = 3 x float4, one per (leaf, triangle) pair in leaf order
//             (triangles are copied into their leaves, no index indirection):
//             v0 - centre, e1 = v1 - v0, e2 = v2 - v0 (computed in f64, rounded
//             to f32), the triangle index and max(|e1|_1, |e2|_1). The brute-force
//             order (all triangles in file order, src/ray.rs:650-671) is one extra leaf.
//   tri64   : 9 x f64 per triangle (the reference's vertices) for the exact re-test.
#pragma once

#include <cstdint>

namespace rt {

enum : uint32_t { kAxisLeaf = 3u, kFlagMisplaced = 4u };

struct alignas(16) KdNode {
    float bmin[3];  // tight bounds of the node's triangles, relative to the mesh centre
    float bmax[3];
    uint32_t a;     // inner: index of the right child; leaf: first TriRec
    uint32_t b;     // bits 0-1 axis (3 = leaf); inner bit 2 = a child holds a triangle
                    // moved there by the empty-side rule (src/mesh.rs:174-178);
                    // leaf: count << 2 | 3
    double split;   // inner: split position (f64, as built by the reference)
    uint32_t pad[2];
};
static_assert(sizeof(KdNode) == 48, "KdNode is 48 bytes");

constexpr int kMaxKdDepth = 16;  // KdTree::new(.., 16, 10) (src/mesh.rs:1036-1039)
constexpr int kStackSize = kMaxKdDepth + 2;

// Mesh BVH (this build's traversal structure; no reference counterpart): a binary tree
// over the mesh's triangles, each triangle in exactly one leaf, built with a binned
// surface-area heuristic. A node holds both children's boxes (f32 relative to the mesh
// centre, rounded outward) and their links: an inner node's index, or kMeshLeaf |
// first TriRec << 6 | count. Depth at most kMeshBvhDepth, so a walk's stack (one entry
// per level) fits the per-lane stack of the k-d walk; hence at most
// kMeshBvhLeafMax << kMeshBvhDepth triangles.
struct alignas(16) MeshBvhNode {
    float lo0[3], hi0[3];
    float lo1[3], hi1[3];
    uint32_t link[2];
    uint32_t pad[2];
};
static_assert(sizeof(MeshBvhNode) == 64, "MeshBvhNode is 64 bytes");
constexpr uint32_t kMeshLeaf = 0x80000000u;
constexpr int kMeshBvhDepth = kMaxKdDepth;
constexpr int kMeshBvhLeafMax = 8;

// The walked form of the mesh BVH: 4-wide nodes made by merging each binary node with
// its children (each slot a grandchild, or a child that is a leaf), so at most
// kMeshBvhDepth / 2 levels. Boxes per slot in SoA order (f32 relative to the mesh centre,
// rounded outward); an unused slot has link kMeshNone. A walk pushes one stack entry per
// node it leaves with children still to visit (the node and a mask of them), so its
// stack holds at most kMeshBvhDepth / 2 entries.
struct alignas(16) MeshBvh4Node {
    float lox[4], loy[4], loz[4];
    float hix[4], hiy[4], hiz[4];
    uint32_t link[4];
    uint32_t pad[4];
};
static_assert(sizeof(MeshBvh4Node) == 128, "MeshBvh4Node is 128 bytes");
constexpr uint32_t kMeshNone = 0xffffffffu;

// Device view of one mesh.
struct MeshDev {
    const KdNode* nodes;
    const float* recs;  // 12 floats per TriRec
    const MeshBvh4Node* bvh;  // null: the k-d walk for every ray
    const float* bvh_recs;   // TriRecs in BVH leaf order
    const double* tri64;
    const double* nb64;  // per node: f64 tight bounds (min xyz, max xyz) for the tie resolver
    uint32_t root_kd;   // 0
    uint32_t root_bf;   // brute-force leaf
    uint32_t n_tris;
    float radius;              // max |v - centre|_1 (scale of the f32 rounding error)
    double cx, cy, cz;         // centre subtracted from the f32 copies
    float bmin[3], bmax[3];    // bounds relative to the centre, rounded outward
    double bmin64[3], bmax64[3];
};

enum : int32_t { kSphere = 0, kPlane = 1, kCube = 2, kMesh = 3 };

// Object-level BVH over the scene's spheres (scenes with many spheres): 32 B nodes,
// f32 boxes rounded outward. Inner: a = left child, b = right child. Leaf: b has the top
// bit set, a = first entry of the leaf's sphere list, b & 0x7fffffff = count.
struct alignas(16) BvhNode {
    float bmin[3];
    float bmax[3];
    uint32_t a;
    uint32_t b;
};
static_assert(sizeof(BvhNode) == 32, "BvhNode is 32 bytes");
constexpr uint32_t kBvhLeaf = 0x80000000u;

struct ObjDev {
    int32_t kind;
    int32_t mat;
    int32_t mesh;
    int32_t t_identity;  // cube: world_to_local == identity (src/ray.rs:474)
    double c[3];         // sphere/cube centre, plane point
    double r;            // sphere radius
    double n[3];         // plane unit normal
    double half[3];      // cube half size
    double w2l[16];      // cube world_to_local (M^-1), column-major
    double l2w[16];      // cube local_to_world (M)
    double ntf[9];       // cube: (inv(w2l 3x3))^T row-major, applied to normals
    float cf[3];         // sphere centre (f32) for the prefilter
    float rf;            // sphere radius (f32)
    float bmin[3];       // cube: world box of its geometry, rounded outward (prefilter)
    float bmax[3];
    float pc[3];         // plane point (f32) for the prefilter
    float pn[3];         // plane normal (f32) for the prefilter
};

struct MatDev {
    double color[3];
    double ambient, diffuse, specular, shininess;
    int32_t has_refl;
    int32_t texture;  // 0 none, 1 grid, 2 checkerboard
    double refl;
    double line_color[3];
    double line_width, cell_size;
    int32_t mat_b;
    int32_t pad;
};

struct LightDev {
    double pos[3];
    double color[3];
    double intensity;
    double diameter;
    int32_t is_area;
    int32_t shape;    // kLightDisk / kLightSquare
    int32_t samples;  // shadow samples per hit (area lights)
    int32_t cols;     // square: jittered-grid columns for `samples` (largest divisor <= sqrt)
    double size;      // square side
};
enum : int32_t { kLightDisk = 0, kLightSquare = 1 };

struct CameraDev {
    double origin[3], horizontal[3], vertical[3], llc[3], view[3];
    int32_t persp;
    int32_t has_grid;
    double pitch, thick;
    double grid_color[3];
};

}  // namespace rt
