// kdbuild_gpu.hpp — the device k-d build (kdbuild_gpu.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../host/frame.hpp"
#include "../host/kdbuild.hpp"

namespace rt {

// The host build's arrays (build_flat_kd), built on the device for a mesh whose f64
// vertices are already in device memory (d_tri64): the node array, the triangle records
// and the f64 node bounds land in d_nodes / d_recs / d_nb64; `out` gets the host metadata
// and host copies of the nodes and bounds (not of the records).
void build_flat_kd_device(const std::vector<double>& verts, const double* d_tri64, FlatKd& out, DevBuf& d_nodes,
                          DevBuf& d_recs, DevBuf& d_nb64, hipStream_t st);

constexpr size_t kDeviceKdMinTris = 50000;  // smaller meshes build on the host

}  // namespace rt
