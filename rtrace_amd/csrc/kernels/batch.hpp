// batch.hpp — launch interface of the persistent multi-job kernels (batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "trace_core.hpp"

namespace rt {

// Device-side description of a batch: per job one scene view (the job's scene with the
// job's camera) and two frame views (even / odd levels: the ray queues swap roles), plus
// the work plan of the current stage.
struct BatchArgs {
    const SceneView* views;   // [n_jobs]
    const FrameView* frames;  // [2 * n_jobs]
    int32_t* base;            // [n_jobs + 1] first chunk of each job in the current stage
    int32_t* counter;         // [kBatchSlices * 32]: chunks handed out of each slice, 128 bytes apart
    int32_t n_jobs;
    int32_t chunk;            // items per chunk (a multiple of the block size)
    unsigned long long* tally;  // [n_jobs][2]: reflection and shadow rays, written at the end
};

struct BatchLaunch {
    int grid;       // persistent workgroups per stage kernel
    int max_depth;  // deepest job
};

// One group of jobs' launch sequence on `st` (capturable into a hipGraph: kernels only);
// its last kernel tallies each job's rays, since the next group reuses the frame buffers.
hipError_t launch_batch(const BatchArgs& a, const BatchLaunch& l, hipStream_t st);

constexpr int kBatchBlock = 64;  // one wavefront per workgroup
constexpr int kBatchSlices = 8;  // chunk-range slices of a stage launch, one per XCD

}  // namespace rt
