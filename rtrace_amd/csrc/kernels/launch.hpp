// launch.hpp — host entry points into the trace kernels (trace.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "trace_core.hpp"

#include <vector>

namespace rt {

// Optional per-stage hipEvent pairs around each launch group of a frame.
struct StageTimer {
    std::vector<hipEvent_t> ev;   // pool
    std::vector<int> stage;       // stage of each (begin, end) pair
    size_t used = 0;
    hipError_t mark(int stage_id, bool begin, hipStream_t st) {
        if (begin) {
            if (used + 2 > ev.size()) {
                for (int i = 0; i < 2; ++i) {
                    hipEvent_t e;
                    hipError_t rc = hipEventCreate(&e);
                    if (rc != hipSuccess) return rc;
                    ev.push_back(e);
                }
            }
            if (stage.size() < used / 2 + 1) stage.resize(used / 2 + 1);
            stage[used / 2] = stage_id;
            return hipEventRecord(ev[used], st);
        }
        hipError_t rc = hipEventRecord(ev[used + 1], st);
        used += 2;
        return rc;
    }
    void reset() { used = 0; }
    ~StageTimer() {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
};

// Enqueue one full frame on `st` (no host synchronisation inside).
hipError_t launch_frame(const SceneView& S, FrameView F, RayRec* q0, RayRec* q1, hipStream_t st, StageTimer* tm);
// Runs the host-side queries launch_frame makes (cached), so a capture of it holds only kernels.
void warm_frame_launches();

// Closest hit of n caller rays (device pointers).
hipError_t launch_intersect(const SceneView& S, const double* o, const double* d, int n, double tmin, double tmax,
                            int use_kd, int32_t* obj, int32_t* tri, double* t, hipStream_t st);

}  // namespace rt
