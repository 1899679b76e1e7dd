// wave_ops.hpp — wavefront-level helpers shared by the trace kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "trace_core.hpp"

namespace rt {

// One atomic per wavefront; returns this lane's slot (or -1 when !pred).
__device__ __forceinline__ int wave_append(int32_t* counter, bool pred) {
    unsigned long long mask = __ballot(pred);
    if (mask == 0) return -1;
    int lane = __lane_id();
    int leader = __ffsll((long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, (int)__popcll(mask));
    base = __shfl(base, leader);
    int rank = (int)__popcll(mask & ((1ull << lane) - 1ull));
    return pred ? base + rank : -1;
}

__device__ __forceinline__ void flush_work(const FrameView& F, const Work& w, int kernel) {
    if (!F.count_work) return;
#pragma unroll
    for (int k = 0; k < kWorkCount; ++k)
        if (w.v[k]) atomicAdd(&F.work[kernel * kWorkCount + k], w.v[k]);
}

}  // namespace rt
