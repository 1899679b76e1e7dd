// batch.hip — persistent kernels that run the trace pipeline of many frames (scenes,
// camera views, resolutions) together.
//
// Every stage of the single-frame pipeline (raygen, closest, resolve, shade, shadow,
// finish, pixel, finalize) becomes one launch for the whole batch. A plan kernel reads
// each job's device-side item count for the stage and lays the jobs' chunks end to end;
// a persistent grid then takes the chunks, each XCD from its own slice of them first
// (ChunkQueue). A chunk never spans two
// jobs, so the scene and frame views a workgroup reads are uniform (scalar loads), and
// jobs of very different cost share the machine without a host round trip. The item
// functions are the single-frame ones (trace_core.hpp), so a batched frame is the same
// image as a frame rendered alone.
#include <hip/hip_runtime.h>

#include "batch.hpp"
#include "trace_core.hpp"
#include "wave_ops.hpp"

namespace rt {
namespace {

enum Stage : int {
    kStRaygen = 0,
    kStClosest,
    kStResolve,
    kStShadePrep,
    kStShadow,
    kStFinish,
    kStPixel,
    kStFinalize
};

__device__ __forceinline__ long long stage_items(int stage, const SceneView& S, const FrameView& F, int level) {
    switch (stage) {
        case kStRaygen: return F.n_q0;
        case kStPixel:
        case kStFinalize: return (long long)F.W * F.H;
        default: break;
    }
    if (level >= F.max_depth) return 0;  // this job has no such level
    switch (stage) {
        case kStClosest: return F.counts[level];
        case kStResolve: return F.tie_count[level];
        case kStShadePrep:
        case kStFinish: return F.hit_count[level];
        case kStShadow: return S.n_lights > 0 ? (long long)F.hit_count[level] * S.shadow_per_hit : 0;
        default: return 0;
    }
}

// One workgroup: chunk counts of every job, exclusive scan into base[], total in
// base[n_jobs], counter reset.
__global__ __launch_bounds__(1024) void k_plan(BatchArgs a, int stage, int level) {
    __shared__ int32_t s_part[1024];
    const int per = (a.n_jobs + 1023) / 1024;
    const int j0 = (int)threadIdx.x * per, j1 = min(j0 + per, a.n_jobs);
    int32_t local = 0;
    for (int j = j0; j < j1; ++j) {
        const long long n = stage_items(stage, a.views[j], a.frames[2 * j], level);
        local += (int32_t)((n + a.chunk - 1) / a.chunk);
    }
    s_part[threadIdx.x] = local;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int32_t v = (int)threadIdx.x >= off ? s_part[threadIdx.x - off] : 0;
        __syncthreads();
        s_part[threadIdx.x] += v;
        __syncthreads();
    }
    int32_t run = s_part[threadIdx.x] - local;
    for (int j = j0; j < j1; ++j) {
        a.base[j] = run;
        const long long n = stage_items(stage, a.views[j], a.frames[2 * j], level);
        run += (int32_t)((n + a.chunk - 1) / a.chunk);
    }
    if (threadIdx.x == 1023) a.base[a.n_jobs] = s_part[1023];
    if (threadIdx.x < kBatchSlices) a.counter[threadIdx.x * 32] = 0;
}

__global__ void k_batch_init(BatchArgs a) {
    const FrameView& F = a.frames[2 * blockIdx.x];
    for (int i = threadIdx.x; i <= F.max_depth; i += blockDim.x) {
        F.counts[i] = i == 0 ? F.n_q0 : 0;
        if (i < F.max_depth) {
            F.hit_count[i] = 0;
            F.tie_count[i] = 0;
        }
    }
}

// Rays each job traced (reflection, shadow), read from its queue counts before another
// group of jobs reuses the frame buffers.
__global__ void k_batch_tally(BatchArgs a) {
    const int j = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (j >= a.n_jobs) return;
    const SceneView& S = a.views[j];
    const FrameView& F = a.frames[2 * j];
    unsigned long long refl = 0, shadow = 0;
    const unsigned long long k = S.n_lights > 0 ? (unsigned long long)S.shadow_per_hit : 0ull;
    for (int l = 0; l < F.max_depth; ++l) {
        if (l > 0) refl += (unsigned long long)F.counts[l];
        shadow += (unsigned long long)F.hit_count[l] * k;
    }
    a.tally[2 * j] = refl;
    a.tally[2 * j + 1] = shadow;
}

// The scene and frame views reach the stage kernels as separate read-only, non-aliased
// kernel arguments so that their wave-uniform fields load through the scalar cache. They
// are in the constant address space (nothing writes them while the batch runs): the compiler
// then takes the buffer pointers read from them as global, as it does for a single frame's
// views passed by value, and the item functions' loads and stores through them are global
// accesses rather than flat ones.
#define RT_AS_CONST __attribute__((address_space(4)))
#define BATCH_PARAMS                                                                        \
    const RT_AS_CONST SceneView *__restrict__ views, const RT_AS_CONST FrameView *__restrict__ frames, \
        const int32_t *__restrict__ base, int32_t *counter, int n_jobs, int chunk, int level

// The chunks of one stage launch, taken one at a time by a one-wavefront workgroup (lane 0's
// atomic, no barrier). The chunk range is cut into kBatchSlices contiguous slices, one per
// XCD (workgroups go to the XCDs round-robin by id), each with its own counter: an XCD works
// through its own part of the jobs, which its L2 then holds, and moves on to the other slices
// once its own is drained. The search for a chunk's job starts from the workgroup's previous
// job (usually the same one). Taking the next chunk before working on the current one, to hide
// the atomic, measured 5 % slower on the 64-scene batch: each wavefront then holds a chunk it
// has not started, and the stage's last chunks wait behind busy wavefronts.
static_assert(kBatchBlock == 64, "a chunk queue belongs to one wavefront");
struct ChunkQueue {
    const int32_t* __restrict__ base;  // [n_jobs + 1]: first chunk of each job, total last
    int32_t* counter;                   // [kBatchSlices * 32]
    int n_jobs, chunk, total, job, slice, left;
    __device__ ChunkQueue(const int32_t* __restrict__ b, int32_t* c, int nj, int ch)
        : base(b), counter(c), n_jobs(nj), chunk(ch), job(0), slice((int)(blockIdx.x % kBatchSlices)),
          left(kBatchSlices) {
        total = base[n_jobs];
    }
    // the next chunk of this workgroup's slice, else of the next slices; total when none is left
    __device__ int take() {
        while (left > 0) {
            const int lo = (int)((long long)total * slice / kBatchSlices);
            const int hi = (int)((long long)total * (slice + 1) / kBatchSlices);
            if (lo < hi) {
                int t = 0;
                if (__lane_id() == 0) t = atomicAdd(&counter[slice * 32], 1);
                t = __builtin_amdgcn_readfirstlane(t);
                if (t < hi - lo) return lo + t;
            }
            slice = (slice + 1) % kBatchSlices;
            --left;
        }
        return total;
    }
    // false when no chunk is left; else the chunk's job and its first item (wave-uniform)
    __device__ bool pop(int& j, long long& first) {
        const int c = take();
        if (c >= total) return false;
        if (c < base[job] || base[job + 1] <= c) {  // another job: the last one whose first chunk is <= c
            int lo = 0, hi = n_jobs;  // base[lo] <= c < base[hi] = total
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (base[mid] <= c) lo = mid;
                else hi = mid;
            }
            job = lo;
        }
        j = job;
        first = (long long)(c - base[job]) * chunk;
        return true;
    }
};

// Held to 4 wavefronts per SIMD (128 VGPRs; the closest stage spills 36 B) as k_closest is:
// 64-scene batch 1479 -> 1521 Mrays/s against 3 wavefronts at 136 VGPRs.
template <int kStage>
__global__ __launch_bounds__(kBatchBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_batch_trace(BATCH_PARAMS) {
    __shared__ uint32_t s_node[kMaxKdDepth * kBatchBlock];
    __shared__ float s_t0[kMaxKdDepth * kBatchBlock];
    StackRef stk{s_node + threadIdx.x, s_t0 + threadIdx.x, kBatchBlock};
    ChunkQueue q(base, counter, n_jobs, chunk);
    int j;
    long long i0;
    while (q.pop(j, i0)) {
        const SceneView& S = *(const SceneView*)(views + j);
        const FrameView& F = *(const FrameView*)(frames + 2 * j + (level & 1));
        const long long n = stage_items(kStage, S, F, level);
        for (int k = 0; k < chunk; k += kBatchBlock) {
            const long long i = i0 + k + threadIdx.x;
            const bool live = i < n;
            if (kStage == kStClosest) {
                Work w;
                bool hit = false, tie = false;
                if (live) hit = closest_item<false>(S, F, (int)i, level, stk, w, &tie);
                const int pos = wave_append(&F.hit_count[level], hit);
                if (hit) F.hitq[pos] = (int)i;
                const int tpos = wave_append(&F.tie_count[level], tie);
                if (tie) F.tieq[tpos] = (int)i;
            } else if (kStage == kStResolve) {
                if (live) resolve_item(S, F, F.tieq[i], level, stk);
            } else {  // kStShadow
                Work w;
                if (live) shadow_item<false>(S, F, (int)i, stk, w);
            }
        }
    }
}

template <int kStage>
__global__ __launch_bounds__(kBatchBlock) void k_batch_plain(BATCH_PARAMS) {
    ChunkQueue q(base, counter, n_jobs, chunk);
    int j;
    long long i0;
    while (q.pop(j, i0)) {
        const SceneView& S = *(const SceneView*)(views + j);
        const FrameView& F = *(const FrameView*)(frames + 2 * j + (level & 1));
        const long long n = stage_items(kStage, S, F, level);
        for (int k = 0; k < chunk; k += kBatchBlock) {
            const long long i = i0 + k + threadIdx.x;
            const bool live = i < n;
            if (kStage == kStRaygen) {
                if (live) F.rays[i] = raygen_item(S, F, (int)i);
            } else if (kStage == kStShadePrep) {
                RayRec refl;
                bool spawn = false;
                if (live) spawn = shade_prepare_item(S, F, (int)i, level, &refl);
                const int pos = wave_append(&F.counts[level + 1], spawn);
                if (spawn) F.rays_next[pos] = refl;
            } else if (kStage == kStFinish) {
                if (live) shade_finish_item(S, F, (int)i);
            } else if (kStage == kStPixel) {
                if (live) pixel_item(F, (int)i);
            } else {  // kStFinalize
                if (live) finalize_item(F, nullptr, (int)i);
            }
        }
    }
}

}  // namespace

hipError_t launch_batch(const BatchArgs& a, const BatchLaunch& l, hipStream_t st) {
    const dim3 blk(kBatchBlock), grid(l.grid);
#define BATCH_ARGS(lvl)                                                                                      \
    (const RT_AS_CONST SceneView*)a.views, (const RT_AS_CONST FrameView*)a.frames, a.base, a.counter, a.n_jobs, \
        a.chunk, (lvl)
    auto plan = [&](int stage, int level) { hipLaunchKernelGGL(k_plan, dim3(1), dim3(1024), 0, st, a, stage, level); };
    hipLaunchKernelGGL(k_batch_init, dim3(a.n_jobs), dim3(64), 0, st, a);
    plan(kStRaygen, 0);
    hipLaunchKernelGGL(k_batch_plain<kStRaygen>, grid, blk, 0, st, BATCH_ARGS(0));
    for (int level = 0; level < l.max_depth; ++level) {
        plan(kStClosest, level);
        hipLaunchKernelGGL(k_batch_trace<kStClosest>, grid, blk, 0, st, BATCH_ARGS(level));
        plan(kStResolve, level);
        hipLaunchKernelGGL(k_batch_trace<kStResolve>, dim3(64), blk, 0, st, BATCH_ARGS(level));
        plan(kStShadePrep, level);
        hipLaunchKernelGGL(k_batch_plain<kStShadePrep>, grid, blk, 0, st, BATCH_ARGS(level));
        plan(kStShadow, level);
        hipLaunchKernelGGL(k_batch_trace<kStShadow>, grid, blk, 0, st, BATCH_ARGS(level));
        plan(kStFinish, level);
        hipLaunchKernelGGL(k_batch_plain<kStFinish>, grid, blk, 0, st, BATCH_ARGS(level));
    }
    plan(kStPixel, 0);
    hipLaunchKernelGGL(k_batch_plain<kStPixel>, grid, blk, 0, st, BATCH_ARGS(0));
    plan(kStFinalize, 0);
    hipLaunchKernelGGL(k_batch_plain<kStFinalize>, grid, blk, 0, st, BATCH_ARGS(0));
    hipLaunchKernelGGL(k_batch_tally, dim3((a.n_jobs + 63) / 64), dim3(64), 0, st, a);
#undef BATCH_ARGS
    return hipGetLastError();
}

}  // namespace rt
