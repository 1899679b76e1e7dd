// trace.hip — gfx950 kernels of the wavefront trace pipeline and their launchers.
//
// Stages per frame (one launch each, grid-stride over a device-side count so the host
// never waits between levels):
//   raygen -> for each level: closest -> shade_prepare -> shadow -> shade_finish
//          -> pixel -> [outline edges -> dilate] -> finalize (blend + quantise)
// Traversal kernels keep a 16-entry per-lane stack in LDS (entry i of lane l at
// [i][l], so the 64 lanes of a wavefront touch 64 consecutive banks). Queue appends
// are wavefront-aggregated: one atomic per wavefront, ranks from a ballot.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "launch.hpp"
#include "trace_core.hpp"
#include "wave_ops.hpp"

namespace rt {

namespace {

constexpr int kBlock = 128;       // shadow and resolve: 2 wavefronts, 16 KiB of LDS stack
constexpr int kClosestBlock = 64;  // closest hit: 1 wavefront, 8 KiB of LDS stack
constexpr int kXcd = 8;      // MI355X: 8 XCDs, workgroups dispatched round-robin
// traversal work queues (WaveQueue): per launch up to kXcd slice counters, kQueuePad
// int32 apart (one 128-byte line each)
constexpr int kQueuePad = 32;
constexpr int kClosestSlices = kXcd;
constexpr int kShadowSlices = kXcd;

// Workgroup id remapped so that each XCD (and its L2) walks one contiguous slice of
// the item range instead of every 8th workgroup.
__device__ __forceinline__ int xcd_block() {
    int nb = gridDim.x, b = blockIdx.x;
    if (nb % kXcd) return b;
    return (b % kXcd) * (nb / kXcd) + b / kXcd;
}

__global__ void k_init_counts(FrameView F) {
    for (int i = threadIdx.x; i <= F.max_depth; i += blockDim.x) {
        F.counts[i] = i == 0 && !F.tmask ? F.n_q0 : 0;
        if (i < F.max_depth) {
            F.hit_count[i] = 0;
            F.tie_count[i] = 0;
        }
    }
    for (int i = threadIdx.x; i < 3 * kXcd * kQueuePad * F.max_depth; i += blockDim.x) F.take[i] = 0;
    if (threadIdx.x < 2 * kWorkCount) F.work[threadIdx.x] = 0;
}

__global__ __launch_bounds__(256) void k_raygen(SceneView S, FrameView F) {
    int stride = gridDim.x * blockDim.x;
    for (int q = xcd_block() * blockDim.x + threadIdx.x; q < F.n_q0; q += stride) {
        const RayRec rec = raygen_item(S, F, q);
        if (F.tmask) {
            // a tile-masked render queues only the rays of its masked pixels
            const int pos = wave_append(&F.counts[0], rec.path >= 0);
            if (rec.path >= 0) F.rays[pos] = rec;
        } else {
            F.rays[q] = rec;
        }
    }
}

// Work queue of a traversal launch: items [0, n) in kSlices contiguous slices (8: one per
// XCD, as workgroups go to the XCDs round-robin by block id, so each L2 serves one part
// of the spatially ordered queue). Items go out in runs of `step`, sized for about 16 runs
// per wavefront (64 to 256 items). Each wavefront's first run is fixed: the r-th wavefront
// of a slice takes the slice's r-th run, with no atomic, so a short launch starts at once
// and ends without queue traffic. The runs after the fixed ones go out from the slice's
// counter, one atomic per run (each counter on its own 128-byte line), and a wavefront
// whose own slice is drained moves on to the others. Called where every lane of the
// wavefront is active.
template <int kSlices>
struct WaveQueue {
    int32_t* ctr;  // [kSlices * kQueuePad] runs handed out after the fixed ones
    int n, step, slice, left, rank;
    bool fixed;  // the fixed first run is still to come
    __device__ WaveQueue(int32_t* c, int items)
        : ctr(c), n(items), slice((int)(blockIdx.x % kSlices)), left(kSlices), fixed(true) {
        const int wpb = blockDim.x / 64;
        const long long waves = (long long)gridDim.x * wpb;
        step = (int)min(256ll, max(64ll, ((long long)n / (waves * 16)) & ~63ll));
        rank = __builtin_amdgcn_readfirstlane((int)(blockIdx.x / kSlices) * wpb + (int)(threadIdx.x / 64));
    }
    // wavefronts whose workgroups map to slice s
    __device__ int slice_waves(int s) const {
        const int blocks = ((int)gridDim.x - s + kSlices - 1) / kSlices;
        return max(blocks, 0) * (int)(blockDim.x / 64);
    }
    // [b, e): the next items of this wavefront; false when every slice is drained
    __device__ bool next(int& b, int& e) {
        if (fixed) {
            fixed = false;
            const int lo = (int)((long long)n * slice / kSlices);
            const int hi = (int)((long long)n * (slice + 1) / kSlices);
            if ((long long)rank * step < hi - lo) {
                b = lo + rank * step;
                e = min(hi, b + step);
                return true;
            }
        }
        while (left > 0) {
            const int lo = (int)((long long)n * slice / kSlices);
            const int hi = (int)((long long)n * (slice + 1) / kSlices);
            const long long fixed_part = (long long)slice_waves(slice) * step;
            if (fixed_part < hi - lo) {
                int32_t t = 0;
                if (__lane_id() == 0) t = atomicAdd(&ctr[slice * kQueuePad], step);
                t = __builtin_amdgcn_readfirstlane(t);
                if (fixed_part + t < hi - lo) {
                    b = lo + (int)fixed_part + t;
                    e = min(hi, b + step);
                    return true;
                }
            }
            slice = (slice + 1) % kSlices;
            --left;
        }
        return false;
    }
};

// Traversal kernels run a resident grid whose wavefronts draw 64 items at a time from
// the launch's queue until none is left: a wavefront that drew short rays takes more,
// so the launch ends with the last rays rather than with the unluckiest static slice.
// Held to 128 VGPRs (4 wavefronts per SIMD instead of 3, for a few bytes of spill).
template <bool kCount>
__global__ __launch_bounds__(kClosestBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_closest(SceneView S,
                                                                                                   FrameView F,
                                                                                                   int level) {
    __shared__ uint32_t s_node[kMaxKdDepth * kClosestBlock];
    __shared__ float s_t0[kMaxKdDepth * kClosestBlock];
    StackRef stk{s_node + threadIdx.x, s_t0 + threadIdx.x, kClosestBlock};
    Work w;
    WaveQueue<kClosestSlices> queue(&F.take[3 * level * kXcd * kQueuePad], F.counts[level]);
    const int lane = __lane_id();
    // the loop is wavefront-uniform, so the ballot in wave_append sees the whole wavefront
    int b, e;
    while (queue.next(b, e))
        for (; b < e; b += 64) {
            const int slot = b + lane;
            bool hit = false, tie = false;
            if (slot < e) hit = closest_item<kCount>(S, F, slot, level, stk, w, &tie);
            int pos = wave_append(&F.hit_count[level], hit);
            if (hit) F.hitq[pos] = slot;
            int tpos = wave_append(&F.tie_count[level], tie);
            if (tie) F.tieq[tpos] = slot;
        }
    if (kCount) flush_work(F, w, 0);
}

// Reference-order tie resolution for the (rare) rays k_closest marked.
__global__ __launch_bounds__(kBlock) void k_resolve(SceneView S, FrameView F, int level) {
    __shared__ uint32_t s_node[kMaxKdDepth * kBlock];
    __shared__ float s_t0[kMaxKdDepth * kBlock];
    StackRef stk{s_node + threadIdx.x, s_t0 + threadIdx.x, kBlock};
    WaveQueue<1> queue(&F.take[(3 * level + 2) * kXcd * kQueuePad], F.tie_count[level]);
    const int lane = __lane_id();
    int b, e;
    while (queue.next(b, e))
        for (; b < e; b += 64)
            if (b + lane < e) resolve_item(S, F, F.tieq[b + lane], level, stk);
}

__global__ __launch_bounds__(256) void k_shade_prepare(SceneView S, FrameView F, int level) {
    const int n = F.hit_count[level];
    const int stride = gridDim.x * blockDim.x;
    const int base = xcd_block() * blockDim.x + (int)threadIdx.x;
    const int iters = (n + stride - 1) / stride;
    for (int it = 0; it < iters; ++it) {
        int hq = base + it * stride;
        RayRec refl;
        bool spawn = false;
        if (hq < n) spawn = shade_prepare_item(S, F, hq, level, &refl);
        int pos = wave_append(&F.counts[level + 1], spawn);
        if (spawn) F.rays_next[pos] = refl;
    }
}

// Shadow rays (any hit), 64 queue items per wavefront step. kByLight (scenes with several
// lights): items go light by light within blocks of 64 hits (shadow_task).
template <bool kCount, bool kByLight>
__global__ __launch_bounds__(kBlock) void k_shadow(SceneView S, FrameView F, int level) {
    __shared__ uint32_t s_node[kMaxKdDepth * kBlock];
    __shared__ float s_t0[kMaxKdDepth * kBlock];
    StackRef stk{s_node + threadIdx.x, s_t0 + threadIdx.x, kBlock};
    Work w;
    // < 2^31 items: checked at frame set-up
    WaveQueue<kShadowSlices> queue(&F.take[(3 * level + 1) * kXcd * kQueuePad], F.hit_count[level] * S.shadow_per_hit);
    const int lane = __lane_id();
    const int n_hits = F.hit_count[level];
    int b, e;
    while (queue.next(b, e))
        for (; b < e; b += 64)
            if (b + lane < e) shadow_item<kCount>(S, F, kByLight ? shadow_task(S, n_hits, b + lane) : b + lane, stk, w);
    if (kCount) flush_work(F, w, 1);
}

// Held to 3 wavefronts per SIMD (168 VGPRs, a few bytes of spill) instead of the
// compiler's 2 at 180: the f64 Phong terms need the latency hiding (9 % on the 1080p icosphere scene).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_shade_finish(SceneView S, FrameView F, int level) {
    const int n = F.hit_count[level];
    const int stride = gridDim.x * blockDim.x;
    for (int hq = xcd_block() * blockDim.x + threadIdx.x; hq < n; hq += stride) shade_finish_item(S, F, hq);
}

// shade_finish for frames with few hits and many light samples (a tile-masked many-sample
// re-render): one wavefront per hit, its lanes compute the area-light sample terms into LDS,
// then one lane sums them in sample order (the same sum as the per-hit loop, which would
// otherwise run 256 samples serially on a few thousand lanes).
constexpr int kWideSlots = 512;
__global__ __launch_bounds__(256) void k_shade_finish_wide(SceneView S, FrameView F, int level) {
    __shared__ double s_term[4][kWideSlots * 3];
    const int wave = (int)threadIdx.x >> 6, lane = __lane_id();
    const int n = F.hit_count[level];
    double* term = s_term[wave];
    const int slots = S.shadow_per_hit;
    for (int hq = (int)blockIdx.x * 4 + wave; hq < n; hq += (int)gridDim.x * 4) {  // wave-uniform
        for (int j = lane; j < slots; j += 64) {
            const d3 t = area_slot_term(S, F, hq, j);
            term[3 * j] = t.x;
            term[3 * j + 1] = t.y;
            term[3 * j + 2] = t.z;
        }
        __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): this lane's LDS writes are done
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) shade_finish_item(S, F, hq, term);
        __builtin_amdgcn_wave_barrier();  // the sum has read the terms before they are replaced
    }
}

__global__ __launch_bounds__(256) void k_pixel(FrameView F) {
    int n = F.W * F.H;
    int stride = gridDim.x * blockDim.x;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) pixel_item(F, p);
}

// Edge strength as a 16x16 stencil with a one-pixel halo staged in LDS
// (compute_edge_strength + detect_edges, src/outline.rs:113-182).
constexpr int kOT = 16;
__global__ __launch_bounds__(kOT * kOT) void k_outline_edges(FrameView F) {
    __shared__ double s_d[kOT + 2][kOT + 2];
    __shared__ double s_n[kOT + 2][kOT + 2][3];
    __shared__ uint8_t s_f[kOT + 2][kOT + 2];
    const int x0 = blockIdx.x * kOT - 1, y0 = blockIdx.y * kOT - 1;
    for (int i = threadIdx.y * kOT + threadIdx.x; i < (kOT + 2) * (kOT + 2); i += kOT * kOT) {
        int lx = i % (kOT + 2), ly = i / (kOT + 2);
        int gx = x0 + lx, gy = y0 + ly;
        bool in = gx >= 0 && gy >= 0 && gx < F.W && gy < F.H;
        double d = -1.0;
        double n0 = 0, n1 = 0, n2 = 0;
        if (in) {
            const double* q = F.px_dn + ((size_t)gy * F.W + gx) * 4;
            d = q[0];
            n0 = q[1];
            n1 = q[2];
            n2 = q[3];
        }
        s_d[ly][lx] = d;
        s_n[ly][lx][0] = n0;
        s_n[ly][lx][1] = n1;
        s_n[ly][lx][2] = n2;
        s_f[ly][lx] = in && !(d < 0.0);  // a hit: depth >= 0, or NaN for a NaN camera ray
    }
    __syncthreads();
    int x = blockIdx.x * kOT + threadIdx.x, y = blockIdx.y * kOT + threadIdx.y;
    if (x >= F.W || y >= F.H) return;
    auto fetch = [&](int gx, int gy, double& d, double* n) -> bool {
        int lx = gx - x0, ly = gy - y0;  // always inside the halo tile
        if (!s_f[ly][lx]) return false;
        d = s_d[ly][lx];
        n[0] = s_n[ly][lx][0];
        n[1] = s_n[ly][lx][1];
        n[2] = s_n[ly][lx][2];
        return true;
    };
    double e = edge_strength(F, x, y, fetch);
    F.edge[(size_t)y * F.W + x] = edge_mask_value(F, e);
}

__global__ __launch_bounds__(256) void k_outline_dilate(FrameView F) {
    int n = F.W * F.H;
    int stride = gridDim.x * blockDim.x;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride)
        F.edge2[p] = dilate_item(F, F.edge, p % F.W, p / F.W);
}

__global__ __launch_bounds__(256) void k_finalize(FrameView F, const double* mask) {
    int n = F.W * F.H;
    int stride = gridDim.x * blockDim.x;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) finalize_item(F, mask, p);
}

// rt_intersect: closest hit of caller rays (World::hit seam).
__global__ __launch_bounds__(kBlock) void k_intersect(SceneView S, const double* o, const double* d, int n, double tmin,
                                                     double tmax, int use_kd, int32_t* obj, int32_t* tri, double* t) {
    __shared__ uint32_t s_node[kMaxKdDepth * kBlock];
    __shared__ float s_t0[kMaxKdDepth * kBlock];
    StackRef stk{s_node + threadIdx.x, s_t0 + threadIdx.x, kBlock};
    Work w;
    int stride = gridDim.x * kBlock;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        RayD r;
        r.o = mk(o[3 * i], o[3 * i + 1], o[3 * i + 2]);
        r.d = unit(mk(d[3 * i], d[3 * i + 1], d[3 * i + 2]));
        double tt;
        int32_t ob, tr;
        bool h = world_closest<false>(S, r, tmin, tmax, use_kd != 0, stk, w, tt, ob, tr);
        if (h && tr == kTiePending) tr = mesh_resolve_tie(S.meshes[S.objs[ob].mesh], r, tmin, tt, stk);
        obj[i] = h ? ob : -1;
        tri[i] = h ? tr : -1;
        t[i] = h ? tt : 0.0;
    }
}

int grid_for(long long items, int block, int cap) {
    long long g = (items + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    // a multiple of the XCD count keeps xcd_block() a permutation
    if (g > kXcd) g -= g % kXcd;
    return (int)g;
}

// Workgroups of `kern` resident on the whole device at once (every CU at the kernel's
// occupancy), and no more than `items` need at 64 items per wavefront. The runtime
// queries run once per (device, kernel): a frame launches these kernels tens of times.
int resident_grid(const void* kern, int block, long long items) {
    struct Entry {
        const void* kern;
        int dev, block, blocks;
    };
    static std::mutex mu;
    static std::vector<Entry> cache;
    int dev = 0;
    (void)hipGetDevice(&dev);
    int blocks = 0;
    {
        std::lock_guard<std::mutex> lock(mu);
        for (const Entry& e : cache)
            if (e.kern == kern && e.dev == dev && e.block == block) blocks = e.blocks;
        if (blocks == 0) {
            int cus = 0, per_cu = 0;
            if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
                cus = 256;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, block, 0) != hipSuccess || per_cu < 1)
                per_cu = 1;
            blocks = cus * per_cu;
            cache.push_back(Entry{kern, dev, block, blocks});
        }
    }
    const long long need = (items + block - 1) / block;
    return (int)std::max<long long>(1, std::min<long long>(need, blocks));
}

}  // namespace

void warm_frame_launches() {
    // the occupancy queries of the traversal kernels, before a stream capture
    (void)resident_grid((const void*)k_closest<false>, kClosestBlock, 1);
    (void)resident_grid((const void*)k_closest<true>, kClosestBlock, 1);
    (void)resident_grid((const void*)k_resolve, kBlock, 1);
    (void)resident_grid((const void*)k_shadow<false, false>, kBlock, 1);
    (void)resident_grid((const void*)k_shadow<false, true>, kBlock, 1);
    (void)resident_grid((const void*)k_shadow<true, false>, kBlock, 1);
}

hipError_t launch_frame(const SceneView& S, FrameView F, RayRec* q0, RayRec* q1, hipStream_t st, StageTimer* tm) {
    constexpr int kCap = 256 * 16;  // 256 CUs x 16 workgroups
    auto mark = [&](int stage, bool begin) {
        if (tm) (void)tm->mark(stage, begin, st);
    };
    mark(0, true);
    hipLaunchKernelGGL(k_init_counts, dim3(1), dim3(64), 0, st, F);
    F.rays = q0;
    hipLaunchKernelGGL(k_raygen, dim3(grid_for(F.n_q0, 256, kCap)), dim3(256), 0, st, S, F);
    mark(0, false);
    for (int level = 0; level < F.max_depth; ++level) {
        F.rays = (level & 1) ? q1 : q0;
        F.rays_next = (level & 1) ? q0 : q1;
        // the device-side counts bound the work; grids are sized for the largest case
        long long cap_rays = level == 0 ? F.n_q0 : F.n_paths;
        mark(1, true);
        if (F.count_work)
            hipLaunchKernelGGL(k_closest<true>, dim3(resident_grid((const void*)k_closest<true>, kClosestBlock, cap_rays)),
                               dim3(kClosestBlock), 0, st, S, F, level);
        else
            hipLaunchKernelGGL(k_closest<false>,
                               dim3(resident_grid((const void*)k_closest<false>, kClosestBlock, cap_rays)),
                               dim3(kClosestBlock), 0, st, S, F, level);
        // ties are rare: a few workgroups per XCD drain the resolver's queue
        hipLaunchKernelGGL(k_resolve, dim3(std::min(256, resident_grid((const void*)k_resolve, kBlock, cap_rays))),
                           dim3(kBlock), 0, st, S, F, level);
        mark(1, false);
        mark(2, true);
        hipLaunchKernelGGL(k_shade_prepare, dim3(grid_for(F.n_paths, 256, kCap)), dim3(256), 0, st, S, F, level);
        mark(2, false);
        mark(3, true);
        const long long cap_shadow = (long long)F.n_paths * S.shadow_per_hit;
        if (S.n_lights == 0) {
            // no shadow rays
        } else if (F.count_work) {  // counts do not depend on the order
            hipLaunchKernelGGL((k_shadow<true, false>),
                               dim3(resident_grid((const void*)k_shadow<true, false>, kBlock, cap_shadow)), dim3(kBlock),
                               0, st, S, F, level);
        } else if (S.n_lights > 1) {
            hipLaunchKernelGGL((k_shadow<false, true>),
                               dim3(resident_grid((const void*)k_shadow<false, true>, kBlock, cap_shadow)), dim3(kBlock),
                               0, st, S, F, level);
        } else {
            hipLaunchKernelGGL((k_shadow<false, false>),
                               dim3(resident_grid((const void*)k_shadow<false, false>, kBlock, cap_shadow)), dim3(kBlock),
                               0, st, S, F, level);
        }
        mark(3, false);
        mark(4, true);
        if (F.tmask && S.n_lights > 0 && S.shadow_per_hit >= 32 && S.shadow_per_hit <= kWideSlots)
            hipLaunchKernelGGL(k_shade_finish_wide, dim3(grid_for((long long)F.n_paths * 64, 256, kCap)), dim3(256), 0, st,
                               S, F, level);
        else
            hipLaunchKernelGGL(k_shade_finish, dim3(grid_for(F.n_paths, 256, kCap)), dim3(256), 0, st, S, F, level);
        mark(4, false);
    }
    mark(5, true);
    int npix = F.W * F.H;
    hipLaunchKernelGGL(k_pixel, dim3(grid_for(npix, 256, kCap)), dim3(256), 0, st, F);
    const double* mask = nullptr;
    if (F.outline) {
        dim3 g((F.W + kOT - 1) / kOT, (F.H + kOT - 1) / kOT);
        hipLaunchKernelGGL(k_outline_edges, g, dim3(kOT, kOT), 0, st, F);
        mask = F.edge;
        if (F.ol_thick > 1.0 && (int)ceil(F.ol_thick - 1.0) > 0) {
            hipLaunchKernelGGL(k_outline_dilate, dim3(grid_for(npix, 256, kCap)), dim3(256), 0, st, F);
            mask = F.edge2;
        }
    }
    hipLaunchKernelGGL(k_finalize, dim3(grid_for(npix, 256, kCap)), dim3(256), 0, st, F, mask);
    mark(5, false);
    return hipGetLastError();
}

hipError_t launch_intersect(const SceneView& S, const double* o, const double* d, int n, double tmin, double tmax,
                            int use_kd, int32_t* obj, int32_t* tri, double* t, hipStream_t st) {
    hipLaunchKernelGGL(k_intersect, dim3(grid_for(n, kBlock, 4096)), dim3(kBlock), 0, st, S, o, d, n, tmin, tmax,
                       use_kd, obj, tri, t);
    return hipGetLastError();
}

}  // namespace rt
