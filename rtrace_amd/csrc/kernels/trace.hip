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

#include "launch.hpp"
#include "trace_core.hpp"
#include "wave_ops.hpp"

namespace rt {

namespace {

constexpr int kBlock = 128;  // traversal kernels: 2 wavefronts, 16 KiB of LDS stack
constexpr int kXcd = 8;      // MI355X: 8 XCDs, workgroups dispatched round-robin

// Workgroup id remapped so that each XCD (and its L2) walks one contiguous slice of
// the item range instead of every 8th workgroup.
__device__ __forceinline__ int xcd_block() {
    int nb = gridDim.x, b = blockIdx.x;
    if (nb % kXcd) return b;
    return (b % kXcd) * (nb / kXcd) + b / kXcd;
}

__global__ void k_init_counts(FrameView F) {
    for (int i = threadIdx.x; i <= F.max_depth; i += blockDim.x) {
        F.counts[i] = i == 0 ? F.n_q0 : 0;
        if (i < F.max_depth) {
            F.hit_count[i] = 0;
            F.tie_count[i] = 0;
        }
    }
    if (threadIdx.x < 2 * kWorkCount) F.work[threadIdx.x] = 0;
}

__global__ __launch_bounds__(256) void k_raygen(SceneView S, FrameView F) {
    int stride = gridDim.x * blockDim.x;
    for (int q = xcd_block() * blockDim.x + threadIdx.x; q < F.n_q0; q += stride) raygen_item(S, F, q);
}

template <bool kCount>
__global__ __launch_bounds__(kBlock) void k_closest(SceneView S, FrameView F, int level) {
    __shared__ uint32_t s_node[kMaxKdDepth * kBlock];
    __shared__ float s_t0[kMaxKdDepth * kBlock];
    StackRef stk{s_node + threadIdx.x, s_t0 + threadIdx.x, kBlock};
    Work w;
    const int n = F.counts[level];
    const int stride = gridDim.x * kBlock;
    // every lane of a wavefront runs the same number of iterations so the ballot in
    // wave_append sees the whole wavefront
    const int base = xcd_block() * kBlock + (int)threadIdx.x;
    const int iters = (n + stride - 1) / stride;
    for (int it = 0; it < iters; ++it) {
        int slot = base + it * stride;
        bool hit = false, tie = false;
        if (slot < n) hit = closest_item<kCount>(S, F, slot, level, stk, w, &tie);
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
    const int n = F.tie_count[level];
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        resolve_item(S, F, F.tieq[i], level, stk);
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

template <bool kCount>
__global__ __launch_bounds__(kBlock) void k_shadow(SceneView S, FrameView F, int level) {
    __shared__ uint32_t s_node[kMaxKdDepth * kBlock];
    __shared__ float s_t0[kMaxKdDepth * kBlock];
    StackRef stk{s_node + threadIdx.x, s_t0 + threadIdx.x, kBlock};
    Work w;
    const long long n = (long long)F.hit_count[level] * S.shadow_per_hit;
    const long long stride = (long long)gridDim.x * kBlock;
    for (long long t = (long long)xcd_block() * kBlock + threadIdx.x; t < n; t += stride)
        shadow_item<kCount>(S, F, (int)t, stk, w);
    if (kCount) flush_work(F, w, 1);
}

__global__ __launch_bounds__(256) void k_shade_finish(SceneView S, FrameView F, int level) {
    const int n = F.hit_count[level];
    const int stride = gridDim.x * blockDim.x;
    for (int hq = xcd_block() * blockDim.x + threadIdx.x; hq < n; hq += stride) shade_finish_item(S, F, hq);
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
        double d = NAN;
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
        s_f[ly][lx] = in && d == d;
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

}  // namespace

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
            hipLaunchKernelGGL(k_closest<true>, dim3(grid_for(cap_rays, kBlock, kCap * 2)), dim3(kBlock), 0, st, S, F,
                               level);
        else
            hipLaunchKernelGGL(k_closest<false>, dim3(grid_for(cap_rays, kBlock, kCap * 2)), dim3(kBlock), 0, st, S, F,
                               level);
        hipLaunchKernelGGL(k_resolve, dim3(64), dim3(kBlock), 0, st, S, F, level);
        mark(1, false);
        mark(2, true);
        hipLaunchKernelGGL(k_shade_prepare, dim3(grid_for(F.n_paths, 256, kCap)), dim3(256), 0, st, S, F, level);
        mark(2, false);
        mark(3, true);
        dim3 gs(grid_for((long long)F.n_paths * S.shadow_per_hit, kBlock, kCap * 2));
        if (S.n_lights == 0) {
            // no shadow rays
        } else if (F.count_work) {
            hipLaunchKernelGGL(k_shadow<true>, gs, dim3(kBlock), 0, st, S, F, level);
        } else {
            hipLaunchKernelGGL(k_shadow<false>, gs, dim3(kBlock), 0, st, S, F, level);
        }
        mark(3, false);
        mark(4, true);
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
