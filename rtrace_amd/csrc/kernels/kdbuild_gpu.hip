// kdbuild_gpu.hip — the reference's k-d build (KdTree::new / build_recursive,
// src/mesh.rs:75-197) on the device, producing exactly the arrays of the host build
// (host/kdbuild.cpp): same nodes, same depth-first pre-order, same triangle records.
//
// The tree is built one level at a time. Every node of a level still to be split is a
// contiguous segment of one element array (triangle ids in the node's list order). Per
// level: the elements are sorted stably by (segment, centroid on the level's axis) with
// two stable radix sorts (centroid key, then segment id), which is the reference's
// per-node stable sort by centroid; the split is the centroid of each segment's middle
// element; per element the left / right tests (triangle min <= split, max > split) set two
// flags whose segmented exclusive scans give each element's rank in the child lists; the
// empty-side rule moves a segment's last element. Children that become leaves (depth 16,
// or at most 10 triangles) go to a leaf store, the others form the next level's segments,
// in the parent's order. The host keeps the tree's shape (a few numbers per node and
// level), lays the nodes out in depth-first pre-order and writes the node array; the
// device writes the triangle records, which are the bulk of the output.
#include <hip/hip_runtime.h>

#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../common/layout.hpp"
#include "../common/vec.hpp"
#include "kdbuild_gpu.hpp"

namespace rt {
namespace {

constexpr int kB = 256;

int blocks_for(long long n) { return (int)std::max<long long>(1, std::min<long long>((n + kB - 1) / kB, 65535ll * 8)); }

// centroid ((v0 + v1) + v2) / 3 and min / max per axis, structure of arrays
__global__ void k_tri_prep(const double* __restrict__ v, uint32_t n, double* cen, double* lo, double* hi,
                           uint32_t* nan_flag) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const double* p = v + (size_t)t * 9;
        for (int a = 0; a < 3; ++a) {
            double s = p[a] + p[3 + a];
            s = s + p[6 + a];
            const double c = s / 3.0;
            if (c != c) *nan_flag = 1;
            cen[(size_t)a * n + t] = c;
            lo[(size_t)a * n + t] = fmin_r(fmin_r(p[a], p[3 + a]), p[6 + a]);
            hi[(size_t)a * n + t] = fmax_r(fmax_r(p[a], p[3 + a]), p[6 + a]);
        }
    }
}

// order-preserving 64-bit key of a (non-NaN) f64; -0.0 and +0.0 compare equal, as
// partial_cmp has them, so both map to the key of +0.0
__device__ __forceinline__ uint64_t order_key(double d) {
    if (d == 0.0) d = 0.0;
    uint64_t u;
    memcpy(&u, &d, 8);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

__global__ void k_keys(const uint32_t* __restrict__ elems, uint32_t m, const double* __restrict__ cen_axis,
                       uint64_t* key, uint32_t* idx) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        key[i] = order_key(cen_axis[elems[i]]);
        idx[i] = i;
    }
}

// segment of each position (segments are contiguous ranges given by off[0..k])
__device__ __forceinline__ uint32_t seg_of(const uint32_t* __restrict__ off, uint32_t k, uint32_t pos) {
    uint32_t lo = 0, hi = k;  // last segment whose start is <= pos
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= pos) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void k_seg_keys(const uint32_t* __restrict__ perm1, uint32_t m, const uint32_t* __restrict__ off,
                           uint32_t k, uint32_t* seg, uint32_t* val) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        const uint32_t p = perm1[i];
        seg[i] = seg_of(off, k, p);
        val[i] = p;
    }
}

__global__ void k_gather(const uint32_t* __restrict__ elems, const uint32_t* __restrict__ perm, uint32_t m,
                         uint32_t* out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x)
        out[i] = elems[perm[i]];
}

__global__ void k_splits(const uint32_t* __restrict__ sorted, const uint32_t* __restrict__ off, uint32_t k,
                         const double* __restrict__ cen_axis, double* split) {
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < k; s += gridDim.x * blockDim.x) {
        const uint32_t size = off[s + 1] - off[s];
        split[s] = cen_axis[sorted[off[s] + size / 2]];
    }
}

// left / right flags of each sorted element (one extra zero at m for the scans' totals)
__global__ void k_flags(const uint32_t* __restrict__ sorted, uint32_t m, const uint32_t* __restrict__ seg,
                        const double* __restrict__ split, const double* __restrict__ lo_axis,
                        const double* __restrict__ hi_axis, uint32_t* lf, uint32_t* rf) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= m; i += gridDim.x * blockDim.x) {
        if (i == m) {
            lf[i] = 0;
            rf[i] = 0;
            continue;
        }
        const uint32_t t = sorted[i];
        const double sp = split[seg[i]];
        lf[i] = lo_axis[t] <= sp ? 1u : 0u;
        rf[i] = hi_axis[t] > sp ? 1u : 0u;
    }
}

// per segment: side counts, and the empty-side rule (src/mesh.rs:174-178) applied to the
// segment's last element (which then has rank 0 on its new side and was last on the old)
__global__ void k_counts(const uint32_t* __restrict__ off, uint32_t k, const uint32_t* __restrict__ ls,
                         const uint32_t* __restrict__ rs, uint32_t* lf, uint32_t* rf, uint32_t* cl, uint32_t* cr,
                         uint32_t* moved) {
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < k; s += gridDim.x * blockDim.x) {
        const uint32_t a = off[s], b = off[s + 1], last = b - 1;
        uint32_t nl = ls[b] - ls[a], nr = rs[b] - rs[a], mv = 0;
        if (nl == 0) {
            lf[last] = 1;
            rf[last] = 0;
            nl = 1;
            nr -= 1;
            mv = 1;
        } else if (nr == 0) {
            rf[last] = 1;
            lf[last] = 0;
            nr = 1;
            nl -= 1;
            mv = 1;
        }
        cl[s] = nl;
        cr[s] = nr;
        moved[s] = mv;
    }
}

// every element to its child lists: dst[side][s] is the child's first slot, with the top
// bit set when the child is a leaf (leaf store) rather than a segment of the next level
constexpr uint32_t kToLeaf = 0x80000000u;
__global__ void k_scatter(const uint32_t* __restrict__ sorted, uint32_t m, const uint32_t* __restrict__ seg,
                          const uint32_t* __restrict__ off, const uint32_t* __restrict__ lf,
                          const uint32_t* __restrict__ rf, const uint32_t* __restrict__ ls,
                          const uint32_t* __restrict__ rs, const uint32_t* __restrict__ dst_l,
                          const uint32_t* __restrict__ dst_r, uint32_t* next, uint32_t* leaf_store) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        const uint32_t s = seg[i], t = sorted[i], a = off[s];
        if (lf[i]) {
            // the moved last element has rank 0 (its side was empty before the move)
            const uint32_t rank = ls[i] - ls[a], d = dst_l[s];
            ((d & kToLeaf) ? leaf_store : next)[(d & ~kToLeaf) + rank] = t;
        }
        if (rf[i]) {
            const uint32_t rank = rs[i] - rs[a], d = dst_r[s];
            ((d & kToLeaf) ? leaf_store : next)[(d & ~kToLeaf) + rank] = t;
        }
    }
}

// tight f64 bounds (min xyz, max xyz) of each list [off[s], off[s+1]) of `elems`, one
// wavefront per list
__global__ void k_bounds(const uint32_t* __restrict__ elems, const uint32_t* __restrict__ off, uint32_t k,
                         const double* __restrict__ lo, const double* __restrict__ hi, uint32_t n, double* out) {
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64, lane = threadIdx.x & 63;
    if (wave >= k) return;
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = off[wave] + lane; i < off[wave + 1]; i += 64) {
        const uint32_t t = elems[i];
        for (int a = 0; a < 3; ++a) {
            mn[a] = fmin(mn[a], lo[(size_t)a * n + t]);
            mx[a] = fmax(mx[a], hi[(size_t)a * n + t]);
        }
    }
    for (int d = 32; d > 0; d >>= 1)
        for (int a = 0; a < 3; ++a) {
            mn[a] = fmin(mn[a], __shfl_xor(mn[a], d));
            mx[a] = fmax(mx[a], __shfl_xor(mx[a], d));
        }
    if (lane == 0)
        for (int a = 0; a < 3; ++a) {  // zero without sign, as the host build stores it
            out[(size_t)wave * 6 + a] = mn[a] == 0.0 ? 0.0 : mn[a];
            out[(size_t)wave * 6 + 3 + a] = mx[a] == 0.0 ? 0.0 : mx[a];
        }
}

// the 48-byte triangle record (the host's push_rec): v0 - c, e1, e2 rounded to f32, the
// triangle index, max(|e1|_1, |e2|_1) * (1 + 1e-6) and a zero word
__device__ __forceinline__ void write_rec(float* r, const double* __restrict__ v, uint32_t t, double cx, double cy,
                                          double cz) {
    const double* p = v + (size_t)t * 9;
    r[0] = (float)(p[0] - cx);
    r[1] = (float)(p[1] - cy);
    r[2] = (float)(p[2] - cz);
    r[3] = (float)(p[3] - p[0]);
    r[4] = (float)(p[4] - p[1]);
    r[5] = (float)(p[5] - p[2]);
    r[6] = (float)(p[6] - p[0]);
    r[7] = (float)(p[7] - p[1]);
    r[8] = (float)(p[8] - p[2]);
    r[9] = __uint_as_float(t);
    const double n1 = fabs(p[3] - p[0]) + fabs(p[4] - p[1]) + fabs(p[5] - p[2]);
    const double n2 = fabs(p[6] - p[0]) + fabs(p[7] - p[1]) + fabs(p[8] - p[2]);
    r[10] = (float)(fmax(n1, n2) * (1.0 + 1e-6));
    r[11] = 0.0f;
}

// records of the leaves: leaf j's elements [loff[j], loff[j+1]) of the store go to
// records [rbase[j], ...) in the same order
__global__ void k_leaf_recs(const uint32_t* __restrict__ store, uint32_t m, const uint32_t* __restrict__ loff,
                            uint32_t nleaf, const uint32_t* __restrict__ rbase, const double* __restrict__ v,
                            double cx, double cy, double cz, float* recs) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        const uint32_t j = seg_of(loff, nleaf, i);
        write_rec(recs + (size_t)(rbase[j] + (i - loff[j])) * 12, v, store[i], cx, cy, cz);
    }
}

// the brute-force leaf: every triangle in file order
__global__ void k_bf_recs(uint32_t n, uint32_t first, const double* __restrict__ v, double cx, double cy, double cz,
                          float* recs) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x)
        write_rec(recs + (size_t)(first + t) * 12, v, t, cx, cy, cz);
}

#define GPU_TRY(expr) HIP_TRY(expr)

// rocPRIM wrappers with a reusable temporary buffer
struct Temp {
    DevBuf buf;
    void* get(size_t bytes) {
        buf.reserve(bytes);
        return buf.p;
    }
};

void sort_u64(Temp& tmp, const uint64_t* kin, uint64_t* kout, const uint32_t* vin, uint32_t* vout, uint32_t n,
              hipStream_t st) {
    size_t bytes = 0;
    GPU_TRY(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, n, 0, 64, st));
    void* p = tmp.get(bytes);
    GPU_TRY(rocprim::radix_sort_pairs(p, bytes, kin, kout, vin, vout, n, 0, 64, st));
}
void sort_u32(Temp& tmp, const uint32_t* kin, uint32_t* kout, const uint32_t* vin, uint32_t* vout, uint32_t n,
              int bits, hipStream_t st) {
    size_t bytes = 0;
    GPU_TRY(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, n, 0, bits, st));
    void* p = tmp.get(bytes);
    GPU_TRY(rocprim::radix_sort_pairs(p, bytes, kin, kout, vin, vout, n, 0, bits, st));
}
void scan_u32(Temp& tmp, const uint32_t* in, uint32_t* out, uint32_t n, hipStream_t st) {
    size_t bytes = 0;
    GPU_TRY(rocprim::exclusive_scan(nullptr, bytes, in, out, 0u, n, rocprim::plus<uint32_t>(), st));
    void* p = tmp.get(bytes);
    GPU_TRY(rocprim::exclusive_scan(p, bytes, in, out, 0u, n, rocprim::plus<uint32_t>(), st));
}

template <class T>
void to_dev(DevBuf& b, const std::vector<T>& v, hipStream_t st) {
    b.reserve(std::max<size_t>(v.size(), 1) * sizeof(T));
    if (!v.empty()) GPU_TRY(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
}
template <class T>
void to_host(std::vector<T>& v, const DevBuf& b, size_t n, hipStream_t st) {
    v.resize(n);
    if (n) GPU_TRY(hipMemcpyAsync(v.data(), b.p, n * sizeof(T), hipMemcpyDeviceToHost, st));
    GPU_TRY(hipStreamSynchronize(st));
}

// f32 copy of f64 bounds relative to the centre, rounded outward (the host's set_bounds)
void f32_bounds(KdNode& nd, const double* nb, const double* c) {
    for (int k = 0; k < 3; ++k) {
        const double lo = nb[k] - c[k], hi = nb[3 + k] - c[k];
        float flo = (float)lo, fhi = (float)hi;
        if ((double)flo > lo) flo = std::nextafter(flo, -INFINITY);
        if ((double)fhi < hi) fhi = std::nextafter(fhi, INFINITY);
        nd.bmin[k] = flo;
        nd.bmax[k] = fhi;
    }
}

// The tree's shape, kept on the host: one entry per node in creation (level) order.
struct ShapeNode {
    bool leaf = false;
    uint32_t flags = 0;  // axis | kFlagMisplaced (inner)
    double split = 0.0;
    int left = -1, right = -1;
    uint32_t leaf_idx = 0;  // leaves: index in the leaf store's list order
    double nb[6] = {0, 0, 0, 0, 0, 0};
};

}  // namespace

void build_flat_kd_device(const std::vector<double>& verts, const double* d_tri64, FlatKd& out, DevBuf& d_nodes,
                          DevBuf& d_recs, DevBuf& d_nb64, hipStream_t st) {
    const uint32_t n = (uint32_t)(verts.size() / 9);
    out = FlatKd{};
    // mesh bounds and centre (Mesh::compute_bounds, src/mesh.rs:1017-1033), as the host build
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = 0; i < (size_t)n * 3; ++i)
        for (int k = 0; k < 3; ++k) {
            mn[k] = fmin_r(mn[k], verts[i * 3 + k]);
            mx[k] = fmax_r(mx[k], verts[i * 3 + k]);
        }
    if (n == 0)
        for (int k = 0; k < 3; ++k) mn[k] = mx[k] = 0.0;
    for (int k = 0; k < 3; ++k) {
        out.bmin[k] = mn[k];
        out.bmax[k] = mx[k];
        out.center[k] = 0.5 * (mn[k] + mx[k]);
    }
    const double* c = out.center;
    if (n <= 10) {  // a root leaf (or no tree): nothing worth a device build
        FlatKd h;
        build_flat_kd(verts, h, false);
        out = std::move(h);
        to_dev(d_nodes, out.nodes, st);
        to_dev(d_recs, out.recs, st);
        to_dev(d_nb64, out.nb64, st);
        GPU_TRY(hipStreamSynchronize(st));
        return;
    }

    Temp tmp;
    DevBuf cen, lo, hi, flag;
    cen.reserve((size_t)n * 3 * sizeof(double));
    lo.reserve((size_t)n * 3 * sizeof(double));
    hi.reserve((size_t)n * 3 * sizeof(double));
    flag.reserve(sizeof(uint32_t));
    GPU_TRY(hipMemsetAsync(flag.p, 0, sizeof(uint32_t), st));
    hipLaunchKernelGGL(k_tri_prep, dim3(blocks_for(n)), dim3(kB), 0, st, d_tri64, n, cen.as<double>(), lo.as<double>(),
                       hi.as<double>(), flag.as<uint32_t>());
    GPU_TRY(hipGetLastError());
    uint32_t nan = 0;
    GPU_TRY(hipMemcpyAsync(&nan, flag.p, sizeof(nan), hipMemcpyDeviceToHost, st));
    GPU_TRY(hipStreamSynchronize(st));
    if (nan) throw Error(RT_E_SCENE, "k-d build: NaN centroid (the reference panics here)");

    std::vector<ShapeNode> shape(1);
    // current level: segment offsets, and the shape node of each segment
    std::vector<uint32_t> h_off = {0, n};
    std::vector<int> seg_node = {0};
    std::vector<uint32_t> h_iota(n);
    for (uint32_t i = 0; i < n; ++i) h_iota[i] = i;
    DevBuf elems, next;
    to_dev(elems, h_iota, st);
    std::vector<uint32_t> leaf_off = {0};  // leaf store list offsets
    std::vector<int> leaf_node;           // shape node of each stored leaf
    DevBuf store;                         // leaf store, grown per level
    size_t store_cap = 0, store_used = 0;
    DevBuf off, key, key_s, idx, perm1, seg, seg_s, perm, sorted, split, lf, rf, ls, rs, cl, cr, moved, dl, dr, bnd;

    for (int depth = 0; !seg_node.empty(); ++depth) {
        const uint32_t k = (uint32_t)seg_node.size(), m = h_off.back();
        const int axis = depth % 3;
        const double* cen_a = cen.as<double>() + (size_t)axis * n;
        to_dev(off, h_off, st);
        // stable sort by (segment, centroid): centroid key first, then segment id
        key.reserve((size_t)m * 8);
        key_s.reserve((size_t)m * 8);
        idx.reserve((size_t)m * 4);
        perm1.reserve((size_t)m * 4);
        hipLaunchKernelGGL(k_keys, dim3(blocks_for(m)), dim3(kB), 0, st, elems.as<uint32_t>(), m, cen_a,
                           key.as<uint64_t>(), idx.as<uint32_t>());
        GPU_TRY(hipGetLastError());
        sort_u64(tmp, key.as<uint64_t>(), key_s.as<uint64_t>(), idx.as<uint32_t>(), perm1.as<uint32_t>(), m, st);
        seg.reserve((size_t)m * 4);
        seg_s.reserve((size_t)m * 4);
        perm.reserve((size_t)m * 4);
        hipLaunchKernelGGL(k_seg_keys, dim3(blocks_for(m)), dim3(kB), 0, st, perm1.as<uint32_t>(), m,
                           off.as<uint32_t>(), k, seg.as<uint32_t>(), idx.as<uint32_t>());
        GPU_TRY(hipGetLastError());
        int bits = 1;
        while ((1u << bits) < k) ++bits;
        sort_u32(tmp, seg.as<uint32_t>(), seg_s.as<uint32_t>(), idx.as<uint32_t>(), perm.as<uint32_t>(), m, bits, st);
        sorted.reserve((size_t)m * 4);
        hipLaunchKernelGGL(k_gather, dim3(blocks_for(m)), dim3(kB), 0, st, elems.as<uint32_t>(), perm.as<uint32_t>(),
                           m, sorted.as<uint32_t>());
        GPU_TRY(hipGetLastError());
        // bounds of every node of this level (its whole list), splits, side flags and ranks
        bnd.reserve((size_t)k * 6 * sizeof(double));
        hipLaunchKernelGGL(k_bounds, dim3((k * 64 + kB - 1) / kB), dim3(kB), 0, st, sorted.as<uint32_t>(),
                           off.as<uint32_t>(), k, lo.as<double>(), hi.as<double>(), n, bnd.as<double>());
        GPU_TRY(hipGetLastError());
        split.reserve((size_t)k * 8);
        hipLaunchKernelGGL(k_splits, dim3(blocks_for(k)), dim3(kB), 0, st, sorted.as<uint32_t>(), off.as<uint32_t>(), k,
                           cen_a, split.as<double>());
        GPU_TRY(hipGetLastError());
        lf.reserve((size_t)(m + 1) * 4);
        rf.reserve((size_t)(m + 1) * 4);
        ls.reserve((size_t)(m + 1) * 4);
        rs.reserve((size_t)(m + 1) * 4);
        hipLaunchKernelGGL(k_flags, dim3(blocks_for(m + 1)), dim3(kB), 0, st, sorted.as<uint32_t>(), m,
                           seg_s.as<uint32_t>(), split.as<double>(), lo.as<double>() + (size_t)axis * n,
                           hi.as<double>() + (size_t)axis * n, lf.as<uint32_t>(), rf.as<uint32_t>());
        GPU_TRY(hipGetLastError());
        scan_u32(tmp, lf.as<uint32_t>(), ls.as<uint32_t>(), m + 1, st);
        scan_u32(tmp, rf.as<uint32_t>(), rs.as<uint32_t>(), m + 1, st);
        cl.reserve((size_t)k * 4);
        cr.reserve((size_t)k * 4);
        moved.reserve((size_t)k * 4);
        hipLaunchKernelGGL(k_counts, dim3(blocks_for(k)), dim3(kB), 0, st, off.as<uint32_t>(), k, ls.as<uint32_t>(),
                           rs.as<uint32_t>(), lf.as<uint32_t>(), rf.as<uint32_t>(), cl.as<uint32_t>(), cr.as<uint32_t>(),
                           moved.as<uint32_t>());
        GPU_TRY(hipGetLastError());
        std::vector<uint32_t> h_cl, h_cr, h_mv;
        std::vector<double> h_split, h_bnd;
        to_host(h_cl, cl, k, st);
        to_host(h_cr, cr, k, st);
        to_host(h_mv, moved, k, st);
        to_host(h_split, split, k, st);
        to_host(h_bnd, bnd, (size_t)k * 6, st);
        // children: shape nodes, destinations in the next level or the leaf store
        std::vector<uint32_t> h_dl(k), h_dr(k), n_off = {0};
        std::vector<int> n_node;
        for (uint32_t s = 0; s < k; ++s) {
            ShapeNode& p = shape[seg_node[s]];
            p.flags = (uint32_t)axis | (h_mv[s] ? kFlagMisplaced : 0u);
            p.split = h_split[s];
            memcpy(p.nb, &h_bnd[(size_t)s * 6], sizeof(p.nb));
            const uint32_t sizes[2] = {h_cl[s], h_cr[s]};
            for (int side = 0; side < 2; ++side) {
                const int child = (int)shape.size();
                shape.emplace_back();
                (side == 0 ? shape[seg_node[s]].left : shape[seg_node[s]].right) = child;
                uint32_t dst;
                if (depth + 1 >= kMaxKdDepth || sizes[side] <= 10) {
                    shape[child].leaf = true;
                    shape[child].leaf_idx = (uint32_t)leaf_node.size();
                    leaf_node.push_back(child);
                    dst = (uint32_t)(leaf_off.back()) | kToLeaf;
                    leaf_off.push_back(leaf_off.back() + sizes[side]);
                } else {
                    dst = n_off.back();
                    n_off.push_back(n_off.back() + sizes[side]);
                    n_node.push_back(child);
                }
                (side == 0 ? h_dl : h_dr)[s] = dst;
            }
        }
        if (leaf_off.back() >= kToLeaf || n_off.back() >= kToLeaf) throw Error(RT_E_OOM, "k-d build: too many records");
        // grow the leaf store (copying what it holds)
        if (leaf_off.back() > store_cap) {
            size_t cap = std::max<size_t>(leaf_off.back(), store_cap * 2);
            DevBuf grown;
            grown.reserve(cap * 4);
            if (store_used) GPU_TRY(hipMemcpyAsync(grown.p, store.p, store_used * 4, hipMemcpyDeviceToDevice, st));
            std::swap(store.p, grown.p);
            std::swap(store.bytes, grown.bytes);
            store_cap = cap;
        }
        store_used = leaf_off.back();
        to_dev(dl, h_dl, st);
        to_dev(dr, h_dr, st);
        next.reserve(std::max<size_t>(n_off.back(), 1) * 4);
        hipLaunchKernelGGL(k_scatter, dim3(blocks_for(m)), dim3(kB), 0, st, sorted.as<uint32_t>(), m,
                           seg_s.as<uint32_t>(), off.as<uint32_t>(), lf.as<uint32_t>(), rf.as<uint32_t>(),
                           ls.as<uint32_t>(), rs.as<uint32_t>(), dl.as<uint32_t>(), dr.as<uint32_t>(),
                           next.as<uint32_t>(), store.as<uint32_t>());
        GPU_TRY(hipGetLastError());
        GPU_TRY(hipStreamSynchronize(st));  // this level's host vectors go out of scope
        std::swap(elems.p, next.p);
        std::swap(elems.bytes, next.bytes);
        h_off = std::move(n_off);
        seg_node = std::move(n_node);
    }

    // leaves: bounds, then the depth-first pre-order layout
    const uint32_t nleaf = (uint32_t)leaf_node.size();
    DevBuf loff;
    to_dev(loff, leaf_off, st);
    bnd.reserve((size_t)nleaf * 6 * sizeof(double));
    hipLaunchKernelGGL(k_bounds, dim3((nleaf * 64 + kB - 1) / kB), dim3(kB), 0, st, store.as<uint32_t>(),
                       loff.as<uint32_t>(), nleaf, lo.as<double>(), hi.as<double>(), n, bnd.as<double>());
    GPU_TRY(hipGetLastError());
    std::vector<double> h_lb;
    to_host(h_lb, bnd, (size_t)nleaf * 6, st);
    for (uint32_t j = 0; j < nleaf; ++j) memcpy(shape[leaf_node[j]].nb, &h_lb[(size_t)j * 6], sizeof(double) * 6);

    // pre-order: node, left subtree, right subtree; records in the leaves' pre-order
    std::vector<uint32_t> rbase(nleaf);
    uint32_t recs_used = 0;
    out.nodes.reserve(shape.size() + 1);
    out.nb64.reserve((shape.size() + 1) * 6);
    std::vector<std::pair<int, int>> todo = {{0, -1}};  // (shape node, index of the parent to patch, -1 = none)
    while (!todo.empty()) {
        const auto [sn, patch] = todo.back();
        todo.pop_back();
        const ShapeNode& s = shape[sn];
        const uint32_t me = (uint32_t)out.nodes.size();
        if (patch >= 0) out.nodes[patch].a = me;  // right child index of an inner node
        KdNode nd{};
        f32_bounds(nd, s.nb, c);
        out.nb64.insert(out.nb64.end(), s.nb, s.nb + 6);
        if (s.leaf) {
            const uint32_t cnt = leaf_off[s.leaf_idx + 1] - leaf_off[s.leaf_idx];
            nd.split = 0.0;
            nd.a = recs_used;
            nd.b = (cnt << 2) | kAxisLeaf;
            rbase[s.leaf_idx] = recs_used;
            recs_used += cnt;
            out.leaves++;
            out.max_leaf = std::max<int>(out.max_leaf, (int)cnt);
            out.nodes.push_back(nd);
        } else {
            nd.split = s.split;
            nd.b = s.flags;
            out.nodes.push_back(nd);
            todo.push_back({s.right, (int)me});
            todo.push_back({s.left, -1});
        }
    }
    // the brute-force leaf over every triangle in file order (src/ray.rs:650-671)
    out.root_kd = 0;
    out.root_bf = (uint32_t)out.nodes.size();
    {
        KdNode bf{};
        double nb[6] = {mn[0], mn[1], mn[2], mx[0], mx[1], mx[2]};
        for (double& x : nb)
            if (x == 0.0) x = 0.0;
        f32_bounds(bf, nb, c);
        bf.a = recs_used;
        bf.b = (n << 2) | kAxisLeaf;
        out.nodes.push_back(bf);
        out.nb64.insert(out.nb64.end(), nb, nb + 6);
    }
    const size_t total_recs = (size_t)recs_used + n;
    d_recs.reserve(total_recs * 48);
    DevBuf d_rbase;
    to_dev(d_rbase, rbase, st);
    hipLaunchKernelGGL(k_leaf_recs, dim3(blocks_for(leaf_off.back())), dim3(kB), 0, st, store.as<uint32_t>(),
                       leaf_off.back(), loff.as<uint32_t>(), nleaf, d_rbase.as<uint32_t>(), d_tri64, c[0], c[1], c[2],
                       d_recs.as<float>());
    GPU_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_bf_recs, dim3(blocks_for(n)), dim3(kB), 0, st, n, recs_used, d_tri64, c[0], c[1], c[2],
                       d_recs.as<float>());
    GPU_TRY(hipGetLastError());
    to_dev(d_nodes, out.nodes, st);
    to_dev(d_nb64, out.nb64, st);
    GPU_TRY(hipStreamSynchronize(st));
}

}  // namespace rt
