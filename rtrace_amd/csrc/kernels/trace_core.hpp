// trace_core.hpp — per-item bodies of the wavefront trace pipeline.
//
// Every stage is a function of (scene view, frame view, item index) so the __global__
// kernels in trace.hip are thin grid-stride wrappers. The semantics are the
// reference's (SeaRyanC/rtrace src/lighting.rs, src/ray.rs, src/renderer.rs); the
// structure is the GPU's: FP32 k-d traversal with an LDS stack, FP64 re-test of
// survivors, compacted ray / hit queues, per-level colour records combined at the end.
#pragma once

#include "device_math.hpp"

namespace rt {

// ------------------------------------------------------------------ views
struct SceneView {
    const ObjDev* objs;
    const MeshDev* meshes;
    const MatDev* mats;
    const LightDev* lights;
    int32_t n_obj;
    int32_t n_lights;
    int32_t shadow_per_hit;  // shadow-ray slots per hit: 1 per point light, samples per area light
    int32_t spt_per_hit;     // stored sample points per hit (point and disk lights)
    int32_t has_square;      // some light is a square (its samples are regenerated, not stored)
    int32_t light_samples;   // > 0 overrides every area light's sample count
    int32_t light_cols;      // jittered-grid columns for light_samples
    int32_t has_fog;
    CameraDev cam;
    double cam_pos[3];
    double amb[3];
    double amb_i;
    double fog_color[3];
    double fog_density, fog_start, fog_end;
    double bg[3];
    // object BVH over spheres (null when the scene does not use one)
    const BvhNode* obvh;
    const int32_t* obvh_items;  // sphere object indices in leaf order
    const int32_t* others;      // every non-sphere object index, in scene order
    int32_t n_others;
    int32_t obvh_closest;       // closest-hit may use the BVH (the scene has no cube)
};

struct RayRec {  // 64 B queue entry
    double o[3];
    double d[3];
    int32_t path;   // -1 = hole (tiled raygen)
    int32_t level;
};

struct HitRec {  // per queue slot
    double t;
    int32_t obj;  // -1 miss
    int32_t tri;
};

struct ShadeRec {  // per hit-queue entry
    double p[3];
    double n[3];
    double mc[3];   // effective material colour (texture applied)
    int32_t mat;    // material whose coefficients apply (checkerboard alternate or base)
    int32_t path;
    int32_t level;
    int32_t base;   // object material (reflectivity source)
    int32_t spawned;  // a reflection ray continues this path
    int32_t pad;
};

// work counters (rt_work field order)
enum : int32_t { kWorkNodes = 0, kWorkTris = 1, kWorkRetests = 2, kWorkSphere = 3, kWorkPlane = 4, kWorkCube = 5,
                 kWorkSkipped = 6, kWorkBvhNodes = 7, kWorkCount = 8 };

struct FrameView {
    int32_t W, H, N;  // image size, samples per pixel (1 for quincunx)
    int32_t aa;       // 0 no-jitter, 1 quincunx, 2 stochastic
    int32_t max_depth;
    int32_t use_kd;
    int32_t outline;      // record first-hit depth/normal
    int32_t grid_miss;    // misses may show the ortho grid (camera passed to ray_color)
    int32_t count_work;
    int32_t n_paths;
    int32_t tile;         // raygen tile edge (pixels)
    int32_t tiles_x;      // raygen tiles per row
    int32_t n_q0;         // level-0 queue length including holes
    uint64_t seed;        // Renderer.seed.unwrap_or(0)
    uint64_t* path_seed;  // [n_paths]
    double* lvl;          // [max_depth][n_paths][4] (r, g, b, rho or -1 = terminal)
    int32_t* prim_obj;    // [n_paths]
    int32_t* prim_tri;
    double* prim_t;
    double* ol_dn;        // [n_paths][4] depth, nx, ny, nz (depth -1 = no hit: a hit's depth is a
                          // length, >= 0 or NaN for a NaN camera ray, which the reference outlines too)
    RayRec* rays;         // current level queue
    RayRec* rays_next;    // next level queue
    int32_t* counts;      // [max_depth + 1] queue lengths; counts[0] = n_q0
    HitRec* hits;         // [queue capacity]
    int32_t* hitq;        // [queue capacity] compacted slots that hit
    int32_t* hit_count;   // [max_depth]
    int32_t* tieq;        // [queue capacity] slots whose mesh tie is pending
    int32_t* tie_count;   // [max_depth]
    int32_t* take;        // [max_depth][closest, shadow, resolve][8 slices][32] items handed out
    ShadeRec* shade;      // [queue capacity]
    double* spt;          // stored sample points, sample-major: coordinate c of point s of hit hq at
                          // spt[(3 s + c) * spt_stride + hq] (consecutive hits, consecutive lanes)
    long long spt_stride; // the queue capacity
    uint8_t* vis;         // [queue capacity * shadow_per_hit]
    uint32_t* skip;       // [queue capacity]: bit l set when light l's queries are not traced
    unsigned long long* work;  // [2][kWorkCount]: closest, shadow
    double* pix;          // [W*H*3] final linear colour
    uint8_t* rgb;         // [W*H*3]
    // outline buffers
    double* px_dn;        // [W*H*4] per-pixel depth/normal (depth -1 = background)
    double* edge;         // [W*H]
    double* edge2;        // [W*H]
    double ol_wd, ol_wn, ol_thr, ol_thick;
    double ol_color[3];
    int32_t ol_8;
    // auxiliary outputs (null when not requested)
    float* feat;              // [W*H][kFeatCount] first-hit features of each pixel
    float* rad;               // [W*H][3] linear colour before quantisation
    const uint8_t* tmask;     // [mtiles_y][mtiles_x]: trace only pixels of tiles != 0
    int32_t mtile, mtiles_x;  // mask tile size in pixels, tiles per row
};

// Feature channels: noisy radiance RGB, albedo RGB, normal XYZ, depth, mean and variance
// of the area-light visibility, fog factor; taken from the centre ray (quincunx) or
// sample 0 of the pixel. Misses leave zeros.
enum : int32_t { kFeatRad = 0, kFeatAlbedo = 3, kFeatNormal = 6, kFeatDepth = 9, kFeatVisMean = 10,
                 kFeatVisVar = 11, kFeatFog = 12, kFeatCount = 13 };

// Pixel that path `path` provides the features of, or -1.
RT_HD int feature_pixel(const FrameView& F, int path) {
    if (F.aa == 1) return path < F.W * F.H ? path : -1;
    return path % F.N == 0 ? path / F.N : -1;
}
RT_HD bool pixel_in_mask(const FrameView& F, int x, int y) {
    if (!F.tmask) return true;
    if (x < 0 || y < 0 || x >= F.W || y >= F.H) return false;
    return F.tmask[(y / F.mtile) * F.mtiles_x + x / F.mtile] != 0;
}

// --------------------------------------------------------------- helpers
struct Work {
    unsigned long long v[kWorkCount] = {};
};

// Per-lane traversal stack: entry i of this lane lives at base[i * stride].
struct StackRef {
    uint32_t* node;
    float* t0;
    int stride;
};

RT_HD d3 ld3(const double* p) { return mk(p[0], p[1], p[2]); }
RT_HD void st3(double* p, d3 v) {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}

// The stored light points of one hit (FrameView::spt): point s at p[3 s * st], p[(3 s + 1) * st],
// p[(3 s + 2) * st], so the lanes of a wavefront (consecutive hits) touch consecutive doubles.
struct SptRef {
    double* p;
    long long st;
    RT_HD void put(int s, d3 v) const {
        p[3ll * s * st] = v.x;
        p[(3ll * s + 1) * st] = v.y;
        p[(3ll * s + 2) * st] = v.z;
    }
    RT_HD d3 get(int s) const { return mk(p[3ll * s * st], p[(3ll * s + 1) * st], p[(3ll * s + 2) * st]); }
    RT_HD SptRef at(int s) const { return SptRef{p + 3ll * s * st, st}; }
};

// HitRecord::new face orientation (src/ray.rs:39-63)
RT_HD d3 face(const d3& dir, const d3& outward) {
    bool front = dot(dir, outward) < 0.0;
    return front ? unit(outward) : unit(neg(outward));
}

// reflect (src/lighting.rs:291-294)
RT_HD d3 reflect_dir(d3 inc, d3 n) {
    double k = 2.0 * dot(inc, n);
    return unit(sub(inc, scale(k, n)));
}

constexpr float kEps32 = 5.9604645e-08f;  // 2^-24
constexpr float kTolC = 64.0f * kEps32;   // slack factor for FP32 error bounds

// ------------------------------------------------------------ mesh traversal
constexpr int32_t kTiePending = -2;  // HitRec.tri of a mesh hit whose tie is not yet resolved

// Per-ray FP32 state for a mesh: origin relative to the mesh centre, direction, guarded
// reciprocal and the per-axis slack that covers rounding of the origin to f32.
struct MeshRay {
    float ox, oy, oz, dx, dy, dz, ix, iy, iz;
    float sx, sy, sz;  // slack in t per axis
};

RT_HD MeshRay make_mesh_ray(const MeshDev& M, const RayD& r) {
    MeshRay q;
    q.ox = (float)(r.o.x - M.cx);
    q.oy = (float)(r.o.y - M.cy);
    q.oz = (float)(r.o.z - M.cz);
    q.dx = (float)r.d.x;
    q.dy = (float)r.d.y;
    q.dz = (float)r.d.z;
    // a zero component gets a tiny stand-in so 0 * inf never makes a NaN slab bound
    q.ix = 1.0f / (fabsf(q.dx) < 1e-30f ? copysignf(1e-30f, q.dx) : q.dx);
    q.iy = 1.0f / (fabsf(q.dy) < 1e-30f ? copysignf(1e-30f, q.dy) : q.dy);
    q.iz = 1.0f / (fabsf(q.dz) < 1e-30f ? copysignf(1e-30f, q.dz) : q.dz);
    float sp = kTolC * (fabsf(q.ox) + fabsf(q.oy) + fabsf(q.oz) + M.radius);
    q.sx = sp * fabsf(q.ix);
    q.sy = sp * fabsf(q.iy);
    q.sz = sp * fabsf(q.iz);
    return q;
}

// The same FP32 ray in absolute world coordinates (for object boxes).
RT_HD MeshRay make_world_ray(const RayD& r) {
    MeshDev z{};
    return make_mesh_ray(z, r);
}

// Conservative slab test of a node's tight box, clipped to [c0, c1]; false = empty.
RT_HOT bool box_clip(const MeshRay& q, const float* bmin, const float* bmax, float& c0, float& c1) {
    float ax = (bmin[0] - q.ox) * q.ix, bx = (bmax[0] - q.ox) * q.ix;
    float ay = (bmin[1] - q.oy) * q.iy, by = (bmax[1] - q.oy) * q.iy;
    float az = (bmin[2] - q.oz) * q.iz, bz = (bmax[2] - q.oz) * q.iz;
    float lo = fmaxf(fmaxf(fminf(ax, bx) - q.sx, fminf(ay, by) - q.sy), fminf(az, bz) - q.sz);
    float hi = fminf(fminf(fmaxf(ax, bx) + q.sx, fmaxf(ay, by) + q.sy), fmaxf(az, bz) + q.sz);
    lo -= kTolC * fabsf(lo);
    hi += kTolC * fabsf(hi);
    c0 = fmaxf(c0, lo);
    c1 = fminf(c1, hi);
    return c0 <= c1;
}

// box_clip with (b - o) * i computed as fma(b, i, off) and the per-axis slack folded into
// the offsets: off = -(o * i) - s for the plane the ray enters first on that axis (bmin
// when i >= 0, else bmax) and -(o * i) + s for the other (box_offsets), so the smaller of
// the two slab distances is the entry minus the slack and the larger the exit plus it, as
// in box_clip. The roundings of o * i and of the offsets (2^-24 |o i| each) lie inside the
// slack, which is 64 times larger. Only for rays with no direction component below 1e-9
// (the mesh BVH walk): there |i| <= 1e9 and o * i stays finite. A zero component's
// stand-in reciprocal (1e30, make_mesh_ray) would overflow o * i for |o| > 3.4e8, so the
// k-d and sphere-BVH walks, which take such rays, keep box_clip's (b - o) * i.
struct BoxOffsets {
    float lo[3], hi[3];  // added to bmin * i and bmax * i
};
RT_HOT BoxOffsets box_offsets(const MeshRay& q) {
    const float nox = -(q.ox * q.ix), noy = -(q.oy * q.iy), noz = -(q.oz * q.iz);
    BoxOffsets b;
    b.lo[0] = q.ix >= 0.0f ? nox - q.sx : nox + q.sx;
    b.hi[0] = q.ix >= 0.0f ? nox + q.sx : nox - q.sx;
    b.lo[1] = q.iy >= 0.0f ? noy - q.sy : noy + q.sy;
    b.hi[1] = q.iy >= 0.0f ? noy + q.sy : noy - q.sy;
    b.lo[2] = q.iz >= 0.0f ? noz - q.sz : noz + q.sz;
    b.hi[2] = q.iz >= 0.0f ? noz + q.sz : noz - q.sz;
    return b;
}
RT_HOT bool box_clip_fma(const MeshRay& q, const BoxOffsets& o, const float* bmin, const float* bmax, float& c0,
                         float& c1) {
    const float ax = fmaf(bmin[0], q.ix, o.lo[0]), bx = fmaf(bmax[0], q.ix, o.hi[0]);
    const float ay = fmaf(bmin[1], q.iy, o.lo[1]), by = fmaf(bmax[1], q.iy, o.hi[1]);
    const float az = fmaf(bmin[2], q.iz, o.lo[2]), bz = fmaf(bmax[2], q.iz, o.hi[2]);
    float lo = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fminf(az, bz));
    float hi = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz));
    lo = fmaf(-kTolC, fabsf(lo), lo);  // widened by 64 ulps: one rounding less than box_clip's form
    hi = fmaf(kTolC, fabsf(hi), hi);
    c0 = fmaxf(c0, lo);
    c1 = fminf(c1, hi);
    return c0 <= c1;
}

// ray_intersects_bounds (src/mesh.rs:218-255) in f64 on a node's exact bounds.
RT_HD bool line_hits_box64(const RayD& r, const double* b) {
    double t0 = -INFINITY, t1 = INFINITY;
    for (int a = 0; a < 3; ++a) {
        double da = a == 0 ? r.d.x : (a == 1 ? r.d.y : r.d.z);
        double oa = a == 0 ? r.o.x : (a == 1 ? r.o.y : r.o.z);
        if (fabs(da) < 1e-9) {
            if (oa < b[a] || oa > b[3 + a]) return false;
        } else {
            double inv = 1.0 / da;
            double ta = (b[a] - oa) * inv, tb = (b[3 + a] - oa) * inv;
            if (ta > tb) {
                double x = ta;
                ta = tb;
                tb = x;
            }
            t0 = fmax_r(t0, ta);
            t1 = fmin_r(t1, tb);
            if (t0 > t1) return false;
        }
    }
    return t1 >= 0.0;
}

// FP32 Moller-Trumbore with error bounds on TriRec `rec`: false when the triangle
// certainly has no hit with t in [lo, hi] (the same bounds the closest-hit walk uses).
RT_HD bool tri_may_hit_in(const MeshRay& q, float R, const float* recs, uint32_t rec, float lo, float hi) {
    const float4* rp = reinterpret_cast<const float4*>(recs) + (size_t)rec * 3;
    const float4 r0 = rp[0], r1 = rp[1], r2 = rp[2];
    const float e1x = r0.w, e1y = r1.x, e1z = r1.y, e2x = r1.z, e2y = r1.w, e2z = r2.x, E = r2.z;
    const float hx = q.dy * e2z - q.dz * e2y, hy = q.dz * e2x - q.dx * e2z, hz = q.dx * e2y - q.dy * e2x;
    const float a = e1x * hx + e1y * hy + e1z * hz;
    const float sx = q.ox - r0.x, sy = q.oy - r0.y, sz = q.oz - r0.z;
    const float S = fabsf(sx) + fabsf(sy) + fabsf(sz) + R;
    if (!(fabsf(a) > 4e-8f + kTolC * E * E)) return true;  // near-parallel: no bound, keep it
    const float f = 1.0f / a, g = E * fabsf(f);
    const float du = kTolC * (S + E) * g + 1e-7f;
    const float u = f * (sx * hx + sy * hy + sz * hz);
    if (u < -du || u > 1.0f + du) return false;
    const float qx = sy * e1z - sz * e1y, qy = sz * e1x - sx * e1z, qz = sx * e1y - sy * e1x;
    const float v = f * (q.dx * qx + q.dy * qy + q.dz * qz);
    if (v < -du || u + v > 1.0f + 2.0f * du) return false;
    const float t = f * (e2x * qx + e2y * qy + e2z * qz);
    const float dt = kTolC * E * (S + fabsf(t)) * g + kTolC * fabsf(t) + 1e-7f;
    return !(t + dt < lo || t - dt > hi);
}

// Exact tie-break: several triangles of this mesh hit at exactly `best`. The reference
// keeps the first one its traversal reaches (strict `<`, src/ray.rs:633), so walk the
// tree in the reference's order (src/mesh.rs:399-465: origin side first, far side when
// t_split >= 0, leaves gated by their line-vs-box test) and return the first triangle
// whose exact t equals `best`. Subtrees whose (conservative FP32) box interval does not
// hold `best` are skipped: a triangle hit at `best` lies inside the box of every node that
// holds it, so the walk still reaches the winner first. Inside a leaf, an FP32 bound
// keeps the f64 test for the triangles that may hit at `best`. Runs only for tied rays.
RT_HD int32_t mesh_resolve_tie(const MeshDev& M, const RayD& r, double tmin, double best, StackRef stk) {
    const double rox = r.o.x, roy = r.o.y, roz = r.o.z, rdx = r.d.x, rdy = r.d.y, rdz = r.d.z;
    const MeshRay q = make_mesh_ray(M, r);
    const float blo = (float)best * (1.0f - kTolC) - kTolC, bhi = (float)best * (1.0f + kTolC) + kTolC;
    uint32_t node = M.root_kd;
    int sp = 0;
    for (;;) {
        const KdNode& nd = M.nodes[node];
        float c0 = blo, c1 = bhi;
        if (box_clip(q, nd.bmin, nd.bmax, c0, c1)) {
            const uint32_t ax = nd.b & 3u;
            if (ax != kAxisLeaf) {
                double oa = ax == 0 ? rox : (ax == 1 ? roy : roz);
                double da = ax == 0 ? rdx : (ax == 1 ? rdy : rdz);
                bool left_near = oa <= nd.split;
                uint32_t nearc = left_near ? node + 1 : nd.a;
                uint32_t farc = left_near ? nd.a : node + 1;
                bool far_ok = !(fabs(da) < 1e-9) && ((oa == nd.split) || ((oa < nd.split) == (da > 0.0)));
                if (far_ok) stk.node[(sp++) * stk.stride] = farc;
                node = nearc;
                continue;
            }
            if (line_hits_box64(r, M.nb64 + (size_t)node * 6)) {
                uint32_t first = nd.a, cnt = nd.b >> 2;
                for (uint32_t k = 0; k < cnt; ++k) {
                    if (!tri_may_hit_in(q, M.radius, M.recs, first + k, blo, bhi)) continue;
                    uint32_t ti = f32_to_bits(M.recs[(size_t)(first + k) * 12 + 9]);
                    double tt, uu, vv;
                    d3 nn;
                    if (tri_exact(M.tri64 + (size_t)ti * 9, r, tmin, INFINITY, tt, nn, uu, vv) && tt == best)
                        return (int32_t)ti;
                }
            }
        }
        if (sp == 0) return -1;
        node = stk.node[(--sp) * stk.stride];
    }
}

// Per-ray state of a k-d walk over one mesh.
//
// Visit order and child selection are the reference's (src/mesh.rs:399-465), decided in
// f64: origin side first (origin <= split -> left), far side only when t_split >= 0
// (sign test, no division), only the origin side when |d_axis| < 1e-9. On top of that,
// FP32 [t0, t1] intervals with conservative slack skip subtrees that cannot hold a hit
// before `best`: the split plane clips the interval (except below a node whose child
// received a triangle by the empty-side rule, kFlagMisplaced) and every node's tight
// box clips it again. Skipped subtrees hold no triangle the reference could accept
// ahead of the current best, so the candidate set and its visit order are the
// reference's.
struct MeshWalk {
    MeshRay q;
    float R;        // max |v - centre|_1 over the mesh (plus the re-based origin's error)
    float tshift;   // 0, or the distance along the ray of the re-based origin (walk_init)
    float tmin_lo;  // lower bound of tmin in f32, in the walk's frame (distances from q's origin)
    float best_hi;  // upper bound of best in f32, in the walk's frame
    float c0, c1;   // interval of the current node
    uint32_t node;
    int sp;
    bool have_node, par0, par1, par2;
    bool bvh;  // walking the mesh BVH (else the k-d tree)
    uint32_t cmask;  // BVH: children of `node` still to visit
};
constexpr uint32_t kNoLeaf = 0xffffffffu;

// The mesh bounds test of MeshObject::hit (src/ray.rs:586-610) in f64 on the exact box,
// with the interval [tmin, tmax] widened by a relative 2^-30: false only when no point of
// the box lies on the ray within [tmin, tmax], so no triangle can be hit there. `enter`:
// where the ray's part within [tmin, tmax] enters the box (at least tmin).
RT_HD bool mesh_bounds_may_hit(const MeshDev& M, const RayD& r, double tmin, double tmax, double& enter) {
    double lo = tmin, hi = tmax;
    const double o[3] = {r.o.x, r.o.y, r.o.z}, d[3] = {r.d.x, r.d.y, r.d.z};
    for (int a = 0; a < 3; ++a) {
        const double inv = 1.0 / d[a];
        double t0 = (M.bmin64[a] - o[a]) * inv, t1 = (M.bmax64[a] - o[a]) * inv;
        if (inv < 0.0) {
            const double x = t0;
            t0 = t1;
            t1 = x;
        }
        lo = fmax_r(lo, t0);  // a NaN bound (0 * inf) leaves the interval as it is
        hi = fmin_r(hi, t1);
    }
    enter = lo;
    if (lo <= hi) return true;
    const double slack = 0x1p-30 * (fabs(lo) + fabs(hi)) + 1e-300;
    return slack < INFINITY && lo <= hi + slack;  // an infinite bound: the ray never meets a slab
}

// Origins farther than this many mesh radii (L1) from the centre take the f64 bounds test
// first, and a BVH walk that goes on starts from where the ray enters the box: the FP32
// walk's slack grows with |origin| and would admit most of the mesh's nodes (a ray leaving
// a far point of a plane toward a light above the scene).
constexpr float kFarOrigin = 256.0f;

// The FP32 ray re-based to the point ts along r (computed in f64): distances in the walk
// are then t - ts. That point lies off the ray by a few ulps of |o| + ts, which the box
// slack and the triangle tests' error scale (R) absorb.
RT_HD MeshRay make_mesh_ray_at(const MeshDev& M, const RayD& r, double ts, float& R) {
    RayD r2;
    r2.o = add(r.o, scale(ts, r.d));
    r2.d = r.d;
    MeshRay q = make_mesh_ray(M, r2);
    const float e = (float)(0x1p-48 * (fabs(r.o.x) + fabs(r.o.y) + fabs(r.o.z) + ts));
    q.sx += e * fabsf(q.ix);
    q.sy += e * fabsf(q.iy);
    q.sz += e * fabsf(q.iz);
    R = M.radius + e / kTolC;
    return q;
}
// Where a walk's node intervals start: below tmin_lo (tmin > 0: half of it), or at tmin_lo
// in a re-based walk's frame, where it is negative.
RT_HOT float walk_start(const MeshWalk& s) { return s.tshift == 0.0f ? s.tmin_lo * 0.5f : s.tmin_lo; }
// Bounds of f64 distance t in a re-based walk's frame (t - tshift): above and below it.
RT_HD float shift_up(float tshift, double t) {
    const float x = (float)(t - (double)tshift);
    return x + kTolC * fabsf(x) + kTolC;
}
RT_HD float shift_down(float tshift, double t) {
    const float x = (float)(t - (double)tshift);
    return x - kTolC * fabsf(x) - kTolC;
}

// kRebase: a BVH walk from a far origin starts where the ray enters the mesh's box (shadow
// walks; closest-hit walks keep the origin, which spares their registers).
template <bool kRebase>
RT_HOT void walk_init(MeshWalk& s, const MeshDev& M, const RayD& r, bool use_kd, double tmin, double best) {
    s.q = make_mesh_ray(M, r);
    s.R = M.radius;
    s.par0 = fabs(r.d.x) < 1e-9;
    s.par1 = fabs(r.d.y) < 1e-9;
    s.par2 = fabs(r.d.z) < 1e-9;
    s.node = use_kd ? M.root_kd : M.root_bf;
    s.have_node = true;
    s.sp = 0;
    s.tshift = 0.0f;
    if (fabsf(s.q.ox) + fabsf(s.q.oy) + fabsf(s.q.oz) > kFarOrigin * M.radius) {
        double enter;
        if (!mesh_bounds_may_hit(M, r, tmin, best, enter))
            s.have_node = false;
        else if (kRebase && use_kd && M.bvh != nullptr && !s.par0 && !s.par1 && !s.par2 && enter > tmin)
            s.tshift = (float)enter;  // the BVH walk (below) starts at the box
        if (s.tshift != 0.0f) s.q = make_mesh_ray_at(M, r, (double)s.tshift, s.R);
    }
    if (s.tshift == 0.0f) {
        s.tmin_lo = (float)tmin * (1.0f - kTolC);
        s.best_hi = (float)best * (1.0f + kTolC) + kTolC;
    } else {
        s.tmin_lo = shift_down(s.tshift, tmin);
        s.best_hi = shift_up(s.tshift, best);
    }
    // the mesh AABB check (src/ray.rs:586-610) is the root's box test
    s.c0 = walk_start(s);
    s.c1 = s.best_hi;
    // The BVH gives the reference's candidates for a ray with no axis-parallel component:
    // the reference's walk then reaches every leaf cell the ray's forward part crosses,
    // and a triangle hit at q lies in the leaf of q's cell. With a component below 1e-9 the
    // reference visits only the origin's side of a split (src/mesh.rs:425-428) and can
    // miss a hit; meshes where the empty-side rule moved a triangle out of its cell
    // (src/mesh.rs:174-178) have no BVH (M.bvh null). Those take the k-d walk.
    s.bvh = use_kd && M.bvh != nullptr && !s.par0 && !s.par1 && !s.par2;
    if (s.bvh) s.node = 0;
    s.cmask = 0xfu;
}

// Walks inner nodes until it holds the next leaf whose interval can still beat `best`
// (returned, with the interval start in leaf_c0) or has no node left (kNoLeaf).
template <bool count>
RT_HOT uint32_t walk_to_leaf(MeshWalk& s, const MeshDev& M, const RayD& r, StackRef stk, Work& w, float& leaf_c0) {
    // selects on values, not on member addresses, keep the ray in registers
    const double rox = r.o.x, roy = r.o.y, roz = r.o.z, rdx = r.d.x, rdy = r.d.y, rdz = r.d.z;
    uint32_t leaf = kNoLeaf;
    while (s.have_node && leaf == kNoLeaf) {
        bool next = false;  // `node` already holds the next node to visit
        if (s.c0 <= s.best_hi) {
            const float4* np = reinterpret_cast<const float4*>(M.nodes + s.node);
            float4 n0 = ldg(np);
            float4 n1 = ldg(np + 1);
            if (count) w.v[kWorkNodes]++;
            const float bmin[3] = {n0.x, n0.y, n0.z}, bmax[3] = {n0.w, n1.x, n1.y};
            uint32_t na = f32_to_bits(n1.z), nb = f32_to_bits(n1.w);
            uint32_t ax = nb & 3u;
            if (box_clip(s.q, bmin, bmax, s.c0, s.c1) && s.c0 <= s.best_hi) {
                if (ax == kAxisLeaf) {
                    leaf = s.node;
                    leaf_c0 = s.c0;
                } else {
                    double split = ldg(&M.nodes[s.node].split);
                    double oa = ax == 0 ? rox : (ax == 1 ? roy : roz);
                    double da = ax == 0 ? rdx : (ax == 1 ? rdy : rdz);
                    float ia = ax == 0 ? s.q.ix : (ax == 1 ? s.q.iy : s.q.iz);
                    bool par = ax == 0 ? s.par0 : (ax == 1 ? s.par1 : s.par2);
                    bool left_near = oa <= split;
                    uint32_t nearc = left_near ? s.node + 1 : na;
                    uint32_t farc = left_near ? na : s.node + 1;
                    bool far_ok = !par && ((oa == split) || ((oa < split) == (da > 0.0)));
                    next = true;
                    if (!far_ok) {
                        s.node = nearc;
                    } else if (nb & kFlagMisplaced) {
                        stk.node[s.sp * stk.stride] = farc;
                        stk.t0[s.sp * stk.stride] = s.c0;
                        ++s.sp;
                        s.node = nearc;
                    } else {
                        float ts = (float)(split - oa) * ia;
                        float tol = kTolC * fabsf(ts) + 1e-30f;
                        bool dn = s.c0 <= ts + tol;
                        bool dfar = ts - tol <= s.c1;
                        if (dn && dfar) {
                            stk.node[s.sp * stk.stride] = farc;
                            stk.t0[s.sp * stk.stride] = fmaxf(s.c0, ts - tol);
                            ++s.sp;
                            s.node = nearc;
                            s.c1 = fminf(s.c1, ts + tol);
                        } else if (dn) {
                            s.node = nearc;
                            s.c1 = fminf(s.c1, ts + tol);
                        } else if (dfar) {
                            s.node = farc;
                            s.c0 = fmaxf(s.c0, ts - tol);
                        } else {
                            next = false;
                        }
                    }
                }
            }
        }
        if (!next) {
            // pop the next pending subtree that can still beat `best`
            s.have_node = false;
            while (s.sp > 0) {
                --s.sp;
                s.node = stk.node[s.sp * stk.stride];
                s.c0 = stk.t0[s.sp * stk.stride];
                if (s.c0 <= s.best_hi) {
                    s.have_node = true;
                    break;
                }
            }
            s.c1 = s.best_hi;  // the popped node's own box clips the far end again
        }
    }
    return leaf;
}

// Tests a leaf's triangles: FP32 Moller-Trumbore with error bounds; survivors re-tested
// in f64 (intersect_triangle, src/ray.rs:527-583). kAny: true on a hit with
// tmin < t < best. Otherwise improves best / best_tri strictly (first visited wins ties)
// and flags `tie` when another triangle hits at exactly `best`.
RT_HOT float f4c(const float4& v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w)); }
// 1 / x for the FP32 prefilter: the hardware reciprocal on the device (1 ulp; the
// triangle test's bounds allow 64), a division on the host.
RT_HOT float rcp_f32(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
RT_HOT uint32_t u4c(const uint4& v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w)); }

// Walks the 4-wide mesh BVH until it holds the next leaf whose interval can still beat
// `best` (true: TriRecs [first, first + cnt) of M.bvh_recs, the interval start in
// leaf_c0), or has no node left (false). At a node the children in s.cmask are tested;
// the nearest hit one is taken and, when others were hit, one stack entry keeps the node,
// the mask of those others and the least of their entry t (re-tested when popped).
template <bool count>
RT_HOT bool bvh_to_leaf(MeshWalk& s, const MeshDev& M, StackRef stk, Work& w, uint32_t& first, uint32_t& cnt,
                        float& leaf_c0) {
    static_assert(kMeshBvhDepth / 2 < kMaxKdDepth, "a 4-wide walk's entries fit the per-lane stack");
    const BoxOffsets off = box_offsets(s.q);
    const float t0 = walk_start(s);
    while (s.have_node) {
        const uint32_t nidx = s.node;
        const float4* np = reinterpret_cast<const float4*>(M.bvh + nidx);
        const float4 lx = ldg(np), ly = ldg(np + 1), lz = ldg(np + 2), hx = ldg(np + 3), hy = ldg(np + 4),
                     hz = ldg(np + 5);
        const uint4 lk = ldg(reinterpret_cast<const uint4*>(np) + 6);
        if (count) w.v[kWorkBvhNodes]++;
        float t[4];
        uint32_t hit = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float lo[3] = {f4c(lx, i), f4c(ly, i), f4c(lz, i)}, hi[3] = {f4c(hx, i), f4c(hy, i), f4c(hz, i)};
            float a = t0, b = s.best_hi;
            const bool h = box_clip_fma(s.q, off, lo, hi, a, b);
            t[i] = a;
            hit |= (h && ((s.cmask >> i) & 1u) && u4c(lk, i) != kMeshNone) ? 1u << i : 0u;
        }
        // the nearest hit child (the first of equal entries) and the least entry of the others
        int bi = -1;
        float bt = INFINITY, tr = INFINITY;
#pragma unroll
        for (int i = 0; i < 4; ++i) {  // selects, no branches (entries are never NaN)
            const bool h = ((hit >> i) & 1u) != 0;
            const bool nearer = h & ((bi < 0) | (t[i] < bt));
            const float joins = nearer ? bt : t[i];  // the entry that goes to the others
            tr = (h & (joins < tr)) ? joins : tr;
            bt = nearer ? t[i] : bt;
            bi = nearer ? i : bi;
        }
        if (bi >= 0) {
            // the other hit children as one entry: written in any case (a node at 4-wide
            // level L has at most L entries below it, L < kMeshBvhDepth / 2, so slot sp is
            // inside the stack) and kept when there are any
            const uint32_t rest = hit & ~(1u << bi);
            stk.node[s.sp * stk.stride] = nidx << 4 | rest;
            stk.t0[s.sp * stk.stride] = tr;
            s.sp += rest ? 1 : 0;
            const uint32_t link = bi == 0 ? lk.x : (bi == 1 ? lk.y : (bi == 2 ? lk.z : lk.w));
            if (!(link & kMeshLeaf)) {
                s.node = link;
                s.cmask = 0xfu;
                continue;
            }
            first = (link & ~kMeshLeaf) >> 6;
            cnt = link & 63u;
            leaf_c0 = bt;
        }
        // next: the stack's last entry that can still beat `best`
        s.have_node = false;
        while (s.sp > 0) {
            --s.sp;
            const uint32_t e = stk.node[s.sp * stk.stride];
            if (stk.t0[s.sp * stk.stride] <= s.best_hi) {
                s.node = e >> 4;
                s.cmask = e & 0xfu;
                s.have_node = true;
                break;
            }
        }
        if (bi >= 0) return true;  // the leaf taken above
    }
    return false;
}

template <bool count>
RT_HOT bool next_leaf(MeshWalk& s, const MeshDev& M, const RayD& r, StackRef stk, Work& w, uint32_t& first,
                      uint32_t& cnt, float& leaf_c0) {
    if (s.bvh) return bvh_to_leaf<count>(s, M, stk, w, first, cnt, leaf_c0);
    const uint32_t leaf = walk_to_leaf<count>(s, M, r, stk, w, leaf_c0);
    if (leaf == kNoLeaf) return false;
    first = ldg(&M.nodes[leaf].a);
    cnt = ldg(&M.nodes[leaf].b) >> 2;
    return true;
}

constexpr bool kDeferRetests = true;

template <bool kAny, bool count>
RT_HOT bool leaf_test(MeshWalk& s, const MeshDev& M, const RayD& r, uint32_t first, uint32_t cnt, double tmin,
                     double& best, int32_t& best_tri, bool& found, bool& tie, Work& w) {
    const float4* recs = reinterpret_cast<const float4*>(s.bvh ? M.bvh_recs : M.recs);
    const MeshRay& q = s.q;
    // closest hit: the f64 re-test of a candidate improves best / best_tri strictly, in visit order
    auto retest = [&](uint32_t ti) {
        double tt, uu, vv;
        d3 nn;
        if (count) w.v[kWorkRetests]++;
        if (tri_exact(M.tri64 + (size_t)ti * 9, r, tmin, INFINITY, tt, nn, uu, vv)) {
            if (tt < best) {
                best = tt;
                best_tri = (int32_t)ti;
                s.best_hi = s.tshift == 0.0f ? (float)best * (1.0f + kTolC) + kTolC : shift_up(s.tshift, best);
                found = true;
                tie = false;
            } else if (tt == best && found && (int32_t)ti != best_tri) {
                tie = true;
            }
        }
    };
    // Closest hit defers the re-tests of up to two candidates to the end of the leaf (or to
    // a third candidate), so the lanes of a wavefront re-test together instead of one
    // triangle index at a time. Candidates are re-tested in visit order; a candidate the
    // FP32 bound would have dropped against the improved best cannot reach t <= best.
    uint32_t c0 = 0, c1 = 0;
    int nc = 0;
    for (uint32_t k = 0; k < cnt; ++k) {
        float4 r0 = ldg(recs + (size_t)(first + k) * 3 + 0);
        float4 r1 = ldg(recs + (size_t)(first + k) * 3 + 1);
        float4 r2 = ldg(recs + (size_t)(first + k) * 3 + 2);
        if (count) w.v[kWorkTris]++;
        float e1x = r0.w, e1y = r1.x, e1z = r1.y;
        float e2x = r1.z, e2y = r1.w, e2z = r2.x;
        float E = r2.z;  // max(|e1|_1, |e2|_1)
        // FP32 with fused multiply-adds (one rounding fewer per step than the bounds allow for)
        const float hx = fmaf(q.dy, e2z, -(q.dz * e2y)), hy = fmaf(q.dz, e2x, -(q.dx * e2z));
        const float hz = fmaf(q.dx, e2y, -(q.dy * e2x));
        const float a = fmaf(e1x, hx, fmaf(e1y, hy, e1z * hz));
        float sx = q.ox - r0.x, sy = q.oy - r0.y, sz = q.oz - r0.z;
        float S = fabsf(sx) + fabsf(sy) + fabsf(sz) + s.R;
        float fa = fabsf(a);
        bool sure_in = false;
        bool maybe = true;
        if (fa > 4e-8f + kTolC * E * E) {
            const float f = rcp_f32(a);
            float g = E * fabsf(f);
            float du = kTolC * (S + E) * g + 1e-7f;
            float u = f * fmaf(sx, hx, fmaf(sy, hy, sz * hz));
            if (u < -du || u > 1.0f + du) {
                maybe = false;
            } else {
                const float qx = fmaf(sy, e1z, -(sz * e1y)), qy = fmaf(sz, e1x, -(sx * e1z));
                const float qz = fmaf(sx, e1y, -(sy * e1x));
                float v = f * fmaf(q.dx, qx, fmaf(q.dy, qy, q.dz * qz));
                if (v < -du || u + v > 1.0f + 2.0f * du) {
                    maybe = false;
                } else {
                    float t = f * fmaf(e2x, qx, fmaf(e2y, qy, e2z * qz));
                    float dt = kTolC * E * (S + fabsf(t)) * g + kTolC * fabsf(t) + 1e-7f;
                    if (t + dt <= s.tmin_lo || t - dt >= s.best_hi) {
                        maybe = false;
                    } else if (kAny) {
                        const float tmin_up = s.tshift == 0.0f ? (float)tmin * (1.0f + kTolC) : shift_up(s.tshift, tmin);
                        const float best_dn = s.tshift == 0.0f ? (float)best * (1.0f - kTolC) : shift_down(s.tshift, best);
                        sure_in = u >= du && v >= du && u + v <= 1.0f - 2.0f * du && t - dt > tmin_up && t + dt < best_dn;
                    }
                }
            }
        }
        if (!maybe) continue;
        if (kAny && sure_in) return true;
        uint32_t ti = f32_to_bits(r2.y);
        if (kAny) {
            double tt, uu, vv;
            d3 nn;
            if (count) w.v[kWorkRetests]++;
            if (tri_exact(M.tri64 + (size_t)ti * 9, r, tmin, best, tt, nn, uu, vv)) return true;
        } else if (!kDeferRetests) {
            retest(ti);
        } else {
            if (nc == 2) {
                retest(c0);
                retest(c1);
                nc = 0;
            }
            if (nc == 0) c0 = ti;
            else c1 = ti;
            ++nc;
        }
    }
    if (!kAny && nc > 0) retest(c0);
    if (!kAny && nc > 1) retest(c1);
    return false;
}

// kAny = shadow (any hit with tmin < t < tmax); otherwise closest hit improving `best`
// strictly (t < best, first visited wins ties, src/ray.rs:620-649).
//
// "while-while" traversal: each lane walks inner nodes until it holds its next leaf,
// then the wavefront tests leaves together, so lanes at inner nodes do not wait through
// other lanes' triangle loops. Leaves are still tested in this lane's visit order.
template <bool kAny, bool count>
RT_HD bool mesh_trace(const MeshDev& M, const RayD& r, bool use_kd, double tmin, double& best, int32_t& best_tri,
                      StackRef stk, Work& w) {
    MeshWalk s;
    walk_init<kAny>(s, M, r, use_kd, tmin, best);
    bool found = false;
    bool tie = false;  // another triangle of this mesh hit at exactly `best`
    for (;;) {
        float leaf_c0 = 0.0f;
        uint32_t first = 0, cnt = 0;
        if (!next_leaf<count>(s, M, r, stk, w, first, cnt, leaf_c0)) break;  // this lane has no node left
        if (leaf_c0 > s.best_hi) continue;  // `best` improved since the leaf was found
        if (leaf_test<kAny, count>(s, M, r, first, cnt, tmin, best, best_tri, found, tie, w)) return true;
    }
    // a tie between triangles of this mesh: the reference-order walk (mesh_resolve_tie)
    // picks the winner in a separate pass; kTiePending marks the hit as pending
    if (!kAny && tie && use_kd) best_tri = kTiePending;
    return found;
}

// ------------------------------------------------------------ World::hit
// Closest hit over all objects in scene order (src/ray.rs:699-711): analytic objects
// accept t == closest (and so replace an equal earlier hit), meshes need t < closest.
// A cube's test runs in its local space and returns a local t (src/ray.rs:413-421), so its
// prefilter is the line against the cube's world box, with no t range.
RT_HD bool cube_may_hit(const MeshRay& wq, const ObjDev& ob) {
    float lo = -3.0e38f, hi = 3.0e38f;
    return box_clip(wq, ob.bmin, ob.bmax, lo, hi);
}

// ------------------------------------------------------------ object BVH
// Visit the spheres whose boxes the ray meets within [tmin, hi()], near child first, and
// call fn(object index) for each; fn returns true to stop. Uses the LDS stack.
template <class Hi, class Fn>
RT_HD void obvh_visit(const SceneView& S, const MeshRay& wq, double tmin, Hi hi, StackRef stk, Fn fn) {
    uint32_t node = 0;
    int sp = 0;
    float c0 = (float)tmin * 0.5f, c1 = hi();
    for (;;) {
        bool next = false;
        if (c0 <= hi()) {
            const BvhNode& nd = S.obvh[node];
            if (box_clip(wq, nd.bmin, nd.bmax, c0, c1) && c0 <= hi()) {
                if (nd.b & kBvhLeaf) {
                    uint32_t first = nd.a, cnt = nd.b & ~kBvhLeaf;
                    for (uint32_t k = 0; k < cnt; ++k)
                        if (fn(S.obvh_items[first + k])) return;
                } else {
                    // near child first by the entry of its box
                    const BvhNode& L = S.obvh[nd.a];
                    const BvhNode& R = S.obvh[nd.b];
                    float l0 = c0, l1 = hi(), r0 = c0, r1 = hi();
                    bool hl = box_clip(wq, L.bmin, L.bmax, l0, l1);
                    bool hr = box_clip(wq, R.bmin, R.bmax, r0, r1);
                    if (hl && hr) {
                        bool left_first = l0 <= r0;
                        stk.node[sp * stk.stride] = left_first ? nd.b : nd.a;
                        stk.t0[sp * stk.stride] = left_first ? r0 : l0;
                        ++sp;
                        node = left_first ? nd.a : nd.b;
                        c0 = left_first ? l0 : r0;
                        next = true;
                    } else if (hl || hr) {
                        node = hl ? nd.a : nd.b;
                        c0 = hl ? l0 : r0;
                        next = true;
                    }
                    c1 = hi();
                }
            }
        }
        if (next) continue;
        bool popped = false;
        while (sp > 0) {
            --sp;
            node = stk.node[sp * stk.stride];
            c0 = stk.t0[sp * stk.stride];
            if (c0 <= hi()) {
                popped = true;
                break;
            }
        }
        if (!popped) return;
        c1 = hi();
    }
}

// World::hit in any object order (spheres through the BVH, the rest linearly): among the
// objects reaching the smallest t, the reference's scene-order fold (analytic objects
// accept t == closest and replace, meshes need t < closest) keeps the largest analytic
// index above the first object reaching it, else that first object. Spheres, planes and
// meshes have hit distances independent of the running `closest`, so the rule is exact
// for them; scenes with cubes use the scene-order fold below instead.
struct ClosestState {
    double best;
    double tmax;
    int32_t first, first_tri, max_analytic;
    RT_HD void offer(double t, int32_t i, int32_t tri, bool analytic) {
        if (!(analytic ? t <= tmax : t < tmax)) return;
        if (first < 0 || t < best) {
            best = t;
            first = i;
            first_tri = tri;
            max_analytic = analytic ? i : -1;
        } else if (t == best) {
            if (i < first) {
                first = i;
                first_tri = tri;
            }
            if (analytic && i > max_analytic) max_analytic = i;
        }
    }
};

template <bool count>
RT_HD bool world_closest_bvh(const SceneView& S, const RayD& r, double tmin, double tmax, bool use_kd, StackRef stk,
                             Work& w, double& t_out, int32_t& obj_out, int32_t& tri_out) {
    ClosestState cs{tmax, tmax, -1, -1, -1};
    RayF f = make_rayf(r);
    const MeshRay wq = make_world_ray(r);
    auto hi = [&]() { return (float)cs.best * (1.0f + kTolC) + kTolC; };
    obvh_visit(S, wq, tmin, hi, stk, [&](int32_t i) {
        const ObjDev& ob = S.objs[i];
        double t;
        if (count) w.v[kWorkSphere]++;
        if (sphere_may_hit(ob, r, f, tmin, cs.best) && sphere_exact(ob, r, tmin, INFINITY, t))
            cs.offer(t, i, -1, true);
        return false;
    });
    for (int k = 0; k < S.n_others; ++k) {
        int32_t i = S.others[k];
        const ObjDev& ob = S.objs[i];
        double t;
        if (ob.kind == kPlane) {
            if (count) w.v[kWorkPlane]++;
            if (plane_may_hit(ob, f, tmin, INFINITY) && plane_exact(ob, r, tmin, INFINITY, t)) cs.offer(t, i, -1, true);
        } else {
            int32_t tri = -1;
            double best = cs.best;
            if (mesh_trace<false, count>(S.meshes[ob.mesh], r, use_kd, tmin, best, tri, stk, w))
                cs.offer(best, i, tri, false);
        }
    }
    if (cs.first < 0) return false;
    bool analytic_wins = cs.max_analytic > cs.first;
    obj_out = analytic_wins ? cs.max_analytic : cs.first;
    tri_out = analytic_wins ? -1 : cs.first_tri;
    t_out = cs.best;
    return true;
}

template <bool count>
RT_HD bool world_closest(const SceneView& S, const RayD& r, double tmin, double tmax, bool use_kd, StackRef stk,
                         Work& w, double& t_out, int32_t& obj_out, int32_t& tri_out) {
    if (S.obvh_closest) return world_closest_bvh<count>(S, r, tmin, tmax, use_kd, stk, w, t_out, obj_out, tri_out);
    double closest = tmax;
    int32_t bo = -1, bt = -1;
    RayF f = make_rayf(r);
    const MeshRay wq = make_world_ray(r);
    for (int i = 0; i < S.n_obj; ++i) {
        const ObjDev& ob = S.objs[i];
        double t;
        if (ob.kind == kSphere) {
            if (count) w.v[kWorkSphere]++;
            if (sphere_may_hit(ob, r, f, tmin, closest) && sphere_exact(ob, r, tmin, closest, t)) {
                closest = t;
                bo = i;
                bt = -1;
            }
        } else if (ob.kind == kPlane) {
            if (count) w.v[kWorkPlane]++;
            if (plane_may_hit(ob, f, tmin, closest) && plane_exact(ob, r, tmin, closest, t)) {
                closest = t;
                bo = i;
                bt = -1;
            }
        } else if (ob.kind == kCube) {
            d3 p, n;
            bool may = cube_may_hit(wq, ob);
            if (count && may) w.v[kWorkCube]++;
            if (may && cube_exact(ob, r, tmin, closest, t, p, n)) {
                closest = t;
                bo = i;
                bt = -1;
            }
        } else {
            int32_t tri = -1;
            double best = closest;
            if (mesh_trace<false, count>(S.meshes[ob.mesh], r, use_kd, tmin, best, tri, stk, w)) {
                closest = best;
                bo = i;
                bt = tri;
            }
        }
    }
    t_out = closest;
    obj_out = bo;
    tri_out = bt;
    return bo >= 0;
}

// Shadow query: World::hit(shadow_ray, 0.001, dist).is_some() (src/lighting.rs:120,185).
// The scene-order fold finds a hit exactly when some object hits in (tmin, tmax): the
// first such object sees the full range. So the order is free, and the analytic objects
// are tested before the meshes, the costly tests.
//
// The analytic part: true when a sphere, plane or cube hits in (tmin, tmax); `meshes`
// tells whether the scene also has meshes to walk.
template <bool count>
RT_HD bool world_any_analytic(const SceneView& S, const RayD& r, double tmin, double tmax, StackRef stk, Work& w,
                              bool& meshes) {
    RayF f = make_rayf(r);
    const MeshRay wq = make_world_ray(r);
    meshes = false;
    if (S.obvh) {
        // any hit in [tmin, tmax] over the spheres, then the other objects
        bool hit = false;
        const float hi_c = (float)tmax * (1.0f + kTolC) + kTolC;
        obvh_visit(S, wq, tmin, [&]() { return hi_c; }, stk, [&](int32_t i) {
            const ObjDev& ob = S.objs[i];
            double t;
            if (count) w.v[kWorkSphere]++;
            hit = sphere_may_hit(ob, r, f, tmin, tmax) && sphere_exact(ob, r, tmin, tmax, t);
            return hit;
        });
        if (hit) return true;
        for (int k = 0; k < S.n_others; ++k) {
            const ObjDev& ob = S.objs[S.others[k]];
            double t;
            if (ob.kind == kPlane) {
                if (count) w.v[kWorkPlane]++;
                if (plane_may_hit(ob, f, tmin, tmax) && plane_exact(ob, r, tmin, tmax, t)) return true;
            } else if (ob.kind == kCube) {
                d3 p, n;
                bool may = cube_may_hit(wq, ob);
                if (count && may) w.v[kWorkCube]++;
                if (may && cube_exact(ob, r, tmin, tmax, t, p, n)) return true;
            } else {
                meshes = true;
            }
        }
        return false;
    }
    for (int i = 0; i < S.n_obj; ++i) {
        const ObjDev& ob = S.objs[i];
        double t;
        if (ob.kind == kSphere) {
            if (count) w.v[kWorkSphere]++;
            if (sphere_may_hit(ob, r, f, tmin, tmax) && sphere_exact(ob, r, tmin, tmax, t)) return true;
        } else if (ob.kind == kPlane) {
            if (count) w.v[kWorkPlane]++;
            if (plane_may_hit(ob, f, tmin, tmax) && plane_exact(ob, r, tmin, tmax, t)) return true;
        } else if (ob.kind == kCube) {
            d3 p, n;
            bool may = cube_may_hit(wq, ob);
            if (count && may) w.v[kWorkCube]++;
            if (may && cube_exact(ob, r, tmin, tmax, t, p, n)) return true;
        } else {
            meshes = true;
        }
    }
    return false;
}

// The meshes of the scene in object order: position k of the list the shadow query walks
// (the non-sphere objects with an object BVH, else all objects); -1 after the last.
RT_HD int next_mesh(const SceneView& S, int k) {
    const int n = S.obvh ? S.n_others : S.n_obj;
    for (++k; k < n; ++k)
        if (S.objs[S.obvh ? S.others[k] : k].kind == kMesh) return k;
    return -1;
}
RT_HD const MeshDev& mesh_at(const SceneView& S, int k) { return S.meshes[S.objs[S.obvh ? S.others[k] : k].mesh]; }

template <bool count>
RT_HD bool world_any(const SceneView& S, const RayD& r, double tmin, double tmax, bool use_kd, StackRef stk, Work& w) {
    bool meshes = false;
    if (world_any_analytic<count>(S, r, tmin, tmax, stk, w, meshes)) return true;
    if (!meshes) return false;
    for (int k = next_mesh(S, -1); k >= 0; k = next_mesh(S, k)) {
        int32_t tri = -1;
        double best = tmax;
        if (mesh_trace<true, count>(mesh_at(S, k), r, use_kd, tmin, best, tri, stk, w)) return true;
    }
    return false;
}

// ------------------------------------------------------------ camera
// Camera::get_ray (src/camera.rs:155-167)
RT_HD RayD camera_ray(const CameraDev& c, double u, double v) {
    d3 vp = add(add(ld3(c.llc), scale(u, ld3(c.horizontal))), scale(v, ld3(c.vertical)));
    RayD r;
    if (c.persp) {
        r.o = ld3(c.origin);
        r.d = unit(unit(sub(vp, r.o)));
    } else {
        r.o = vp;
        r.d = unit(ld3(c.view));
    }
    return r;
}

// Camera::get_grid_color (src/camera.rs:171-253)
RT_HD bool grid_color(const CameraDev& c, const RayD& r, d3& out) {
    if (c.persp || !c.has_grid) return false;
    double ht = c.thick / 2.0, p = c.pitch;
    for (int plane = 0; plane < 3; ++plane) {
        double dcomp = plane == 0 ? r.d.z : (plane == 1 ? r.d.y : r.d.x);
        double ocomp = plane == 0 ? r.o.z : (plane == 1 ? r.o.y : r.o.x);
        if (fabs(dcomp) > 1e-10) {
            double t = -ocomp / dcomp;
            if (t > 0.0) {
                d3 ip = add(r.o, scale(t, r.d));
                double a = plane == 2 ? ip.y : ip.x;
                double b = plane == 0 ? ip.y : ip.z;
                double am = fabs(fract_r(a / p)), bm = fabs(fract_r(b / p));
                double ad = fmin_r(am * p, (1.0 - am) * p);
                double bd = fmin_r(bm * p, (1.0 - bm) * p);
                if (ad <= ht || bd <= ht) {
                    out = ld3(c.grid_color);
                    return true;
                }
            }
        }
    }
    return false;
}

// ------------------------------------------------------------ raygen

// Level-0 queue slot q -> its camera ray record. Slots enumerate T x T tiles (holes outside
// image get path = -1); quincunx puts centres first, then the (W+1)x(H+1) corners.
RT_HD RayRec raygen_item(const SceneView& S, const FrameView& F, int q) {
    const int T = F.tile, T2 = T * T;  // slots run tile by tile, T x T pixels each
    RayRec rec;
    rec.level = 0;
    rec.path = -1;
    const CameraDev& cam = S.cam;
    if (F.aa == 1) {
        // render_quincunx (src/renderer.rs:752-897)
        int tiles_c = F.tiles_x * ((F.H + T - 1) / T) * T2;
        bool corner = q >= tiles_c;
        int qq = corner ? q - tiles_c : q;
        int gw = corner ? F.W + 1 : F.W, gh = corner ? F.H + 1 : F.H;
        int tx = (gw + T - 1) / T;
        int tile = qq / (T2), in = qq % (T2);
        int x = (tile % tx) * T + in % T, y = (tile / tx) * T + in / T;
        bool want = x < gw && y < gh;
        if (want && F.tmask)
            want = corner ? (pixel_in_mask(F, x - 1, y - 1) || pixel_in_mask(F, x, y - 1) ||
                             pixel_in_mask(F, x - 1, y) || pixel_in_mask(F, x, y))
                          : pixel_in_mask(F, x, y);
        if (want) {
            double pw = 1.0 / (double)F.W, ph = 1.0 / (double)F.H;
            double u, v;
            uint64_t s;
            int path;
            if (corner) {
                u = clamp_r((double)x * pw, 0.0, 1.0);
                v = clamp_r(1.0 - (double)y * ph, 0.0, 1.0);
                s = F.seed * 0x9E3779B97F4A7C15ull + (uint64_t)x + (uint64_t)y * 0x85EBCA6Bull;
                path = F.W * F.H + y * (F.W + 1) + x;
            } else {
                u = ((double)x + 0.5) * pw;
                v = 1.0 - ((double)y + 0.5) * ph;
                s = F.seed * 0x9E3779B97F4A7C15ull + (uint64_t)x * 0x85EBCA6Bull + (uint64_t)y * 0xC2B2AE35ull +
                    0x12345678ull;
                path = y * F.W + x;
            }
            RayD r = camera_ray(cam, u, v);
            st3(rec.o, r.o);
            st3(rec.d, r.d);
            rec.path = path;
            F.path_seed[path] = s;
            if (F.feat && !corner) {
                float* fp = F.feat + (size_t)path * kFeatCount;
                for (int c = 0; c < kFeatCount; ++c) fp[c] = 0.0f;
            }
        }
    } else {
        // render_standard / render_standard_with_outline (src/renderer.rs:445-749);
        // slot = sample * (tiled pixels) + tiled pixel
        int tiled = F.tiles_x * ((F.H + T - 1) / T) * T2;
        int smp = q / tiled, qq = q % tiled;
        int tile = qq / (T2), in = qq % (T2);
        int x = (tile % F.tiles_x) * T + in % T, y = (tile / F.tiles_x) * T + in / T;
        if (x < F.W && y < F.H && smp < F.N && pixel_in_mask(F, x, y)) {
            double pu = (double)x / (double)(F.W - 1);
            double pv = (double)(F.H - 1 - y) / (double)(F.H - 1);
            double pw = 1.0 / (double)(F.W - 1), ph = 1.0 / (double)(F.H - 1);
            uint64_t ps = F.seed * 0x9E3779B97F4A7C15ull + (uint64_t)x * 0x85EBCA6Bull + (uint64_t)y * 0xC2B2AE35ull;
            double su = pu, sv = pv;
            if (F.aa == 2) {
                StdRngDev rng;
                if (F.N == 1) {
                    rng.init(ps, 0);
                    double ju = rng.gen_f64() - 0.5;
                    double jv = rng.gen_f64() - 0.5;
                    su = pu + ju * pw;
                    sv = pv + jv * ph;
                } else {
                    // sample s draws u64 #2s and #2s+1 of the pixel's stream
                    rng.init(ps, (uint64_t)smp * 4);
                    double angle = 2.0 * 3.141592653589793 * (double)smp / (double)F.N;
                    double phase = rng.gen_f64() * 2.0 * 3.141592653589793;
                    double rot = angle + phase;
                    double rad = 0.5 * rng.gen_f64();
                    double ju = rad * cos(rot);
                    double jv = rad * sin(rot);
                    su = pu + ju * pw;
                    sv = pv + jv * ph;
                }
            }
            RayD r = camera_ray(cam, su, sv);
            st3(rec.o, r.o);
            st3(rec.d, r.d);
            int path = (y * F.W + x) * F.N + smp;
            rec.path = path;
            F.path_seed[path] = ps + (uint64_t)smp * 0x1F845FEDull;
            if (F.feat && smp == 0) {
                float* fp = F.feat + (size_t)(y * F.W + x) * kFeatCount;
                for (int c = 0; c < kFeatCount; ++c) fp[c] = 0.0f;
            }
        }
    }
    return rec;
}

// ------------------------------------------------------------ closest hit stage
// Returns true when the ray hit (the caller appends the slot to the hit queue);
// *tie is set when the hit waits for the tie resolver.
template <bool count>
RT_HD bool closest_item(const SceneView& S, const FrameView& F, int slot, int level, StackRef stk, Work& w, bool* tie) {
    const RayRec& rr = F.rays[slot];
    HitRec h;
    h.t = 0.0;
    h.obj = -1;
    h.tri = -1;
    *tie = false;
    if (rr.path < 0) {
        F.hits[slot] = h;
        return false;
    }
    RayD r;
    r.o = ld3(rr.o);
    r.d = ld3(rr.d);
    double t;
    int32_t obj, tri;
    bool hit = world_closest<count>(S, r, 0.001, INFINITY, F.use_kd != 0, stk, w, t, obj, tri);
    if (hit) {
        h.t = t;
        h.obj = obj;
        h.tri = tri;
        *tie = tri == kTiePending;
    }
    F.hits[slot] = h;
    if (level == 0) {
        F.prim_obj[rr.path] = hit ? obj : -1;
        F.prim_tri[rr.path] = hit ? tri : -1;
        F.prim_t[rr.path] = hit ? t : 0.0;
    }
    if (!hit) {
        // miss: background, or the ortho grid when the camera is passed (src/lighting.rs:497-505)
        d3 c = ld3(S.bg);
        if (F.grid_miss) {
            d3 g;
            if (grid_color(S.cam, r, g)) c = g;
        }
        double* L = F.lvl + ((size_t)level * F.n_paths + rr.path) * 4;
        L[0] = c.x;
        L[1] = c.y;
        L[2] = c.z;
        L[3] = -1.0;
        if (level == 0 && F.outline) {
            F.ol_dn[(size_t)rr.path * 4] = -1.0;
        }
    }
    return hit;
}

// Tie pass: the reference-order winner for a hit marked kTiePending.
RT_HD void resolve_item(const SceneView& S, const FrameView& F, int slot, int level, StackRef stk) {
    const RayRec& rr = F.rays[slot];
    HitRec& h = F.hits[slot];
    RayD r;
    r.o = ld3(rr.o);
    r.d = ld3(rr.d);
    h.tri = mesh_resolve_tie(S.meshes[S.objs[h.obj].mesh], r, 0.001, h.t, stk);
    if (level == 0) F.prim_tri[rr.path] = h.tri;
}

// ------------------------------------------------------------ area-light samples
// n x sample_disk_light_point (src/lighting.rs:69-97; the reference's n is SAMPLES = 16 at
// :158): rejection sampling with two gen_range(-r..r) draws per trial from
// StdRng::seed_from_u64(seed). Each draw is one u64 = two consecutive ChaCha12 words, so a
// trial takes exactly 4 words and a 16-word block holds exactly 4 trials; the block is
// consumed with static indices. A draw that lands on `high` (gen_range retries with a
// smaller scale, probability ~2^-52) makes the fast path give up so the sequential reader
// redoes the light.
RT_HD bool disk_samples_fast(uint64_t seed, int count, double radius, d3 lp, d3 ub, d3 vb, SptRef out) {
    uint32_t key[8];
    stdrng_key(seed, key);
    const double low = -radius, high = radius, scl = high - low;
    int n = 0;
    for (uint64_t blk = 0; n < count; ++blk) {
        uint32_t b[16];
        chacha12_block(key, blk, b);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            uint64_t wx = (uint64_t)b[4 * t] | ((uint64_t)b[4 * t + 1] << 32);
            uint64_t wy = (uint64_t)b[4 * t + 2] | ((uint64_t)b[4 * t + 3] << 32);
            double x = (bits_to_f64((wx >> 12) | (1023ull << 52)) - 1.0) * scl + low;
            double y = (bits_to_f64((wy >> 12) | (1023ull << 52)) - 1.0) * scl + low;
            if (!(x < high) || !(y < high)) return false;
            if (n < count && x * x + y * y <= radius * radius) {
                out.put(n, add(add(lp, scale(x, ub)), scale(y, vb)));
                ++n;
            }
        }
    }
    return true;
}

RT_COLD void disk_samples_slow(uint64_t seed, int count, double radius, d3 lp, d3 ub, d3 vb, SptRef out) {
    StdRngDev rng;
    rng.init(seed, 0);
    for (int s = 0; s < count; ++s) {
        double x, y;
        for (;;) {  // sample_disk_point (src/lighting.rs:57-66)
            x = rng.gen_range(-radius, radius);
            y = rng.gen_range(-radius, radius);
            if (x * x + y * y <= radius * radius) break;
        }
        out.put(s, add(add(lp, scale(x, ub)), scale(y, vb)));
    }
}

// The per-hit light seed and the light-facing basis of calculate_diffuse_light_contribution
// (src/lighting.rs:161-166) and sample_disk_light_point (:74-82); squares reuse both.
struct AreaFrame {
    uint64_t seed;
    d3 ub, vb;
};
RT_HD AreaFrame area_frame(uint64_t path_seed, d3 p, d3 lp) {
    AreaFrame a;
    uint64_t ls = path_seed * 0x9E3779B97F4A7C15ull;
    ls = ls + to_u64_sat(p.x * 1000.0) * 0x85EBCA6Bull;
    ls = ls + to_u64_sat(p.y * 1000.0) * 0xC2B2AE35ull;
    ls = ls + to_u64_sat(p.z * 1000.0) * 0x6C8E9CF5ull;
    a.seed = ls;
    d3 ldir = unit(sub(lp, p));
    d3 up = fabs(ldir.x) < 0.9 ? mk(1, 0, 0) : mk(0, 1, 0);
    a.ub = unit(cross(up, ldir));
    a.vb = unit(cross(ldir, a.ub));
    return a;
}

// Square light (extension, no reference counterpart): sample s of n on a cols x (n/cols)
// jittered grid over the side x side square centred on the light, in the same
// light-facing basis as the disk. The jitter of sample s is outputs 2s+1 and 2s+2 of a
// SplitMix64 stream seeded with the light seed, so any sample is computed on its own.
RT_HD uint64_t splitmix_at(uint64_t seed, uint64_t k) {
    uint64_t z = seed + k * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
RT_HD d3 square_sample(const LightDev& L, const AreaFrame& a, int s, int n, int cols) {
    const int rows = n / cols;
    const int ix = s % cols, iy = s / cols;
    const double jx = (double)(splitmix_at(a.seed, 2ull * (uint64_t)s + 1) >> 11) * 0x1.0p-53;
    const double jy = (double)(splitmix_at(a.seed, 2ull * (uint64_t)s + 2) >> 11) * 0x1.0p-53;
    const double x = (((double)ix + jx) / (double)cols - 0.5) * L.size;
    const double y = (((double)iy + jy) / (double)rows - 0.5) * L.size;
    return add(add(ld3(L.pos), scale(x, a.ub)), scale(y, a.vb));
}

RT_HD int light_count(const SceneView& S, const LightDev& L) {
    return L.is_area ? (S.light_samples > 0 ? S.light_samples : L.samples) : 1;
}
RT_HD int light_cols(const SceneView& S, const LightDev& L) {
    return S.light_samples > 0 ? S.light_cols : L.cols;
}

// diffuse_strength of a light point (src/lighting.rs:128,189), as phong_sample computes it
RT_HD double diffuse_strength(d3 p, d3 n, d3 sp) { return fmax_r(dot(n, unit(sub(sp, p))), 0.0); }

// Whether some sample of area light L (its frame `a`; disk samples at pts) has
// diffuse_strength > 0 at the hit (p, n). When none has, the light's shadow queries need
// not be traced: its contribution is diffuse + specular summed over the visible samples
// (src/lighting.rs:178-216), every term is kd * 0 * I * colour = +-0 plus a zero
// specular, so the sum is +0 and so is (sum / n) * (visible / n), whatever the visibility
// (a point light alike, src/lighting.rs:119-140). Samples lie in the light's plane within
// `half` of its centre, so n . (sample - p) lies in c +- e; the slack covers the
// rounding of the samples and of the normalised dot product, and the samples themselves
// are tested only when the interval straddles 0.
RT_HD bool area_lit(const SceneView& S, const LightDev& L, const AreaFrame& a, d3 p, d3 n, SptRef pts,
                    int cnt) {
    const d3 lp = ld3(L.pos);
    const bool sq = L.shape == kLightSquare;
    const double half = sq ? L.size / 2.0 : L.diameter / 2.0;
    const double nu = fabs(dot(n, a.ub)), nv = fabs(dot(n, a.vb));
    const double c = dot(n, sub(lp, p));
    const double e = sq ? half * (nu + nv) : half * sqrt(nu * nu + nv * nv);
    const double slack =
        1e-9 * (fabs(lp.x) + fabs(lp.y) + fabs(lp.z) + fabs(p.x) + fabs(p.y) + fabs(p.z) + fabs(half) + 1.0);
    if (c + e < -slack) return false;
    if (c - e > slack && cnt > 0) return true;
    const int cols = sq ? light_cols(S, L) : 1;
    for (int q = 0; q < cnt; ++q) {
        const d3 sp = sq ? square_sample(L, a, q, cnt, cols) : pts.get(q);
        if (diffuse_strength(p, n, sp) > 0.0) return true;
    }
    return false;
}

// Shadow query `task` belongs to a light whose queries are not traced (area_lit).
RT_HD bool shadow_culled(const SceneView& S, const FrameView& F, int task) {
    const int hq = task / S.shadow_per_hit;
    const uint32_t m = F.skip[hq];
    if (m == 0) return false;
    int j = task - hq * S.shadow_per_hit;
    for (int li = 0; li < S.n_lights; ++li) {
        const int cnt = light_count(S, S.lights[li]);
        if (j < cnt) return li < 32 && ((m >> li) & 1u);
        j -= cnt;
    }
    return false;
}

// Phong term of one visible light sample (src/lighting.rs:123-138 and :184-208); vd is
// the hit's view direction, normalize(camera_pos - p), the same for every sample.
RT_HD d3 phong_sample(const MatDev& m, const LightDev& L, d3 lc, d3 mc, d3 p, d3 n, d3 vd, d3 sp) {
    d3 ld = unit(sub(sp, p));
    double ds = fmax_r(dot(n, ld), 0.0);
    d3 diffuse = scale((m.diffuse * ds) * L.intensity, hadamard(lc, mc));
    d3 spec = mk(0, 0, 0);
    if (ds > 0.0) {
        d3 rd = reflect_dir(neg(ld), n);
        double ss = pow(fmax_r(dot(vd, rd), 0.0), m.shininess);
        spec = scale((m.specular * ss) * L.intensity, lc);
    }
    return add(diffuse, spec);
}

// ------------------------------------------------------------ shade, part 1
// Exact hit record, effective material, shadow sample points, reflection ray.
// Returns true when a reflection ray was written to *refl.
RT_HD bool shade_prepare_item(const SceneView& S, const FrameView& F, int hq, int level, RayRec* refl) {
    int slot = F.hitq[hq];
    const RayRec& rr = F.rays[slot];
    const HitRec& h = F.hits[slot];
    RayD r;
    r.o = ld3(rr.o);
    r.d = ld3(rr.d);
    const ObjDev& ob = S.objs[h.obj];
    d3 p, n;
    bool has_uv = false;
    double u = 0, v = 0;
    if (ob.kind == kSphere) {
        p = add(r.o, scale(h.t, r.d));
        n = face(r.d, divs(sub(p, ld3(ob.c)), ob.r));
    } else if (ob.kind == kPlane) {
        // Plane::hit uv (src/ray.rs:150-173)
        p = add(r.o, scale(h.t, r.d));
        d3 pn = ld3(ob.n);
        n = face(r.d, pn);
        d3 ua = fabs(pn.x) > 0.9 ? mk(0, 1, 0) : mk(1, 0, 0);
        ua = unit(cross(ua, pn));
        d3 va = unit(cross(pn, ua));
        d3 rel = sub(p, ld3(ob.c));
        u = dot(rel, ua);
        v = dot(rel, va);
        has_uv = true;
    } else if (ob.kind == kCube) {
        double t;
        d3 wn;
        cube_exact(ob, r, 0.001, h.t, t, p, wn);  // same t, point and normal as the hit
        n = face(r.d, wn);
    } else {
        const MeshDev& M = S.meshes[ob.mesh];
        double t;
        d3 tn;
        tri_exact(M.tri64 + (size_t)h.tri * 9, r, 0.001, INFINITY, t, tn, u, v);
        p = add(r.o, scale(h.t, r.d));
        n = face(r.d, tn);
        has_uv = true;
    }
    // effective material (src/lighting.rs:229-242, apply_texture :9-54)
    const MatDev& base = S.mats[ob.mat];
    int eff = ob.mat;
    d3 mc = ld3(base.color);
    if (base.texture != 0 && has_uv) {
        if (base.texture == 1) {
            double hw = base.line_width / 2.0;
            double um = fabs(fract_r(u / base.cell_size));
            double vm = fabs(fract_r(v / base.cell_size));
            bool on_u = um <= hw || um >= (1.0 - hw);
            bool on_v = vm <= hw || vm >= (1.0 - hw);
            if (on_u || on_v) mc = ld3(base.line_color);
        } else {
            int32_t cu = to_i32_sat(floor(u)), cv = to_i32_sat(floor(v));
            int32_t sum = (int32_t)((uint32_t)cu + (uint32_t)cv);
            if (sum % 2 != 0) {
                eff = base.mat_b;
                mc = ld3(S.mats[eff].color);
            }
        }
    }
    ShadeRec sr;
    st3(sr.p, p);
    st3(sr.n, n);
    st3(sr.mc, mc);
    sr.mat = eff;
    sr.path = rr.path;
    sr.level = level;
    sr.base = ob.mat;
    sr.spawned = base.has_refl && base.refl > 0.0 && (F.max_depth - level) > 1;
    sr.pad = 0;
    F.shade[hq] = sr;
    if (level == 0 && F.outline) {
        double* o = F.ol_dn + (size_t)rr.path * 4;
        o[0] = length(sub(p, ld3(S.cam_pos)));
        o[1] = n.x;
        o[2] = n.y;
        o[3] = n.z;
    }
    // shadow sample points of point and disk lights, in scene order (src/lighting.rs:251-285),
    // and the lights whose shadow queries are not traced (light_lit; with features, such a
    // light's samples count as not visible)
    uint64_t seed = F.path_seed[rr.path];
    const SptRef spt{F.spt + hq, F.spt_stride};
    uint32_t skip = 0;
    int j = 0;
    for (int li = 0; li < S.n_lights; ++li) {
        const LightDev& L = S.lights[li];
        d3 lp = ld3(L.pos);
        const bool cull = li < 32;
        if (!L.is_area) {
            spt.put(j, lp);
            ++j;
            if (cull && !(diffuse_strength(p, n, lp) > 0.0)) skip |= 1u << li;
            continue;
        }
        const int cnt = light_count(S, L);
        AreaFrame a = area_frame(seed, p, lp);
        if (L.shape != kLightSquare) {  // squares are regenerated where needed
            double radius = L.diameter / 2.0;
            if (!disk_samples_fast(a.seed, cnt, radius, lp, a.ub, a.vb, spt.at(j)))
                disk_samples_slow(a.seed, cnt, radius, lp, a.ub, a.vb, spt.at(j));
        }
        if (cull && !area_lit(S, L, a, p, n, spt.at(j), cnt)) skip |= 1u << li;
        if (L.shape != kLightSquare) j += cnt;
    }
    F.skip[hq] = skip;
    // reflection (src/lighting.rs:469-494): mirror of the view direction from the camera
    if (sr.spawned) {
        d3 vd = unit(sub(ld3(S.cam_pos), p));
        d3 rd = reflect_dir(neg(vd), n);
        RayRec o;
        st3(o.o, add(p, scale(0.001, n)));
        st3(o.d, unit(rd));
        o.path = rr.path;
        o.level = level + 1;
        *refl = o;
        return true;
    }
    return false;
}

// ------------------------------------------------------------ shadow stage
// Shadow ray of task `task` (hit task / slots, slot task % slots): from p + 0.001 n
// toward the light point, with the distance to it (src/lighting.rs:117-120, 178-185).
RT_HD void shadow_ray(const SceneView& S, const FrameView& F, int task, RayD& r, double& dist) {
    int hq = task / S.shadow_per_hit;
    const ShadeRec& sr = F.shade[hq];
    d3 p = ld3(sr.p), n = ld3(sr.n);
    d3 sp;
    if (!S.has_square) {
        sp = SptRef{F.spt + hq, F.spt_stride}.get(task - hq * S.shadow_per_hit);  // slots == stored points
    } else {
        int j = task - hq * S.shadow_per_hit, k = 0;
        for (int li = 0;; ++li) {
            const LightDev& L = S.lights[li];
            const int cnt = light_count(S, L);
            if (j < cnt) {
                if (L.shape == kLightSquare)
                    sp = square_sample(L, area_frame(F.path_seed[sr.path], p, ld3(L.pos)), j, cnt, light_cols(S, L));
                else
                    sp = SptRef{F.spt + hq, F.spt_stride}.get(k + j);
                break;
            }
            j -= cnt;
            if (L.shape != kLightSquare) k += cnt;
        }
    }
    d3 ld = unit(sub(sp, p));
    dist = length(sub(sp, p));
    r.o = add(p, scale(0.001, n));
    r.d = unit(ld);
}

// Queue item -> shadow task (hit hq's slot j, task hq * slots + j). Hits go in blocks of 64
// and, within a block, light by light: the slots of one light for hit 0, hit 1, ... So the
// lanes of a wavefront trace rays toward one light from neighbouring hits (a point light:
// 64 rays toward the same point) or the samples of one area light from a few neighbouring
// hits, instead of rays toward several lights far apart. The last partial block of hits
// keeps hit-major order.
RT_HD int shadow_task(const SceneView& S, int n_hits, int item) {
    if (S.n_lights <= 1) return item;  // one light: already that order
    const int slots = S.shadow_per_hit;
    const int blk = item / (64 * slots), in = item - blk * 64 * slots;
    if ((blk + 1) * 64 > n_hits) return item;  // hit-major: the same task index
    int off = 0;
    for (int li = 0; li < S.n_lights; ++li) {
        const int cnt = light_count(S, S.lights[li]);
        if (in < 64 * (off + cnt)) {
            const int r = in - 64 * off;
            return (blk * 64 + r / cnt) * slots + off + r % cnt;
        }
        off += cnt;
    }
    return item;  // not reached: the lights' slots add up to `slots`
}

template <bool count>
RT_HD void shadow_item(const SceneView& S, const FrameView& F, int task, StackRef stk, Work& w) {
    if (shadow_culled(S, F, task)) {
        if (count) w.v[kWorkSkipped]++;
        return;
    }
    RayD r;
    double dist;
    shadow_ray(S, F, task, r, dist);
    bool blocked = world_any<count>(S, r, 0.001, dist, F.use_kd != 0, stk, w);
    F.vis[task] = blocked ? 0 : 1;
}

// ------------------------------------------------------------ shade, part 2
// Phong term of shadow slot j of hit hq when that slot belongs to an area light and its
// sample is visible, else zero: the term shade_finish_item adds for that sample, from the
// same inputs (precomputed across the lanes of a wavefront by k_shade_finish_wide).
RT_HD d3 area_slot_term(const SceneView& S, const FrameView& F, int hq, int j) {
    const ShadeRec& sr = F.shade[hq];
    const uint8_t* vis = F.vis + (size_t)hq * S.shadow_per_hit;
    if (!vis[j]) return mk(0, 0, 0);
    int k = 0;  // stored (point and disk) sample points before this light
    for (int li = 0, jj = 0; li < S.n_lights; ++li) {
        const LightDev& L = S.lights[li];
        const int cnt = L.is_area ? light_count(S, L) : 1;
        if (j < jj + cnt) {
            if (!L.is_area) return mk(0, 0, 0);
            const int smp = j - jj;
            const d3 p = ld3(sr.p), n = ld3(sr.n), mc = ld3(sr.mc);
            const MatDev& m = S.mats[sr.mat];
            const d3 vd = unit(sub(ld3(S.cam_pos), p));
            const d3 lc = ld3(L.color);
            d3 sp;
            if (L.shape == kLightSquare) {
                const AreaFrame a = area_frame(F.path_seed[sr.path], p, ld3(L.pos));
                sp = square_sample(L, a, smp, cnt, light_cols(S, L));
            } else {
                sp = SptRef{F.spt + hq, F.spt_stride}.get(k + smp);
            }
            return phong_sample(m, L, lc, mc, p, n, vd, sp);
        }
        jj += cnt;
        if (!L.is_area || L.shape != kLightSquare) k += cnt;
    }
    return mk(0, 0, 0);
}

// phong_lighting (src/lighting.rs:220-288) with visibility from the shadow stage, then
// apply_fog (:297-319). Writes the level record (colour, reflectivity or -1).
// pre (nullable): the area-light sample terms per shadow slot (area_slot_term), summed here
// in sample order exactly as the terms computed in place would be.
RT_HD void shade_finish_item(const SceneView& S, const FrameView& F, int hq, const double* pre = nullptr) {
    const ShadeRec& sr = F.shade[hq];
    d3 p = ld3(sr.p), n = ld3(sr.n), mc = ld3(sr.mc);
    const MatDev& m = S.mats[sr.mat];
    d3 color = scale(m.ambient * S.amb_i, hadamard(ld3(S.amb), mc));
    d3 cam = ld3(S.cam_pos);
    const d3 vd = unit(sub(cam, p));
    const SptRef spt{F.spt + hq, F.spt_stride};
    const uint8_t* vis = F.vis + (size_t)hq * S.shadow_per_hit;
    uint64_t path_seed = 0;
    if (S.has_square) path_seed = F.path_seed[sr.path];
    int j = 0, k = 0;
    const uint32_t culled = F.skip[hq];
    for (int li = 0; li < S.n_lights; ++li) {
        const LightDev& L = S.lights[li];
        d3 lc = ld3(L.color);
        if (li < 32 && ((culled >> li) & 1u)) {
            // no sample of this light faces the hit (area_lit): its contribution is +0
            // whatever the visibility, and its shadow queries were not traced
            color = add(color, mk(0, 0, 0));
            const int cnt = L.is_area ? light_count(S, L) : 1;
            j += cnt;
            if (!L.is_area || L.shape != kLightSquare) k += cnt;
            continue;
        }
        if (!L.is_area) {
            // calculate_point_light_contribution (src/lighting.rs:100-142)
            if (vis[j]) color = add(color, phong_sample(m, L, lc, mc, p, n, vd, ld3(L.pos)));
            else color = add(color, mk(0, 0, 0));
            ++j;
            ++k;
            continue;
        }
        // calculate_diffuse_light_contribution (src/lighting.rs:145-217), n samples
        const int cnt = light_count(S, L);
        d3 total = mk(0, 0, 0);
        int visible = 0;
        if (L.shape == kLightSquare) {
            AreaFrame a = area_frame(path_seed, p, ld3(L.pos));
            const int cols = light_cols(S, L);
            for (int s = 0; s < cnt; ++s) {
                if (!vis[j + s]) continue;
                ++visible;
                total = add(total, pre ? ld3(pre + 3 * (j + s))
                                       : phong_sample(m, L, lc, mc, p, n, vd, square_sample(L, a, s, cnt, cols)));
            }
        } else {
            for (int s = 0; s < cnt; ++s) {
                if (!vis[j + s]) continue;
                ++visible;
                total = add(total, pre ? ld3(pre + 3 * (j + s)) : phong_sample(m, L, lc, mc, p, n, vd, spt.get(k + s)));
            }
            k += cnt;
        }
        j += cnt;
        // (total / n) * (visible / n)
        color = add(color, scale((double)visible / (double)cnt, divs(total, (double)cnt)));
    }
    double fog_f = 0.0;
    if (S.has_fog) {
        double dist = length(sub(p, cam));
        double f;
        if (dist <= S.fog_start) f = 0.0;
        else if (dist >= S.fog_end) f = 1.0;
        else f = (dist - S.fog_start) / (S.fog_end - S.fog_start);
        f = 1.0 - exp(-S.fog_density * f);
        f = clamp_r(f, 0.0, 1.0);
        d3 fc = ld3(S.fog_color);
        color = add(scale(1.0 - f, color), scale(f, fc));  // color * (1 - f) + fog * f
        fog_f = f;
    }
    if (F.feat && sr.level == 0) {
        int fpix = feature_pixel(F, sr.path);
        if (fpix >= 0) {
            float* fp = F.feat + (size_t)fpix * kFeatCount;
            fp[kFeatAlbedo + 0] = (float)mc.x;
            fp[kFeatAlbedo + 1] = (float)mc.y;
            fp[kFeatAlbedo + 2] = (float)mc.z;
            fp[kFeatNormal + 0] = (float)n.x;
            fp[kFeatNormal + 1] = (float)n.y;
            fp[kFeatNormal + 2] = (float)n.z;
            fp[kFeatDepth] = (float)length(sub(p, cam));
            // visibility of the area-light samples (all shadow slots when there is no area light);
            // a light none of whose samples faces the hit was not traced and counts as unseen
            int nv = 0, nt = 0, pv = 0, pt = 0;
            for (int li = 0, jj = 0; li < S.n_lights; ++li) {
                const int cnt = light_count(S, S.lights[li]);
                int v = 0;
                if (!(li < 32 && ((culled >> li) & 1u)))
                    for (int q = 0; q < cnt; ++q) v += vis[jj + q];
                if (S.lights[li].is_area) {
                    nv += v;
                    nt += cnt;
                } else {
                    pv += v;
                    pt += cnt;
                }
                jj += cnt;
            }
            if (nt == 0) {
                nv = pv;
                nt = pt;
            }
            float mean = nt > 0 ? (float)nv / (float)nt : 1.0f;
            fp[kFeatVisMean] = mean;
            fp[kFeatVisVar] = nt > 0 ? mean * (1.0f - mean) / (float)nt : 0.0f;
            fp[kFeatFog] = (float)fog_f;
        }
    }
    double* Lr = F.lvl + ((size_t)sr.level * F.n_paths + sr.path) * 4;
    Lr[0] = color.x;
    Lr[1] = color.y;
    Lr[2] = color.z;
    Lr[3] = sr.spawned ? S.mats[sr.base].refl : -1.0;
}

}  // namespace rt

namespace rt {

// ------------------------------------------------------------ path colour
// ray_color recursion unrolled: c_L = c_L * (1 - rho_L) + c_{L+1} * rho_L from the
// deepest level up (src/lighting.rs:469-494), in the reference's evaluation order.
RT_HD d3 path_color(const FrameView& F, int path) {
    int last = 0;
    while (last + 1 < F.max_depth) {
        const double* L = F.lvl + ((size_t)last * F.n_paths + path) * 4;
        if (!(L[3] >= 0.0)) break;
        ++last;
    }
    const double* Ld = F.lvl + ((size_t)last * F.n_paths + path) * 4;
    d3 c = mk(Ld[0], Ld[1], Ld[2]);
    for (int l = last - 1; l >= 0; --l) {
        const double* L = F.lvl + ((size_t)l * F.n_paths + path) * 4;
        double rho = L[3];
        c = add(scale(1.0 - rho, mk(L[0], L[1], L[2])), scale(rho, c));
    }
    return c;
}

// Per pixel: quincunx (centre + 4 shared corners) / 5, or the mean over samples; with
// the outline on, the nearest first-hit depth/normal over the samples
// (src/renderer.rs:552, :692-698, :881-886).
RT_HD void pixel_item(const FrameView& F, int pix) {
    int x = pix % F.W, y = pix / F.W;
    if (!pixel_in_mask(F, x, y)) return;
    d3 c;
    if (F.max_depth <= 0) {
        c = mk(0, 0, 0);
    } else if (F.aa == 1) {
        int base = F.W * F.H;
        c = path_color(F, pix);
        c = add(c, path_color(F, base + y * (F.W + 1) + x));
        c = add(c, path_color(F, base + y * (F.W + 1) + x + 1));
        c = add(c, path_color(F, base + (y + 1) * (F.W + 1) + x));
        c = add(c, path_color(F, base + (y + 1) * (F.W + 1) + x + 1));
        c = divs(c, 5.0);
    } else {
        d3 tot = mk(0, 0, 0);
        for (int s = 0; s < F.N; ++s) tot = add(tot, path_color(F, pix * F.N + s));
        c = divs(tot, (double)F.N);
    }
    st3(F.pix + (size_t)pix * 3, c);
    if (F.outline) {
        double* o = F.px_dn + (size_t)pix * 4;
        o[0] = -1.0;
        o[1] = o[2] = o[3] = 0.0;
        if (F.max_depth > 0) {
            // the nearest sample with a hit, the first one on equal (or NaN) depths, as the
            // reference picks it (src/renderer.rs:692-697)
            bool have = false;
            for (int s = 0; s < F.N; ++s) {
                const double* q = F.ol_dn + ((size_t)pix * F.N + s) * 4;
                double d = q[0];
                if (!(d < 0.0) && (!have || d < o[0])) {
                    have = true;
                    o[0] = d;
                    o[1] = q[1];
                    o[2] = q[2];
                    o[3] = q[3];
                }
            }
        }
    }
}

// ------------------------------------------------------------ outline
// compute_edge_strength (src/outline.rs:132-182) on per-pixel depth/normal. `fetch`
// reads (depth, normal) of a pixel; false = background or outside the image (the
// reference reads x+1 / y+1 past the last column/row as background, :68-82).
template <class Fetch>
RT_HD double edge_strength(const FrameView& F, int x, int y, Fetch fetch) {
    double cd, cn[3];
    bool cf = fetch(x, y, cd, cn);
    double mx = 0.0;
    int nbr[8][2];
    int k = 0;
    if (F.ol_8) {
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                if (!dx && !dy) continue;
                if (x + dx >= 0 && y + dy >= 0) {
                    nbr[k][0] = x + dx;
                    nbr[k][1] = y + dy;
                    ++k;
                }
            }
    } else {
        if (x > 0) { nbr[k][0] = x - 1; nbr[k][1] = y; ++k; }
        nbr[k][0] = x + 1; nbr[k][1] = y; ++k;
        if (y > 0) { nbr[k][0] = x; nbr[k][1] = y - 1; ++k; }
        nbr[k][0] = x; nbr[k][1] = y + 1; ++k;
    }
    for (int i = 0; i < k; ++i) {
        double nd, nn[3];
        bool nf = fetch(nbr[i][0], nbr[i][1], nd, nn);
        double e = 0.0;
        if (cf && nf) {
            double dd = fabs(cd - nd);
            double ndd = dd / fmax_r(cd * 0.1, 0.1);
            double dp = clamp_r(cn[0] * nn[0] + cn[1] * nn[1] + cn[2] * nn[2], -1.0, 1.0);
            e = F.ol_wd * ndd + F.ol_wn * (1.0 - dp);
        } else if (cf != nf) {
            e = F.ol_wd * 2.0 + F.ol_wn * 2.0;
        }
        mx = fmax_r(mx, e);
    }
    return mx;
}

// detect_edges threshold/normalise (src/outline.rs:113-129)
RT_HD double edge_mask_value(const FrameView& F, double e) {
    if (e > F.ol_thr) return (e - F.ol_thr) / fmax_r(1.0 - F.ol_thr, 0.001);
    return 0.0;
}

// dilate_edges (src/outline.rs:227-268) for one pixel, reading the undilated mask
RT_HD double dilate_item(const FrameView& F, const double* mask, int x, int y) {
    int radius = (int)ceil(F.ol_thick - 1.0);
    double m = mask[(size_t)y * F.W + x];
    for (int dy = -radius; dy <= radius; ++dy)
        for (int dx = -radius; dx <= radius; ++dx) {
            int nx = x + dx, ny = y + dy;
            if (nx >= 0 && ny >= 0 && nx < F.W && ny < F.H) {
                double dist = sqrt((double)(dx * dx + dy * dy));
                if (dist <= F.ol_thick) {
                    double fall = 1.0 - (dist / F.ol_thick);
                    m = fmax_r(m, mask[(size_t)ny * F.W + nx] * fall);
                }
            }
        }
    return m;
}

// blend_colors (src/outline.rs:271-278) then create_image_from_data quantisation
// (src/renderer.rs:899-912)
RT_HD void finalize_item(const FrameView& F, const double* mask, int pix) {
    if (!pixel_in_mask(F, pix % F.W, pix / F.W)) return;  // keep what the output holds
    double* c = F.pix + (size_t)pix * 3;
    double r = c[0], g = c[1], b = c[2];
    if (mask && mask[pix] > 0.0) {
        double w = clamp_r(fmin_r(mask[pix], 1.0), 0.0, 1.0);
        r = r * (1.0 - w) + F.ol_color[0] * w;
        g = g * (1.0 - w) + F.ol_color[1] * w;
        b = b * (1.0 - w) + F.ol_color[2] * w;
    }
    if (F.rad) {
        F.rad[(size_t)pix * 3 + 0] = (float)r;
        F.rad[(size_t)pix * 3 + 1] = (float)g;
        F.rad[(size_t)pix * 3 + 2] = (float)b;
    }
    if (F.feat) {
        float* fp = F.feat + (size_t)pix * kFeatCount;
        fp[kFeatRad + 0] = (float)r;
        fp[kFeatRad + 1] = (float)g;
        fp[kFeatRad + 2] = (float)b;
    }
    F.rgb[(size_t)pix * 3 + 0] = to_u8_sat(clamp_r(r, 0.0, 1.0) * 255.0);
    F.rgb[(size_t)pix * 3 + 1] = to_u8_sat(clamp_r(g, 0.0, 1.0) * 255.0);
    F.rgb[(size_t)pix * 3 + 2] = to_u8_sat(clamp_r(b, 0.0, 1.0) * 255.0);
}

}  // namespace rt
