// Runs the device mesh walk (mesh_trace in kernels/trace_core.hpp, compiled here for the
// host) on seeded meshes built by the library's own host builders (k-d tree and mesh BVH)
// and checks every answer against brute force over all triangles with the exact f64 test
// (tri_exact, intersect_triangle of src/ray.rs:527-583):
//  * any hit (shadow rays): the same yes/no for every ray;
//  * closest hit: the same distance, and the same triangle unless another one reaches
//    the same distance (the walk then reports a tie for the reference-order resolver).
// Rays start near the mesh, or 1e3 to 3e6 mesh radii away aimed at points around it, the
// case where the walk re-bases its FP32 origin to the box entry (shadow walks) or takes
// the f64 bounds test only (closest-hit walks). Also reports the triangle tests per ray
// and fails when far shadow rays cost much more than near ones. Prints "ok <rays>".
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "../../rtrace_amd/csrc/host/kdbuild.hpp"
#include "../../rtrace_amd/csrc/host/mesh_bvh.hpp"
#include "../../rtrace_amd/csrc/kernels/trace_core.hpp"

using namespace rt;

static float up(double x) {
    float f = (float)x;
    return (double)f < x ? std::nextafter(f, INFINITY) : f;
}
static float down(double x) {
    float f = (float)x;
    return (double)f > x ? std::nextafter(f, -INFINITY) : f;
}

struct HostMesh {
    std::vector<double> v;
    FlatKd kd;
    MeshBvh bvh;
    MeshDev md{};
};

// the fields rt_scene_upload fills, pointing at host arrays
static bool make_mesh(HostMesh& H) {
    build_flat_kd(H.v, H.kd, false);
    for (const KdNode& nd : H.kd.nodes)
        if ((nd.b & 3u) != kAxisLeaf && (nd.b & kFlagMisplaced)) return false;  // no BVH for such meshes
    if (!build_mesh_bvh(H.v, H.kd.center, H.bvh)) return false;
    MeshDev& m = H.md;
    m.nodes = H.kd.nodes.data();
    m.recs = H.kd.recs.data();
    m.bvh = H.bvh.nodes4.data();
    m.bvh_recs = H.bvh.recs.data();
    m.tri64 = H.v.data();
    m.nb64 = H.kd.nb64.data();
    m.root_kd = H.kd.root_kd;
    m.root_bf = H.kd.root_bf;
    m.n_tris = (uint32_t)(H.v.size() / 9);
    m.cx = H.kd.center[0];
    m.cy = H.kd.center[1];
    m.cz = H.kd.center[2];
    double R = 0;
    for (size_t i = 0; i < H.v.size(); i += 3)
        R = std::max(R, std::fabs(H.v[i] - m.cx) + std::fabs(H.v[i + 1] - m.cy) + std::fabs(H.v[i + 2] - m.cz));
    m.radius = up(R * (1.0 + 1e-6));
    for (int k = 0; k < 3; ++k) {
        const double c = k == 0 ? m.cx : (k == 1 ? m.cy : m.cz);
        m.bmin[k] = down(H.kd.bmin[k] - c);
        m.bmax[k] = up(H.kd.bmax[k] - c);
        m.bmin64[k] = H.kd.bmin[k];
        m.bmax64[k] = H.kd.bmax[k];
    }
    return true;
}

static void add_tri(std::vector<double>& v, const double* a, const double* b, const double* c) {
    for (int k = 0; k < 3; ++k) v.push_back(a[k]);
    for (int k = 0; k < 3; ++k) v.push_back(b[k]);
    for (int k = 0; k < 3; ++k) v.push_back(c[k]);
}

// latitude-longitude sphere with radial noise, centred off the origin
static std::vector<double> noisy_sphere(int nlat, int nlon, double radius, std::mt19937_64& g) {
    std::uniform_real_distribution<double> U(-0.03, 0.03);
    const double c[3] = {2.0, -1.0, 0.5};
    std::vector<double> P((nlat + 1) * nlon * 3);
    for (int i = 0; i <= nlat; ++i)
        for (int j = 0; j < nlon; ++j) {
            const double th = M_PI * i / nlat, ph = 2 * M_PI * j / nlon;
            const double r = radius * (1.0 + ((i == 0 || i == nlat) ? 0.0 : U(g)));
            double* p = &P[(i * nlon + j) * 3];
            p[0] = c[0] + r * std::sin(th) * std::cos(ph);
            p[1] = c[1] + r * std::sin(th) * std::sin(ph);
            p[2] = c[2] + r * std::cos(th);
        }
    std::vector<double> v;
    for (int i = 0; i < nlat; ++i)
        for (int j = 0; j < nlon; ++j) {
            const int j1 = (j + 1) % nlon;
            const double* a = &P[(i * nlon + j) * 3];
            const double* b = &P[(i * nlon + j1) * 3];
            const double* c2 = &P[((i + 1) * nlon + j) * 3];
            const double* d = &P[((i + 1) * nlon + j1) * 3];
            if (i > 0) add_tri(v, a, b, d);
            if (i < nlat - 1) add_tri(v, a, d, c2);
        }
    return v;
}

static std::vector<double> soup(int n, std::mt19937_64& g) {
    std::uniform_real_distribution<double> U(-1.0, 1.0), S(0.02, 0.4);
    std::vector<double> v;
    for (int t = 0; t < n; ++t) {
        const double c[3] = {U(g), U(g), U(g)}, s = S(g);
        double p[9];
        for (int k = 0; k < 9; ++k) p[k] = c[k % 3] + s * U(g);
        add_tri(v, p, p + 3, p + 6);
    }
    return v;
}

static d3 unit_vec(std::mt19937_64& g) {
    std::normal_distribution<double> N(0.0, 1.0);
    for (;;) {
        d3 d = mk(N(g), N(g), N(g));
        const double l = length(d);
        if (l > 1e-3 && std::fabs(d.x) > 1e-6 * l && std::fabs(d.y) > 1e-6 * l && std::fabs(d.z) > 1e-6 * l)
            return unit(d);  // no axis-parallel component: the walk takes the BVH
    }
}

struct Brute {
    bool any;
    double t;
    int32_t tri;
    bool tie;
};
static Brute brute(const std::vector<double>& v, const RayD& r, double tmin, double tmax) {
    Brute b{false, INFINITY, -1, false};
    for (size_t i = 0; i * 9 < v.size(); ++i) {
        double tt, uu, vv;
        d3 nn;
        if (!tri_exact(v.data() + i * 9, r, tmin, INFINITY, tt, nn, uu, vv)) continue;
        if (tt > tmin && tt < tmax) b.any = true;
        if (tt < b.t) {
            b.t = tt;
            b.tri = (int32_t)i;
            b.tie = false;
        } else if (tt == b.t) {
            b.tie = true;
        }
    }
    return b;
}

int main() {
    std::mt19937_64 g(20261204);
    std::vector<HostMesh> meshes(2);
    meshes[0].v = noisy_sphere(40, 64, 1.5, g);  // 5,040 triangles
    meshes[1].v = soup(3000, g);
    uint32_t snode[kMaxKdDepth];
    float st0[kMaxKdDepth];
    const StackRef stk{snode, st0, 1};
    long long rays = 0;
    for (HostMesh& H : meshes) {
        if (!make_mesh(H)) {
            std::printf("FAIL no BVH for a test mesh\n");
            return 1;
        }
        const MeshDev& M = H.md;
        const d3 c = mk((M.bmin64[0] + M.bmax64[0]) / 2, (M.bmin64[1] + M.bmax64[1]) / 2, (M.bmin64[2] + M.bmax64[2]) / 2);
        const d3 ext = mk(M.bmax64[0] - c.x, M.bmax64[1] - c.y, M.bmax64[2] - c.z);
        std::uniform_real_distribution<double> U(-1.3, 1.3), L(-2.0, 3.0);
        const double dists[4] = {0.0, 1e3, 1e5, 3e6};  // in mesh radii; 0: near the mesh
        double tris_near = 0, tris_far = 0;
        long long n_near = 0, n_far = 0;
        for (double D : dists) {
            for (int i = 0; i < 4000; ++i) {
                const d3 target = add(c, mk(U(g) * ext.x, U(g) * ext.y, U(g) * ext.z));
                RayD r;
                if (D == 0.0) {
                    r.o = add(c, mk(2 * U(g) * ext.x, 2 * U(g) * ext.y, 2 * U(g) * ext.z));
                    r.d = unit_vec(g);
                } else {
                    r.o = add(target, scale(D * M.radius, unit_vec(g)));
                    r.d = unit(sub(target, r.o));
                    if (std::fabs(r.d.x) < 1e-9 || std::fabs(r.d.y) < 1e-9 || std::fabs(r.d.z) < 1e-9) continue;
                }
                const double tmin = 0.001;
                // shadow rays end somewhere around the target: before, inside or beyond the mesh
                const double tmax = std::max(0.01, length(sub(target, r.o)) + L(g) * M.radius);
                const Brute b = brute(H.v, r, tmin, tmax);
                Work w;
                double best = tmax;
                int32_t tri = -1;
                const bool any = mesh_trace<true, true>(M, r, true, tmin, best, tri, stk, w);
                if (any != b.any) {
                    std::printf("FAIL any-hit D=%g ray %d: walk %d brute %d\n", D, i, (int)any, (int)b.any);
                    return 1;
                }
                (D == 0.0 ? tris_near : tris_far) += (double)w.v[kWorkTris];
                ++(D == 0.0 ? n_near : n_far);
                Work w2;
                double cbest = INFINITY;
                int32_t ctri = -1;
                const bool found = mesh_trace<false, false>(M, r, true, tmin, cbest, ctri, stk, w2);
                const bool bfound = b.tri >= 0;
                if (found != bfound || (found && (cbest != b.t || (!b.tie && ctri != b.tri)))) {
                    std::printf("FAIL closest D=%g ray %d: walk %d t=%.17g tri %d, brute %d t=%.17g tri %d tie %d\n", D, i,
                                (int)found, cbest, ctri, (int)bfound, b.t, b.tri, (int)b.tie);
                    return 1;
                }
                ++rays;
            }
        }
        const double near = tris_near / n_near, far = tris_far / n_far;
        std::printf("mesh %zu triangles: shadow walk triangle tests per ray near %.2f far %.2f\n", H.v.size() / 9, near,
                    far);
        if (far > 3.0 * near + 8.0) {
            std::printf("FAIL far shadow rays test %.1f triangles per ray against %.1f near\n", far, near);
            return 1;
        }
    }
    std::printf("ok %lld\n", rays);
    return 0;
}
