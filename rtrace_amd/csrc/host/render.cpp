// render.cpp — device scene upload, frame buffers and the Renderer::render driver
// (src/renderer.rs:120-382) on the HIP pipeline, plus the C ABI (include/rtrace_hip.h).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../../include/rtrace_hip.h"
#include "../common/layout.hpp"
#include "../common/vec.hpp"
#include "../kernels/kdbuild_gpu.hpp"
#include "../kernels/launch.hpp"
#include "frame.hpp"
#include "kdbuild.hpp"
#include "mesh_bvh.hpp"
#include "object_bvh.hpp"
#include "scene_host.hpp"

namespace rt {
namespace {

thread_local std::string g_tls_error;

template <class T>
void upload(DevBuf& b, const std::vector<T>& v, hipStream_t st) {
    b.reserve(v.size() * sizeof(T));
    if (!v.empty()) HIP_TRY(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
}

}  // namespace

// Columns of the jittered grid for n square-light samples: the largest divisor of n that
// is at most sqrt(n) (16 -> 4x4, 8 -> 2x4, 7 -> 1x7).
int grid_cols(int n) {
    int c = (int)std::sqrt((double)n);
    while (c > 1 && n % c != 0) --c;
    return c < 1 ? 1 : c;
}

namespace {

float round_down(double x) {
    float f = (float)x;
    if ((double)f > x) f = std::nextafter(f, -INFINITY);
    return f;
}
float round_up(double x) {
    float f = (float)x;
    if ((double)f < x) f = std::nextafter(f, INFINITY);
    return f;
}

// (inv(A))^T for the upper-left 3x3 of a column-major 4x4, with nalgebra's do_inverse3
// formulas (applied to cube normals, src/ray.rs:474-481); row-major out.
void normal_matrix(const double* w2l, double out[9]) {
    auto a = [&](int r, int c) { return w2l[c * 4 + r]; };
    double m12_23 = a(1, 1) * a(2, 2) - a(2, 1) * a(1, 2);
    double m11_23 = a(1, 0) * a(2, 2) - a(2, 0) * a(1, 2);
    double m11_22 = a(1, 0) * a(2, 1) - a(2, 0) * a(1, 1);
    double det = a(0, 0) * m12_23 - a(0, 1) * m11_23 + a(0, 2) * m11_22;
    double inv[3][3];
    if (det == 0.0) {
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) inv[r][c] = r == c ? 1.0 : 0.0;
    } else {
        inv[0][0] = m12_23 / det;
        inv[0][1] = (a(0, 2) * a(2, 1) - a(2, 2) * a(0, 1)) / det;
        inv[0][2] = (a(0, 1) * a(1, 2) - a(0, 2) * a(1, 1)) / det;
        inv[1][0] = -m11_23 / det;
        inv[1][1] = (a(0, 0) * a(2, 2) - a(0, 2) * a(2, 0)) / det;
        inv[1][2] = (a(0, 2) * a(1, 0) - a(0, 0) * a(1, 2)) / det;
        inv[2][0] = m11_22 / det;
        inv[2][1] = (a(0, 1) * a(2, 0) - a(0, 0) * a(2, 1)) / det;
        inv[2][2] = (a(0, 0) * a(1, 1) - a(0, 1) * a(1, 0)) / det;
    }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) out[r * 3 + c] = inv[c][r];
}

// World box of a cube's hit geometry: Cube::hit intersects the box [-h, h] in the
// local frame local = A^-1 (x - c) (src/ray.rs:413-421), A^-1 = upper 3x3 of
// world_to_local, so the geometry is c + A [-h, h]^3 with A = ntf^T. Padded and rounded
// outward for the FP32 prefilter.
void cube_world_box(ObjDev& d) {
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int corner = 0; corner < 8; ++corner) {
        double l[3] = {(corner & 1) ? d.half[0] : -d.half[0], (corner & 2) ? d.half[1] : -d.half[1],
                       (corner & 4) ? d.half[2] : -d.half[2]};
        for (int i = 0; i < 3; ++i) {
            double x = d.c[i];
            for (int j = 0; j < 3; ++j) x += d.ntf[j * 3 + i] * l[j];  // A[i][j] = ntf[j][i]
            mn[i] = std::min(mn[i], x);
            mx[i] = std::max(mx[i], x);
        }
    }
    for (int i = 0; i < 3; ++i) {
        double pad = 1e-6 * (std::fabs(mn[i]) + std::fabs(mx[i]) + (mx[i] - mn[i])) + 1e-12;
        d.bmin[i] = std::isfinite(mn[i]) ? round_down(mn[i] - pad) : -3.0e38f;
        d.bmax[i] = std::isfinite(mx[i]) ? round_up(mx[i] + pad) : 3.0e38f;
    }
}

bool is_identity(const double* m) {
    for (int i = 0; i < 16; ++i)
        if (m[i] != ((i % 5 == 0) ? 1.0 : 0.0)) return false;
    return true;
}

}  // namespace

}  // namespace rt

namespace rt {
namespace {

void build_device_scene(rt_ctx* ctx, SceneDevice& D) {
    const HostScene& H = D.host;
    auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = ctx->stream;
    std::vector<ObjDev> objs;
    std::vector<MeshDev> meshes;
    D.kd.resize(H.meshes.size());
    D.mesh_bufs.clear();
    for (size_t m = 0; m < H.meshes.size(); ++m) {
        const std::vector<double>& v = H.meshes[m];
        FlatKd& kd = D.kd[m];
        auto nodes = std::make_unique<DevBuf>(), recs = std::make_unique<DevBuf>(), tri64 = std::make_unique<DevBuf>(),
             nb64 = std::make_unique<DevBuf>();
        upload(*tri64, v, st);
        if (v.size() / 9 >= kDeviceKdMinTris) {
            // big meshes: the same tree, built on the device from the uploaded vertices
            build_flat_kd_device(v, tri64->as<double>(), kd, *nodes, *recs, *nb64, st);
        } else {
            build_flat_kd(v, kd);
            upload(*nodes, kd.nodes, st);
            upload(*recs, kd.recs, st);
            upload(*nb64, kd.nb64, st);
        }
        // the BVH walk needs every triangle in the leaf of its own cell (MeshWalk::bvh)
        bool moved = false;
        for (const KdNode& nd : kd.nodes) moved = moved || ((nd.b & 3u) != kAxisLeaf && (nd.b & kFlagMisplaced));
        MeshBvh bvh;
        auto bvh_nodes = std::make_unique<DevBuf>(), bvh_recs = std::make_unique<DevBuf>();
        if (!moved && build_mesh_bvh(v, kd.center, bvh)) {
            upload(*bvh_nodes, bvh.nodes4, st);
            upload(*bvh_recs, bvh.recs, st);
        }
        MeshDev md{};
        md.nodes = nodes->as<KdNode>();
        md.recs = recs->as<float>();
        md.bvh = bvh.nodes.empty() ? nullptr : bvh_nodes->as<MeshBvh4Node>();
        md.bvh_recs = bvh.nodes.empty() ? nullptr : bvh_recs->as<float>();
        md.tri64 = tri64->as<double>();
        md.nb64 = nb64->as<double>();
        md.root_kd = kd.root_kd;
        md.root_bf = kd.root_bf;
        md.n_tris = (uint32_t)(v.size() / 9);
        md.cx = kd.center[0];
        md.cy = kd.center[1];
        md.cz = kd.center[2];
        double R = 0;
        for (size_t i = 0; i < v.size(); i += 3)
            R = std::max(R, std::fabs(v[i] - md.cx) + std::fabs(v[i + 1] - md.cy) + std::fabs(v[i + 2] - md.cz));
        md.radius = round_up(R * (1.0 + 1e-6));
        for (int k = 0; k < 3; ++k) {
            double c = k == 0 ? md.cx : (k == 1 ? md.cy : md.cz);
            md.bmin[k] = round_down(kd.bmin[k] - c);
            md.bmax[k] = round_up(kd.bmax[k] - c);
            md.bmin64[k] = kd.bmin[k];
            md.bmax64[k] = kd.bmax[k];
        }
        meshes.push_back(md);
        D.mesh_bufs.push_back(std::move(nodes));
        D.mesh_bufs.push_back(std::move(recs));
        D.mesh_bufs.push_back(std::move(tri64));
        D.mesh_bufs.push_back(std::move(nb64));
        D.mesh_bufs.push_back(std::move(bvh_nodes));
        D.mesh_bufs.push_back(std::move(bvh_recs));
    }
    for (const rt_object_desc& o : H.objects) {
        ObjDev d{};
        d.kind = o.kind;
        d.mat = o.material;
        d.mesh = o.mesh;
        for (int k = 0; k < 3; ++k) {
            d.c[k] = o.kind == RT_OBJ_PLANE ? o.point[k] : o.center[k];
            d.n[k] = o.normal[k];
            d.half[k] = o.half_size[k];
            d.cf[k] = (float)o.center[k];
            d.pc[k] = (float)d.c[k];
            d.pn[k] = (float)d.n[k];
        }
        d.r = o.radius;
        d.rf = round_up(o.radius);
        memcpy(d.w2l, o.world_to_local, sizeof(d.w2l));
        memcpy(d.l2w, o.local_to_world, sizeof(d.l2w));
        d.t_identity = is_identity(o.world_to_local);
        normal_matrix(o.world_to_local, d.ntf);
        if (o.kind == RT_OBJ_CUBE) cube_world_box(d);
        objs.push_back(d);
    }
    std::vector<MatDev> mats;
    for (const rt_material_desc& m : H.materials) {
        MatDev d{};
        memcpy(d.color, m.color, sizeof(d.color));
        d.ambient = m.ambient;
        d.diffuse = m.diffuse;
        d.specular = m.specular;
        d.shininess = m.shininess;
        d.has_refl = m.has_reflectivity;
        d.refl = m.reflectivity;
        d.texture = m.texture;
        memcpy(d.line_color, m.grid_line_color, sizeof(d.line_color));
        d.line_width = m.line_width;
        d.cell_size = m.cell_size;
        d.mat_b = m.material_b;
        mats.push_back(d);
    }
    std::vector<LightDev> lights;
    for (const rt_light_desc& l : H.lights) {
        LightDev d{};
        memcpy(d.pos, l.position, sizeof(d.pos));
        memcpy(d.color, l.color, sizeof(d.color));
        d.intensity = l.intensity;
        d.is_area = l.is_area;
        d.diameter = l.diameter;
        d.shape = l.shape == RT_LIGHT_SQUARE ? kLightSquare : kLightDisk;
        d.samples = l.samples > 0 ? l.samples : 16;
        d.size = l.size;
        d.cols = grid_cols(d.samples);
        if (l.is_area && d.shape == kLightDisk && !(l.diameter / 2.0 > -l.diameter / 2.0))
            throw Error(RT_E_INVALID_ARG, "cannot sample empty range");  // gen_range(-r..r) with r <= 0
        if (d.shape == kLightSquare && !(l.size > 0.0 && std::isfinite(l.size)))
            throw Error(RT_E_INVALID_ARG, "square light size must be positive");
        lights.push_back(d);
    }
    // spheres through an object BVH once there are enough of them
    std::vector<SphereBox> sph;
    std::vector<int32_t> others;
    bool has_cube = false;
    for (size_t i = 0; i < objs.size(); ++i) {
        if (objs[i].kind == kSphere) {
            SphereBox b{};
            for (int k = 0; k < 3; ++k) {
                double pad = 1e-6 * (std::fabs(objs[i].c[k]) + std::fabs(objs[i].r)) + 1e-12;
                b.lo[k] = round_down(objs[i].c[k] - std::fabs(objs[i].r) - pad);
                b.hi[k] = round_up(objs[i].c[k] + std::fabs(objs[i].r) + pad);
                b.c[k] = objs[i].c[k];
            }
            b.obj = (int32_t)i;
            sph.push_back(b);
        } else {
            others.push_back((int32_t)i);
            has_cube |= objs[i].kind == kCube;
        }
    }
    D.has_obvh = sph.size() >= 16;
    if (D.has_obvh) {
        std::vector<BvhNode> nodes;
        std::vector<int32_t> order;
        if (build_obvh(sph, nodes, order) > kObvhMaxDepth)
            throw Error(RT_E_INTERNAL, "object BVH deeper than its traversal stack");
        upload(D.obvh, nodes, st);
        upload(D.obvh_items, order, st);
        upload(D.others, others, st);
        D.n_others = (int)others.size();
        D.obvh_closest = !has_cube;
    }
    upload(D.objs, objs, st);
    upload(D.meshes, meshes, st);
    upload(D.mats, mats, st);
    upload(D.lights, lights, st);
    HIP_TRY(hipStreamSynchronize(st));
    D.n_obj = (int)objs.size();
    D.n_lights = (int)lights.size();
    D.ms_upload = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Camera::from_config (src/camera.rs:23-152) for this image's aspect ratio.
}  // namespace

CameraDev make_camera(const rt_camera_desc& c, double aspect) {
    d3 origin = mk(c.position[0], c.position[1], c.position[2]);
    d3 target = mk(c.target[0], c.target[1], c.target[2]);
    d3 up = mk(c.up[0], c.up[1], c.up[2]);
    d3 w = unit(sub(origin, target));
    d3 u = unit(cross(up, w));
    d3 v = cross(w, u);
    d3 view = unit(neg(w));
    CameraDev cam{};
    d3 hor, ver, llc;
    if (c.kind == 0) {
        double vh = c.height;
        double vw = fmax_r(c.width, vh * aspect);
        hor = scale(vw, u);
        ver = scale(vh, v);
        llc = sub(sub(origin, divs(hor, 2.0)), divs(ver, 2.0));
        cam.persp = 0;
    } else {
        double fov = c.has_fov ? c.fov : 45.0;
        if (fov <= 0.0 || fov >= 180.0) throw Error(RT_E_CAMERA, "Field of view must be between 0 and 180 degrees");
        double focal = 1.0;
        double theta = fov * (M_PI / 180.0);
        double hh = std::tan(theta / 2.0);
        double hw = aspect * hh;
        double vph = (2.0 * hh) * focal;
        double vpw = (2.0 * hw) * focal;
        hor = scale(vpw, u);
        ver = scale(vph, v);
        d3 center = add(origin, scale(focal, view));
        llc = sub(sub(center, divs(hor, 2.0)), divs(ver, 2.0));
        cam.persp = 1;
    }
    auto put = [](double* dst, d3 x) {
        dst[0] = x.x;
        dst[1] = x.y;
        dst[2] = x.z;
    };
    put(cam.origin, origin);
    put(cam.horizontal, hor);
    put(cam.vertical, ver);
    put(cam.llc, llc);
    put(cam.view, view);
    cam.has_grid = c.has_grid;
    cam.pitch = c.grid_pitch;
    cam.thick = c.grid_thickness;
    memcpy(cam.grid_color, c.grid_color, sizeof(cam.grid_color));
    return cam;
}

SceneView make_view(const SceneDevice& D, const CameraDev& cam, int light_samples, const double* cam_pos) {
    const HostScene& H = D.host;
    SceneView S{};
    S.objs = D.objs.as<ObjDev>();
    S.meshes = D.meshes.as<MeshDev>();
    S.mats = D.mats.as<MatDev>();
    S.lights = D.lights.as<LightDev>();
    S.n_obj = D.n_obj;
    S.n_lights = D.n_lights;
    // shadow-ray slots per hit: 1 per point light, `samples` per area light; sample points
    // are stored for point and disk lights (square samples are regenerated from the seed)
    int slots = 0, spt = 0;
    bool square = false;
    for (const rt_light_desc& l : H.lights) {
        int n = l.is_area ? (light_samples > 0 ? light_samples : (l.samples > 0 ? l.samples : 16)) : 1;
        slots += n;
        if (l.shape == RT_LIGHT_SQUARE) square = true;
        else spt += n;
    }
    S.shadow_per_hit = slots > 0 ? slots : 1;
    S.spt_per_hit = spt;
    S.has_square = square;
    S.light_samples = light_samples;
    S.light_cols = grid_cols(light_samples > 0 ? light_samples : 1);
    S.has_fog = H.has_fog;
    S.cam = cam;
    memcpy(S.cam_pos, cam_pos ? cam_pos : H.camera.position, sizeof(S.cam_pos));  // the rendering camera's
    memcpy(S.amb, H.ambient_color, sizeof(S.amb));
    S.amb_i = H.ambient_intensity;
    memcpy(S.fog_color, H.fog_color, sizeof(S.fog_color));
    S.fog_density = H.fog_density;
    S.fog_start = H.fog_start;
    S.fog_end = H.fog_end;
    memcpy(S.bg, H.background, sizeof(S.bg));
    if (D.has_obvh) {
        S.obvh = D.obvh.as<BvhNode>();
        S.obvh_items = D.obvh_items.as<int32_t>();
        S.others = D.others.as<int32_t>();
        S.n_others = D.n_others;
        S.obvh_closest = D.obvh_closest;
    }
    return S;
}

constexpr int kRaygenTile = 8;  // 8x8 pixel tile = one 64-lane wavefront

// Scalar fields of the frame view for options `o` with raygen tiles of `tile` pixels;
// validates the options and returns the queue sizes.
FrameLayout frame_fill(FrameView& F, const SceneView& S, const rt_render_opts* o, int tile) {
    if (o->samples == 0) throw Error(RT_E_INVALID_ARG, "Samples must be greater than 0");
    if (o->samples < 0 || o->width <= 0 || o->height <= 0) throw Error(RT_E_INVALID_ARG, "invalid image size or samples");
    if (o->aa_mode < 0 || o->aa_mode > 2) throw Error(RT_E_INVALID_ARG, "unknown anti-aliasing mode");
    if (tile != 4 && tile != 8 && tile != 16) throw Error(RT_E_INVALID_ARG, "raygen tile must be 4, 8 or 16");
    const int W = o->width, H = o->height;
    const bool quin = o->aa_mode == RT_AA_QUINCUNX;
    F.W = W;
    F.H = H;
    F.N = quin ? 1 : o->samples;
    F.aa = o->aa_mode;
    F.use_kd = o->use_kdtree;
    F.outline = (!quin && o->outline_enabled) ? 1 : 0;
    F.grid_miss = quin ? 0 : 1;  // quincunx calls ray_color without the camera
    F.count_work = o->count_work;
    F.seed = o->has_seed ? o->seed : 0;
    F.tile = tile;
    F.tiles_x = (W + tile - 1) / tile;
    const long long tiled = (long long)F.tiles_x * ((H + tile - 1) / tile) * tile * tile;
    FrameLayout L;
    if (quin) {
        const long long tiled_c = (long long)((W + tile) / tile) * ((H + tile) / tile) * tile * tile;
        L.n_paths = (long long)W * H + (long long)(W + 1) * (H + 1);
        L.n_q0 = tiled + tiled_c;
    } else {
        L.n_paths = (long long)W * H * F.N;
        L.n_q0 = tiled * F.N;
    }
    if (L.n_q0 > (1ll << 30)) throw Error(RT_E_INVALID_ARG, "image too large for one launch");
    F.n_paths = (int)L.n_paths;
    F.n_q0 = (int)L.n_q0;
    L.depth = std::max(o->max_depth, 0);
    F.max_depth = L.depth;
    L.cap = (size_t)L.n_q0;
    L.K = (size_t)S.shadow_per_hit;
    // shadow items are int32 indices; the work counters overshoot by one take per wavefront
    if (L.cap * L.K >= (1ull << 31) - (1ull << 24))
        throw Error(RT_E_INVALID_ARG, "too many shadow samples per frame for one launch");
    F.ol_wd = o->outline_depth_weight;
    F.ol_wn = o->outline_normal_weight;
    F.ol_thr = o->outline_threshold;
    F.ol_thick = o->outline_thickness;
    memcpy(F.ol_color, o->outline_color, sizeof(F.ol_color));
    F.ol_8 = o->outline_use_8_neighbors;
    return L;
}

// Places the frame's buffers in one allocation starting at `base` (256-byte aligned
// pieces); with base == nullptr only measures. Returns the bytes needed.
size_t frame_carve(FrameView& F, const SceneView& S, const FrameLayout& L, char* base, RayRec** q0, RayRec** q1,
                   bool own_rgb) {
    size_t off = 0;
    auto take = [&](size_t bytes) -> char* {
        char* p = base ? base + off : nullptr;
        off += (std::max(bytes, (size_t)16) + 255) & ~(size_t)255;
        return p;
    };
    const size_t np = (size_t)L.n_paths, cap = L.cap, px = (size_t)F.W * F.H;
    const size_t depth = (size_t)L.depth;
    F.path_seed = (uint64_t*)take(np * sizeof(uint64_t));
    F.lvl = (double*)take(std::max(depth, (size_t)1) * np * 4 * sizeof(double));
    F.prim_obj = (int32_t*)take(np * sizeof(int32_t));
    F.prim_tri = (int32_t*)take(np * sizeof(int32_t));
    F.prim_t = (double*)take(np * sizeof(double));
    F.ol_dn = (double*)take(F.outline ? np * 4 * sizeof(double) : 16);
    *q0 = (RayRec*)take(cap * sizeof(RayRec));
    *q1 = (RayRec*)take(cap * sizeof(RayRec));
    F.counts = (int32_t*)take((depth + 2) * sizeof(int32_t));
    F.hits = (HitRec*)take(cap * sizeof(HitRec));
    F.hitq = (int32_t*)take(cap * sizeof(int32_t));
    F.hit_count = (int32_t*)take((depth + 1) * sizeof(int32_t));
    F.tieq = (int32_t*)take(cap * sizeof(int32_t));
    F.tie_count = (int32_t*)take((depth + 1) * sizeof(int32_t));
    F.take = (int32_t*)take((3 * depth + 3) * 8 * 32 * sizeof(int32_t));
    F.shade = (ShadeRec*)take(cap * sizeof(ShadeRec));
    F.spt = (double*)take(cap * std::max(S.spt_per_hit, 1) * 3 * sizeof(double));
    F.spt_stride = (long long)cap;
    F.vis = (uint8_t*)take(cap * L.K);
    F.skip = (uint32_t*)take(cap * sizeof(uint32_t));
    F.work = (unsigned long long*)take(2 * kWorkCount * sizeof(unsigned long long));
    F.pix = (double*)take(px * 3 * sizeof(double));
    F.rgb = own_rgb ? (uint8_t*)take(px * 3) : nullptr;
    F.px_dn = (double*)take(F.outline ? px * 4 * sizeof(double) : 16);
    F.edge = (double*)take(F.outline ? px * sizeof(double) : 16);
    F.edge2 = (double*)take(F.outline ? px * sizeof(double) : 16);
    return off;
}

namespace {

// Camera rays traced: every path, or with a tile mask the masked pixels' rays (quincunx:
// their centres and every corner that touches one).
uint64_t primary_rays(const FrameView& F, long long n_paths) {
    if (!F.tmask) return (uint64_t)n_paths;
    const int tx = F.mtiles_x, ty = (F.H + F.mtile - 1) / F.mtile;
    std::vector<uint8_t> m((size_t)tx * ty);
    HIP_TRY(hipMemcpy(m.data(), F.tmask, m.size(), hipMemcpyDeviceToHost));
    // mask tile of pixel (x, y), or -1 outside the image or the mask
    auto tile_of = [&](int x, int y) -> long long {
        if (x < 0 || y < 0 || x >= F.W || y >= F.H) return -1;
        const long long t = (long long)(y / F.mtile) * tx + x / F.mtile;
        return m[t] ? t : -1;
    };
    // Visits the flagged tiles only. A corner is counted by the tile of its first masked
    // neighbour pixel (up-left, up-right, down-left, down-right), which is one of the
    // flagged tiles whose closed pixel range holds the corner.
    uint64_t n = 0;
    for (int j = 0; j < ty; ++j)
        for (int i = 0; i < tx; ++i) {
            const long long t = (long long)j * tx + i;
            if (!m[t]) continue;
            const int x0 = i * F.mtile, y0 = j * F.mtile;
            const int x1 = std::min(F.W, x0 + F.mtile), y1 = std::min(F.H, y0 + F.mtile);
            n += (uint64_t)(x1 - x0) * (y1 - y0) * (F.aa == 1 ? 1 : F.N);
            if (F.aa != 1) continue;
            for (int y = y0; y <= y1; ++y)
                for (int x = x0; x <= x1; ++x) {
                    long long owner = tile_of(x - 1, y - 1);
                    if (owner < 0) owner = tile_of(x, y - 1);
                    if (owner < 0) owner = tile_of(x - 1, y);
                    if (owner < 0) owner = tile_of(x, y);
                    n += owner == t ? 1 : 0;
                }
        }
    return n;
}

void render(rt_ctx* ctx, rt_scene* scene, const rt_render_opts* o, uint8_t* host_rgb, void* dev_rgb,
            rt_render_stats* stats, const rt_render_aux* aux = nullptr, const FixedScratch* fixed = nullptr,
            PendingFrame* pending = nullptr, hipStream_t on_stream = nullptr) {
    if (!scene || !scene->dev || !o) throw Error(RT_E_INVALID_ARG, "null scene or options");
    if (o->samples == 0) throw Error(RT_E_INVALID_ARG, "Samples must be greater than 0");
    if (o->samples < 0 || o->width <= 0 || o->height <= 0) throw Error(RT_E_INVALID_ARG, "invalid image size or samples");
    SceneDevice& D = *scene->dev;
    const int W = o->width, H = o->height;
    if (o->aa_mode < 0 || o->aa_mode > 2) throw Error(RT_E_INVALID_ARG, "unknown anti-aliasing mode");
    CameraDev cam = make_camera(D.host.camera, (double)W / (double)H);
    if (o->light_samples < 0 || o->light_samples > 65536) throw Error(RT_E_INVALID_ARG, "light_samples out of range");
    SceneView S = make_view(D, cam, o->light_samples);
    // a frame launched on another stream (a pending one with its own buffers) leaves the
    // context's stream, graph and timing events alone
    if (on_stream && (!pending || !fixed || o->profile))
        throw Error(RT_E_INTERNAL, "a frame on another stream must be pending, with its own buffers and untimed");
    hipStream_t st = on_stream ? on_stream : ctx->stream;

    FrameView F{};
    const FrameLayout L = frame_fill(F, S, o, kRaygenTile);
    const int depth = L.depth;
    const long long n_paths = L.n_paths;
    // every option check comes before the arena changes, so a refused render leaves the
    // previous frame's buffers (and rt_last_primary_hits) as they were
    if (aux && aux->tile_mask) {
        if (aux->mask_tile <= 0) throw Error(RT_E_INVALID_ARG, "mask_tile must be positive");
        if (F.outline) throw Error(RT_E_INVALID_ARG, "a tile mask cannot be combined with the outline pass");
        if (host_rgb) throw Error(RT_E_INVALID_ARG, "a tile mask needs a device output");
    }
    if (fixed && depth > kNoAllocMaxDepth)
        throw Error(RT_E_INVALID_ARG, "max_depth above " + std::to_string(kNoAllocMaxDepth) +
                                          " needs the context's own frame buffers (rt_render_device)");
    RayRec *q0 = nullptr, *q1 = nullptr;
    const size_t bytes = frame_carve(F, S, L, nullptr, &q0, &q1, dev_rgb == nullptr);
    char* base = nullptr;
    if (fixed) {
        if (bytes > fixed->bytes)
            throw Error(RT_E_INVALID_ARG, "frame needs " + std::to_string(bytes) +
                                              " bytes of scratch; the no-allocation entry point holds at most " +
                                              std::to_string(fixed->bytes));
        base = fixed->base;
    } else {
        if (bytes > ctx->fb.arena.bytes) {
            // the arena moves: the last frame's primary records go with it
            ctx->last_paths = 0;
            ctx->last_prim_obj = ctx->last_prim_tri = nullptr;
            ctx->last_prim_t = nullptr;
        }
        ctx->fb.arena.reserve(bytes);
        base = ctx->fb.arena.as<char>();
    }
    frame_carve(F, S, L, base, &q0, &q1, dev_rgb == nullptr);
    if (dev_rgb) F.rgb = static_cast<uint8_t*>(dev_rgb);
    if (aux) {
        F.feat = aux->features;
        F.rad = aux->radiance;
        if (aux->tile_mask) {
            F.tmask = aux->tile_mask;
            F.mtile = aux->mask_tile;
            F.mtiles_x = (W + aux->mask_tile - 1) / aux->mask_tile;
        }
    }

    if (depth == 0) {
        // ray_color returns black at max_depth <= 0 before any World::hit; mark the
        // primary records as misses
        HIP_TRY(hipMemsetAsync(F.prim_obj, 0xff, n_paths * sizeof(int32_t), st));
        HIP_TRY(hipMemsetAsync(F.prim_tri, 0xff, n_paths * sizeof(int32_t), st));
    }
    ctx->timer.reset();
    const bool timed_stages = o->profile && !fixed;
    // Without per-stage timing, a frame rendered twice in a row with the same kernel
    // arguments is captured into a hipGraph (kernels only) and replayed while they stay the
    // same; a one-off frame launches eagerly and pays no capture. The no-allocation entry
    // point always launches eagerly (instantiating a graph allocates).
    std::vector<char> key;
    bool replay = false;
    if (!o->profile && !fixed) {
        key.resize(sizeof(S) + sizeof(F) + 2 * sizeof(RayRec*));
        memcpy(key.data(), &S, sizeof(S));
        memcpy(key.data() + sizeof(S), &F, sizeof(F));
        memcpy(key.data() + sizeof(S) + sizeof(F), &q0, sizeof(q0));
        memcpy(key.data() + sizeof(S) + sizeof(F) + sizeof(q0), &q1, sizeof(q1));
        replay = ctx->frame_exec && key == ctx->frame_key;
        const bool repeat = key == ctx->seen_key;
        ctx->seen_key = key;
        if (!replay && repeat) {
            ctx->drop_frame_graph();
            warm_frame_launches();
            HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            hipError_t le = launch_frame(S, F, q0, q1, st, nullptr);
            hipGraph_t g = nullptr;
            hipError_t ce = hipStreamEndCapture(st, &g);
            ctx->frame_graph = g;
            HIP_TRY(le);
            HIP_TRY(ce);
            HIP_TRY(hipGraphInstantiate(&ctx->frame_exec, ctx->frame_graph, nullptr, nullptr, 0));
            ctx->frame_key = key;
            replay = true;
        }
    }
    if (!on_stream) HIP_TRY(hipEventRecord(ctx->ev0, st));
    if (replay) HIP_TRY(hipGraphLaunch(ctx->frame_exec, st));
    else HIP_TRY(launch_frame(S, F, q0, q1, st, timed_stages ? &ctx->timer : nullptr));
    if (!on_stream) HIP_TRY(hipEventRecord(ctx->ev1, st));
    if (host_rgb) HIP_TRY(hipMemcpyAsync(host_rgb, F.rgb, (size_t)W * H * 3, hipMemcpyDeviceToHost, st));
    PendingFrame pf{F, S, n_paths, depth, timed_stages, fixed != nullptr, &D};
    if (pending) {
        *pending = pf;  // the caller reads the stats once it has synchronised the stream
        return;
    }
    frame_stats(ctx, pf, stats, true);  // waits for the stream
}

}  // namespace

// Stats of a launched frame (waits for ctx's stream): per-level queue lengths, work
// counters and, for the context's last frame, its device time and stage times.
// `record_primary` makes it the frame rt_last_primary_hits reads.
void frame_stats(rt_ctx* ctx, const PendingFrame& pf, rt_render_stats* stats, bool record_primary) {
    hipStream_t st = ctx->stream;
    const FrameView& F = pf.F;
    const int depth = pf.depth;
    // per-level queue lengths (fixed host arrays on the no-allocation path)
    std::vector<int32_t> counts_v, hits_v;
    int32_t counts_a[kNoAllocMaxDepth + 2] = {}, hits_a[kNoAllocMaxDepth + 1] = {};
    int32_t* counts = counts_a;
    int32_t* hit_count = hits_a;
    if (!pf.fixed) {
        counts_v.assign(depth + 2, 0);
        hits_v.assign(depth + 1, 0);
        counts = counts_v.data();
        hit_count = hits_v.data();
    }
    unsigned long long work[2 * kWorkCount] = {};
    HIP_TRY(hipMemcpyAsync(counts, F.counts, (depth + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    if (depth > 0)
        HIP_TRY(hipMemcpyAsync(hit_count, F.hit_count, depth * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(work, F.work, sizeof(work), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (record_primary) {
        ctx->last_paths = (int)pf.n_paths;
        ctx->last_prim_obj = F.prim_obj;
        ctx->last_prim_tri = F.prim_tri;
        ctx->last_prim_t = F.prim_t;
    }
    if (!stats) return;
    memset(stats, 0, sizeof(*stats));
    stats->rays_primary = depth > 0 ? primary_rays(F, pf.n_paths) : 0;
    for (int l = 1; l < depth; ++l) stats->rays_reflect += (uint64_t)counts[l];
    for (int l = 0; l < depth; ++l)
        stats->rays_shadow += (uint64_t)hit_count[l] * (uint64_t)(pf.S.n_lights > 0 ? pf.S.shadow_per_hit : 0);
    for (int k = 0; k < 2; ++k) {
        const unsigned long long* wk = work + k * kWorkCount;
        stats->work[k].nodes = wk[kWorkNodes];
        stats->work[k].tris = wk[kWorkTris];
        stats->work[k].retests = wk[kWorkRetests];
        stats->work[k].spheres = wk[kWorkSphere];
        stats->work[k].planes = wk[kWorkPlane];
        stats->work[k].cubes = wk[kWorkCube];
        stats->work[k].skipped = wk[kWorkSkipped];
        stats->work[k].bvh_nodes = wk[kWorkBvhNodes];
    }
    stats->ms_upload = pf.D->ms_upload;
    stats->levels = depth;
    if (!record_primary) return;  // a pending frame: its events may have been recorded over
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    stats->ms_render = ms;
    if (pf.timed_stages) {
        for (size_t i = 0; i < ctx->timer.used / 2; ++i) {
            float e = 0;
            HIP_TRY(hipEventElapsedTime(&e, ctx->timer.ev[2 * i], ctx->timer.ev[2 * i + 1]));
            int sidx = ctx->timer.stage[i];
            stats->ms_stage[sidx] += e;
            stats->launches_stage[sidx] += 1;
        }
    }
}

namespace {

int fail(rt_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    g_tls_error = msg;
    return code;
}

template <class F>
int guarded(rt_ctx* ctx, F fn) {
    try {
        fn();
        return RT_OK;
    } catch (const Error& e) {
        return fail(ctx, e.code, e.what());
    } catch (const std::bad_alloc&) {
        return fail(ctx, RT_E_OOM, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(ctx, RT_E_INTERNAL, e.what());
    }
}

void check_device(int device) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) throw Error(RT_E_DEVICE, "no HIP device available");
    if (device < 0 || device >= n) throw Error(RT_E_DEVICE, "device index out of range");
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (std::string(prop.gcnArchName).rfind("gfx950", 0) != 0)
        throw Error(RT_E_DEVICE, std::string("librtrace_hip is built for gfx950, device is ") + prop.gcnArchName);
}

rt_scene* make_scene(rt_ctx* ctx, HostScene&& hs) {
    auto s = std::make_unique<rt_scene>();
    s->ctx = ctx;
    s->dev = std::make_unique<SceneDevice>();
    s->dev->host = std::move(hs);
    s->dev->device = ctx->device;
    HIP_TRY(hipSetDevice(ctx->device));
    build_device_scene(ctx, *s->dev);
    return s.release();
}

}  // namespace
}  // namespace rt

// Frame buffers of rt_render_device_noalloc: a fixed device array, so that entry point
// allocates nothing. Each device holds its own copy at its own address until the process
// ends; the address is resolved once per device (scratch_on_device) and only that address,
// never the array's name, is passed to kernels or copies. Every call on one device shares
// the array and finds the last call's contents: the frame kernels clear what must start at
// zero (k_init_counts), and calls that overlap on one device race.
static __device__ __align__(256) char scratch[RT_NOALLOC_SCRATCH_BYTES];

namespace rt {
namespace {

std::mutex g_scratch_mu;
std::vector<char*> g_scratch_addr;  // per device, resolved once

// Device address of `scratch` on `device`, which must be the current device.
char* scratch_on_device(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((size_t)device >= g_scratch_addr.size()) g_scratch_addr.resize((size_t)device + 1, nullptr);
    if (!g_scratch_addr[device]) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw Error(RT_E_DEVICE, "hipGetSymbolAddress returned no address for the no-allocation scratch");
        g_scratch_addr[device] = static_cast<char*>(p);
    }
    return g_scratch_addr[device];
}

}  // namespace
}  // namespace rt

namespace rt {
// Context access for the other host modules (host/context.hpp).
hipStream_t ctx_stream(rt_ctx* ctx) { return ctx->stream; }
int ctx_device(rt_ctx* ctx) { return ctx->device; }
rt_ctx* scene_ctx(rt_scene* scene) { return scene->ctx; }
int ctx_fail(rt_ctx* ctx, int code, const std::string& msg) { return fail(ctx, code, msg); }
void ctx_render_aux(rt_ctx* ctx, rt_scene* scene, const rt_render_opts* o, void* dev_rgb, const rt_render_aux* aux,
                    rt_render_stats* stats) {
    render(ctx, scene, o, nullptr, dev_rgb, stats, aux);
}
void ctx_render_launch(rt_ctx* ctx, rt_scene* scene, const rt_render_opts* o, void* dev_rgb, const rt_render_aux* aux,
                       const FixedScratch* arena, PendingFrame* pending, hipStream_t on_stream) {
    render(ctx, scene, o, nullptr, dev_rgb, nullptr, aux, arena, pending, on_stream);
}
size_t ctx_frame_bytes(rt_scene* scene, const rt_render_opts* o) {
    SceneDevice& D = *scene->dev;
    CameraDev cam = make_camera(D.host.camera, (double)o->width / (double)o->height);
    SceneView S = make_view(D, cam, o->light_samples);
    FrameView F{};
    const FrameLayout L = frame_fill(F, S, o, kRaygenTile);
    RayRec *q0 = nullptr, *q1 = nullptr;
    return frame_carve(F, S, L, nullptr, &q0, &q1, false);
}
}  // namespace rt

using namespace rt;

extern "C" {

int rt_ctx_create(int device, rt_ctx** out) {
    if (!out) return fail(nullptr, RT_E_INVALID_ARG, "out is null");
    *out = nullptr;
    auto c = std::make_unique<rt_ctx>();
    int rc = guarded(nullptr, [&] {
        check_device(device);
        HIP_TRY(hipSetDevice(device));
        HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        HIP_TRY(hipEventCreate(&c->ev0));
        HIP_TRY(hipEventCreate(&c->ev1));
        c->device = device;
        c->noalloc_scratch = scratch_on_device(device);
    });
    if (rc == RT_OK) *out = c.release();
    return rc;
}

void rt_ctx_destroy(rt_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);

    ctx->drop_frame_graph();
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* rt_last_error(rt_ctx* ctx) { return ctx ? ctx->err.c_str() : g_tls_error.c_str(); }

int rt_scene_load_json_file(rt_ctx* ctx, const char* path, rt_scene** out) {
    if (!ctx || !path || !out) return fail(ctx, RT_E_INVALID_ARG, "null argument");
    *out = nullptr;
    return guarded(ctx, [&] { *out = make_scene(ctx, load_scene_file(path)); });
}

int rt_scene_load_json_str(rt_ctx* ctx, const char* json, const char* base_dir, rt_scene** out) {
    if (!ctx || !json || !out) return fail(ctx, RT_E_INVALID_ARG, "null argument");
    *out = nullptr;
    return guarded(ctx, [&] { *out = make_scene(ctx, load_scene_json(json, base_dir ? base_dir : ".")); });
}

int rt_scene_upload(rt_ctx* ctx, const rt_scene_desc* desc, rt_scene** out) {
    if (!ctx || !desc || !out) return fail(ctx, RT_E_INVALID_ARG, "null argument");
    *out = nullptr;
    return guarded(ctx, [&] { *out = make_scene(ctx, from_desc(*desc)); });
}

void rt_scene_free(rt_scene* scene) {
    if (!scene) return;
    if (scene->ctx && scene->ctx->stream) (void)hipStreamSynchronize(scene->ctx->stream);
    delete scene;
}

int rt_scene_info(rt_scene* scene, double* cam_width, double* cam_height, int32_t* outline_enabled,
                  rt_render_opts* outline_out) {
    if (!scene || !scene->dev) return fail(nullptr, RT_E_INVALID_ARG, "null scene");
    const HostScene& H = scene->dev->host;
    if (cam_width) *cam_width = H.camera.width;
    if (cam_height) *cam_height = H.camera.height;
    if (outline_enabled) *outline_enabled = H.outline_enabled;
    if (H.outline_enabled && !H.outline_color_ok)
        return fail(scene->ctx, RT_E_CAMERA, "Invalid hex color format");
    if (outline_out && H.outline_enabled) {
        outline_out->outline_enabled = 1;
        outline_out->outline_use_8_neighbors = H.outline.outline_use_8_neighbors;
        outline_out->outline_depth_weight = H.outline.outline_depth_weight;
        outline_out->outline_normal_weight = H.outline.outline_normal_weight;
        outline_out->outline_threshold = H.outline.outline_threshold;
        memcpy(outline_out->outline_color, H.outline.outline_color, sizeof(outline_out->outline_color));
        outline_out->outline_thickness = H.outline.outline_thickness;
    }
    return RT_OK;
}

int rt_scene_counts(rt_scene* scene, int32_t* n_objects, int32_t* n_lights) {
    if (!scene || !scene->dev) return fail(nullptr, RT_E_INVALID_ARG, "null scene");
    if (n_objects) *n_objects = (int32_t)scene->dev->host.objects.size();
    if (n_lights) *n_lights = (int32_t)scene->dev->host.lights.size();
    return RT_OK;
}

void rt_render_opts_default(rt_render_opts* o, int32_t width, int32_t height) {
    if (!o) return;
    memset(o, 0, sizeof(*o));
    o->width = width;
    o->height = height;
    o->max_depth = 10;
    o->use_kdtree = 1;
    o->samples = 1;
    o->aa_mode = RT_AA_QUINCUNX;
    o->has_seed = 1;
    o->seed = 0;
    o->outline_depth_weight = 1.0;
    o->outline_normal_weight = 1.0;
    o->outline_threshold = 0.1;
    o->outline_thickness = 1.0;
}

int rt_render(rt_ctx* ctx, rt_scene* scene, const rt_render_opts* opts, uint8_t* rgb_out, rt_render_stats* stats) {
    if (!ctx || !rgb_out) return fail(ctx, RT_E_INVALID_ARG, "null argument");
    return guarded(ctx, [&] { render(ctx, scene, opts, rgb_out, nullptr, stats); });
}

int rt_render_device(rt_ctx* ctx, rt_scene* scene, const rt_render_opts* opts, void* d_rgb_out,
                     rt_render_stats* stats) {
    if (!ctx || !d_rgb_out) return fail(ctx, RT_E_INVALID_ARG, "null argument");
    return guarded(ctx, [&] { render(ctx, scene, opts, nullptr, d_rgb_out, stats); });
}

int rt_render_device_noalloc(rt_ctx* ctx, rt_scene* scene, const rt_render_opts* opts, void* d_rgb_out,
                             rt_render_stats* stats) {
    if (!ctx || !d_rgb_out) return fail(ctx, RT_E_INVALID_ARG, "null argument");
    return guarded(ctx, [&] {
        if (!ctx->noalloc_scratch) throw Error(RT_E_DEVICE, "the context has no no-allocation scratch");
        const FixedScratch fixed{ctx->noalloc_scratch, (size_t)RT_NOALLOC_SCRATCH_BYTES};
        render(ctx, scene, opts, nullptr, d_rgb_out, stats, nullptr, &fixed);
    });
}

int rt_render_device_aux(rt_ctx* ctx, rt_scene* scene, const rt_render_opts* opts, void* d_rgb_out,
                         const rt_render_aux* aux, rt_render_stats* stats) {
    if (!ctx || !d_rgb_out) return fail(ctx, RT_E_INVALID_ARG, "null argument");
    return guarded(ctx, [&] { render(ctx, scene, opts, nullptr, d_rgb_out, stats, aux); });
}

int rt_last_primary_hits(rt_ctx* ctx, int32_t* obj, int32_t* tri, double* t, size_t n) {
    if (!ctx) return fail(nullptr, RT_E_INVALID_ARG, "null context");
    return guarded(ctx, [&] {
        size_t m = std::min(n, (size_t)ctx->last_paths);
        if (obj) HIP_TRY(hipMemcpy(obj, ctx->last_prim_obj, m * sizeof(int32_t), hipMemcpyDeviceToHost));
        if (tri) HIP_TRY(hipMemcpy(tri, ctx->last_prim_tri, m * sizeof(int32_t), hipMemcpyDeviceToHost));
        if (t) HIP_TRY(hipMemcpy(t, ctx->last_prim_t, m * sizeof(double), hipMemcpyDeviceToHost));
    });
}

int rt_kd_build(rt_ctx* ctx, const double* vertices, int64_t n_tris, int32_t on_device, rt_kd_arrays* out) {
    if ((!ctx && on_device) || !out || n_tris < 0 || (n_tris > 0 && !vertices))
        return fail(ctx, RT_E_INVALID_ARG, "null argument");
    return guarded(ctx, [&] {
        if (n_tris > (int64_t)(1u << 28)) throw Error(RT_E_INVALID_ARG, "too many triangles");
        const std::vector<double> v(vertices, vertices + (size_t)n_tris * 9);
        FlatKd kd;
        std::vector<float> recs;
        if (on_device) {
            hipStream_t st = ctx->stream;
            DevBuf tri64, nodes, drecs, nb64;
            upload(tri64, v, st);
            build_flat_kd_device(v, tri64.as<double>(), kd, nodes, drecs, nb64, st);
            const size_t n_recs = (size_t)kd.nodes[kd.root_bf].a + (size_t)n_tris;  // the brute-force leaf is last
            const size_t want = std::min<size_t>(n_recs, (size_t)std::max<int64_t>(out->recs_cap, 0));
            recs.resize(n_recs * 12);
            if (want) HIP_TRY(hipMemcpyAsync(recs.data(), drecs.p, want * 48, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
        } else {
            build_flat_kd(v, kd);
            recs = std::move(kd.recs);
        }
        out->n_nodes = (int64_t)kd.nodes.size();
        out->n_recs = (int64_t)(recs.size() / 12);
        out->root_kd = kd.root_kd;
        out->root_bf = kd.root_bf;
        out->leaves = kd.leaves;
        out->max_leaf = kd.max_leaf;
        if (out->nodes)
            memcpy(out->nodes, kd.nodes.data(), std::min<size_t>(kd.nodes.size(), (size_t)std::max<int64_t>(out->nodes_cap, 0)) * 48);
        if (out->recs)
            memcpy(out->recs, recs.data(), std::min<size_t>(recs.size() / 12, (size_t)std::max<int64_t>(out->recs_cap, 0)) * 48);
        if (out->bounds)
            memcpy(out->bounds, kd.nb64.data(),
                   std::min<size_t>(kd.nb64.size() / 6, (size_t)std::max<int64_t>(out->bounds_cap, 0)) * 48);
    });
}

int rt_intersect(rt_ctx* ctx, rt_scene* scene, const double* origins, const double* dirs, size_t n, double t_min,
                 double t_max, int use_kdtree, int32_t* obj_out, int32_t* tri_out, double* t_out) {
    if (!ctx || !scene || !scene->dev || (n && (!origins || !dirs || !obj_out || !tri_out || !t_out)))
        return fail(ctx, RT_E_INVALID_ARG, "null argument");
    if (n == 0) return RT_OK;
    if (n > (size_t)INT32_MAX) return fail(ctx, RT_E_INVALID_ARG, "too many rays");
    return guarded(ctx, [&] {
        SceneDevice& D = *scene->dev;
        SceneView S = make_view(D, CameraDev{}, 0);  // World::hit needs no camera
        DevBuf o, d, ob, tr, tt;
        o.reserve(n * 3 * sizeof(double));
        d.reserve(n * 3 * sizeof(double));
        ob.reserve(n * sizeof(int32_t));
        tr.reserve(n * sizeof(int32_t));
        tt.reserve(n * sizeof(double));
        hipStream_t st = ctx->stream;
        HIP_TRY(hipMemcpyAsync(o.p, origins, n * 3 * sizeof(double), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(d.p, dirs, n * 3 * sizeof(double), hipMemcpyHostToDevice, st));
        HIP_TRY(launch_intersect(S, o.as<double>(), d.as<double>(), (int)n, t_min, t_max, use_kdtree, ob.as<int32_t>(),
                                 tr.as<int32_t>(), tt.as<double>(), st));
        HIP_TRY(hipMemcpyAsync(obj_out, ob.p, n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(tri_out, tr.p, n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(t_out, tt.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    });
}

}  // extern "C"
