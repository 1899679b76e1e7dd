// scene_host.hpp — host-side scene model for the trace path: the owning form of
// rt_scene_desc plus the outline block of scene_settings, and the loaders that fill
// it from Scene JSON (src/scene.rs) and STL (src/mesh.rs:751-1003).
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/rtrace_hip.h"

namespace rt {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

struct ObjBounds {
    bool finite = false;
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
};

struct HostScene {
    rt_camera_desc camera{};
    std::vector<rt_object_desc> objects;
    std::vector<rt_material_desc> materials;  // [0, n_objects) per object, then checkerboard alternates
    std::vector<rt_light_desc> lights;
    std::vector<std::vector<double>> meshes;  // f64 vertices, 9 per triangle, transform applied
    double ambient_color[3] = {1, 1, 1};
    double ambient_intensity = 0.1;
    bool has_fog = false;
    double fog_color[3] = {0.5, 0.5, 0.5};
    double fog_density = 0, fog_start = 0, fog_end = 0;
    double background[3] = {0, 0, 0};
    // scene_settings.outline (src/scene.rs:382-417, 515-534)
    bool outline_enabled = false;
    bool outline_color_ok = true;
    rt_render_opts outline{};

    // Scene::compute_finite_bounds per object (src/scene.rs:538-660): planes are not finite;
    // a transformed mesh contributes its untransformed box's 8 corners, transformed
    std::vector<ObjBounds> bounds;

    rt_scene_desc desc(std::vector<rt_mesh_desc>& mesh_storage) const;
};

HostScene load_scene_json(const std::string& text, const std::string& base_dir);
HostScene load_scene_file(const std::string& path);
HostScene from_desc(const rt_scene_desc& d);
// AutoCamera::generate_cameras (src/auto_camera.rs:9-130) from HostScene::bounds:
// left, front, top, perspective.
void auto_cameras(const HostScene& S, rt_camera_desc out[4]);

// STL loader (src/mesh.rs:751-1003): returns 9 f64 per triangle.
std::vector<double> load_stl_file(const std::string& path);

// Transform strings -> 4x4 column-major (src/scene.rs:117-203). false on a parse error.
bool parse_transform_list(const std::vector<std::string>& list, double out[16]);
// nalgebra try_inverse (4x4 cofactor form). false when singular.
bool invert4(const double m[16], double out[16]);
bool hex_color(const std::string& s, double out[3]);

}  // namespace rt
