// rtrace_napi.cpp — Node addon over librtrace_hip with the reference binding's eight
// functions (index.d.ts:7-21, bindings/node/src/lib.rs): the same names, arguments,
// sizing, error statuses and result strings, rendering on the HIP path. The thread
// counts of the *Threaded variants are accepted and reported as the reference does; the
// image does not depend on them (as the reference's does not, src/renderer.rs:1119-1171).
#define NAPI_VERSION 3
#include <node_api.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../../include/rtrace_hip.h"
#include "../cli/png.hpp"

namespace {

rt_ctx* g_ctx = nullptr;

napi_value make_string(napi_env env, const std::string& s) {
    napi_value v;
    napi_create_string_utf8(env, s.c_str(), s.size(), &v);
    return v;
}

napi_value fail(napi_env env, const char* status, const std::string& msg) {
    napi_throw_error(env, status, msg.c_str());
    return nullptr;
}

bool arg_string(napi_env env, napi_value v, std::string& out) {
    size_t len = 0;
    if (napi_get_value_string_utf8(env, v, nullptr, 0, &len) != napi_ok) return false;
    out.resize(len);
    std::vector<char> buf(len + 1);
    napi_get_value_string_utf8(env, v, buf.data(), len + 1, &len);
    out.assign(buf.data(), len);
    return true;
}

// Option<u32>: absent, undefined or null -> none
bool arg_opt_u32(napi_env env, size_t argc, napi_value* argv, size_t i, bool& has, uint32_t& out) {
    has = false;
    if (i >= argc) return true;
    napi_valuetype t;
    napi_typeof(env, argv[i], &t);
    if (t == napi_undefined || t == napi_null) return true;
    if (t != napi_number) return false;
    double d = 0;
    napi_get_value_double(env, argv[i], &d);
    if (!(d >= 0 && d <= 4294967295.0)) return false;
    out = (uint32_t)d;
    has = true;
    return true;
}

uint32_t sat_u32(double x) {  // Rust `as u32`
    if (!(x > 0)) return 0;
    return x >= 4294967295.0 ? 4294967295u : (uint32_t)x;
}

enum class Src { kJson, kFile };
enum class Mode { kDefault, kThreaded, kBrute };

// The common body of the reference's render_scene* functions (bindings/node/src/lib.rs:18-330).
napi_value render(napi_env env, napi_callback_info info, Src src, Mode mode) {
    size_t argc = 4;
    napi_value argv[4];
    napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr);
    std::string scene_arg, output;
    if (argc < 2 || !arg_string(env, argv[0], scene_arg) || !arg_string(env, argv[1], output))
        return fail(env, "InvalidArg", "expected (string, string, size?)");
    bool has_size = false, has_threads = false;
    uint32_t size = 1000, threads = 0;
    if (!arg_opt_u32(env, argc, argv, 2, has_size, size)) return fail(env, "InvalidArg", "size must be a u32");
    if (!has_size) size = 1000;
    if (mode == Mode::kThreaded && !arg_opt_u32(env, argc, argv, 3, has_threads, threads))
        return fail(env, "InvalidArg", "threadCount must be a u32");
    if (!g_ctx && rt_ctx_create(0, &g_ctx) != RT_OK) {
        g_ctx = nullptr;
        return fail(env, "GenericFailure", std::string("Failed to render scene: ") + rt_last_error(nullptr));
    }
    rt_scene* scene = nullptr;
    int rc = src == Src::kFile ? rt_scene_load_json_file(g_ctx, scene_arg.c_str(), &scene)
                               : rt_scene_load_json_str(g_ctx, scene_arg.c_str(), ".", &scene);
    if (rc != RT_OK)
        return fail(env, "InvalidArg",
                    std::string(src == Src::kFile ? "Failed to load scene file: " : "Failed to parse scene JSON: ") +
                        rt_last_error(g_ctx));
    // pixel size from the diagonal and the camera aspect (bindings/node/src/lib.rs:31-45)
    double cw = 0, ch = 0;
    rt_scene_info(scene, &cw, &ch, nullptr, nullptr);
    const double aspect = cw / ch, diag = (double)size;
    const double hf = diag / std::sqrt(aspect * aspect + 1.0);
    const double wf = aspect * hf;
    const uint32_t W = sat_u32(std::round(wf)), H = sat_u32(std::round(hf));
    rt_render_opts o;
    rt_render_opts_default(&o, (int32_t)W, (int32_t)H);  // Renderer::new / new_with_threads
    if (mode == Mode::kBrute) o.use_kdtree = 0;            // Renderer::new_brute_force
    std::vector<uint8_t> rgb((size_t)W * H * 3);
    rc = rt_render(g_ctx, scene, &o, rgb.data(), nullptr);
    if (rc != RT_OK) {
        std::string e = rt_last_error(g_ctx);
        rt_scene_free(scene);
        return fail(env, "GenericFailure", "Failed to render scene: " + e);
    }
    rt_scene_free(scene);
    if (!rtpng::write_png(output.c_str(), rgb.data(), (int)W, (int)H))
        return fail(env, "GenericFailure", "Failed to render scene: could not write '" + output + "'");
    std::string msg = "Successfully rendered " + std::to_string(W) + "×" + std::to_string(H) + " image (diagonal " +
                      std::to_string(size) + ") to '" + output + "'";
    if (mode == Mode::kDefault) msg += " (multi-threaded)";
    else if (mode == Mode::kBrute) msg += " (brute force)";
    else msg += has_threads ? " with " + std::to_string(threads) + " threads" : " with all available threads";
    return make_string(env, msg);
}

napi_value HelloWorld(napi_env env, napi_callback_info) { return make_string(env, "hello world"); }
napi_value GreetWithName(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1];
    napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr);
    std::string name;
    if (argc < 1 || !arg_string(env, argv[0], name)) return fail(env, "InvalidArg", "expected a string");
    return make_string(env, "hello world, " + name);
}
napi_value RenderScene(napi_env e, napi_callback_info i) { return render(e, i, Src::kJson, Mode::kDefault); }
napi_value RenderSceneThreaded(napi_env e, napi_callback_info i) { return render(e, i, Src::kJson, Mode::kThreaded); }
napi_value RenderSceneBruteForce(napi_env e, napi_callback_info i) { return render(e, i, Src::kJson, Mode::kBrute); }
napi_value RenderSceneFromFile(napi_env e, napi_callback_info i) { return render(e, i, Src::kFile, Mode::kDefault); }
napi_value RenderSceneFromFileThreaded(napi_env e, napi_callback_info i) {
    return render(e, i, Src::kFile, Mode::kThreaded);
}
napi_value RenderSceneFromFileBruteForce(napi_env e, napi_callback_info i) {
    return render(e, i, Src::kFile, Mode::kBrute);
}

napi_value Init(napi_env env, napi_value exports) {
    // plain data properties, like the exports napi-rs generates
    const auto kAttr = static_cast<napi_property_attributes>(napi_writable | napi_enumerable | napi_configurable);
    const napi_property_descriptor fns[] = {
        {"helloWorld", nullptr, HelloWorld, nullptr, nullptr, nullptr, kAttr, nullptr},
        {"greetWithName", nullptr, GreetWithName, nullptr, nullptr, nullptr, kAttr, nullptr},
        {"renderScene", nullptr, RenderScene, nullptr, nullptr, nullptr, kAttr, nullptr},
        {"renderSceneThreaded", nullptr, RenderSceneThreaded, nullptr, nullptr, nullptr, kAttr, nullptr},
        {"renderSceneBruteForce", nullptr, RenderSceneBruteForce, nullptr, nullptr, nullptr, kAttr, nullptr},
        {"renderSceneFromFile", nullptr, RenderSceneFromFile, nullptr, nullptr, nullptr, kAttr, nullptr},
        {"renderSceneFromFileThreaded", nullptr, RenderSceneFromFileThreaded, nullptr, nullptr, nullptr, kAttr,
         nullptr},
        {"renderSceneFromFileBruteForce", nullptr, RenderSceneFromFileBruteForce, nullptr, nullptr, nullptr,
         kAttr, nullptr},
    };
    napi_define_properties(env, exports, sizeof(fns) / sizeof(fns[0]), fns);
    return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
