// frame.hpp — host-side types shared by the frame drivers (render.cpp, batch.cpp):
// device buffers, the uploaded scene, the context, and the frame set-up helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "../../../include/rtrace_hip.h"
#include "../common/layout.hpp"
#include "../kernels/launch.hpp"
#include "../kernels/trace_core.hpp"
#include "kdbuild.hpp"
#include "scene_host.hpp"

namespace rt {

#define HIP_TRY(expr)                                                                             \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            throw Error(e_ == hipErrorOutOfMemory ? RT_E_OOM : RT_E_DEVICE,                       \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                       \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    void reserve(size_t n) {
        if (n <= bytes && p) return;
        if (p) {
            (void)hipFree(p);
            p = nullptr;
            bytes = 0;
        }
        size_t want = n ? n : 16;
        HIP_TRY(hipMalloc(&p, want));
        bytes = want;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};


struct SceneDevice {
    HostScene host;
    int device = 0;
    DevBuf objs, meshes, mats, lights, obvh, obvh_items, others;
    bool has_obvh = false, obvh_closest = false;
    int n_others = 0;
    std::vector<std::unique_ptr<DevBuf>> mesh_bufs;
    int n_obj = 0, n_lights = 0;
    double ms_upload = 0;
    std::vector<FlatKd> kd;  // host copies (stats)
};

struct FrameBuffers {
    DevBuf arena;  // every per-frame buffer, carved by frame_carve
};

struct FrameLayout {
    long long n_paths = 0, n_q0 = 0;
    int depth = 0;
    size_t cap = 0;  // level-0 queue slots (tiled)
    size_t K = 1;    // shadow slots per hit
};

// Where a frame's buffers live when they are not in the context's arena: the fixed
// device array of the no-allocation entry point, or an arena a caller owns.
struct FixedScratch {
    char* base;
    size_t bytes;
};

// Deepest max_depth a frame in fixed buffers takes (its per-level counts are copied back
// through fixed host arrays).
constexpr int kNoAllocMaxDepth = 64;

// A frame whose launches are enqueued and whose stats are read once its stream is done.
struct SceneDevice;
struct PendingFrame {
    FrameView F;
    SceneView S;
    long long n_paths;
    int depth;
    bool timed_stages;
    bool fixed;
    const SceneDevice* D;
};

CameraDev make_camera(const rt_camera_desc& c, double aspect);
// cam_pos: position of the camera `cam` was built from (nullptr = the scene's camera)
SceneView make_view(const SceneDevice& D, const CameraDev& cam, int light_samples, const double* cam_pos = nullptr);
int grid_cols(int n);
// Scalar fields of the frame view for options `o` with raygen tiles of `tile` pixels;
// validates the options and returns the queue sizes.
FrameLayout frame_fill(FrameView& F, const SceneView& S, const rt_render_opts* o, int tile);
// Places the frame's buffers in one allocation at `base` (nullptr: only measures); returns bytes.
size_t frame_carve(FrameView& F, const SceneView& S, const FrameLayout& L, char* base, RayRec** q0, RayRec** q1,
                   bool own_rgb);

// Stats of a launched frame; waits for ctx's stream (render.cpp).
void frame_stats(rt_ctx* ctx, const PendingFrame& pf, rt_render_stats* stats, bool record_primary);
// rt_render_device_aux enqueued without waiting: frame buffers in `arena` (nullable = the
// context's arena); the stats come from frame_stats(*pending) after the stream is done.
// on_stream: launch on that stream instead of the context's (pending frames with their own
// buffers only; the caller orders it against the context's stream with events).
void ctx_render_launch(rt_ctx* ctx, rt_scene* scene, const rt_render_opts* o, void* dev_rgb, const rt_render_aux* aux,
                       const FixedScratch* arena, PendingFrame* pending, hipStream_t on_stream = nullptr);
// Device bytes of a frame's buffers for these options.
size_t ctx_frame_bytes(rt_scene* scene, const rt_render_opts* o);

}  // namespace rt

struct rt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    rt::FrameBuffers fb;
    rt::StageTimer timer;
    int last_paths = 0;
    const int32_t* last_prim_obj = nullptr;  // primary-hit records of the last render (in fb.arena)
    const int32_t* last_prim_tri = nullptr;
    const double* last_prim_t = nullptr;
    char* noalloc_scratch = nullptr;  // this device's address of the no-allocation scratch array
    // the last frame's launches as a hipGraph, replayed while its views and buffers
    // (the key: SceneView, FrameView and the two ray queues, byte for byte) are the same
    hipGraph_t frame_graph = nullptr;
    hipGraphExec_t frame_exec = nullptr;
    std::vector<char> frame_key;
    std::vector<char> seen_key;  // the previous frame's key: a frame is captured the second time in a row
    void drop_frame_graph() {
        if (frame_exec) (void)hipGraphExecDestroy(frame_exec);
        if (frame_graph) (void)hipGraphDestroy(frame_graph);
        frame_exec = nullptr;
        frame_graph = nullptr;
        frame_key.clear();
    }
};

struct rt_scene {
    std::unique_ptr<rt::SceneDevice> dev;
    rt_ctx* ctx = nullptr;
};
