// context.hpp — what host modules other than render.cpp may use of a context.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../../include/rtrace_hip.h"

namespace rt {
hipStream_t ctx_stream(rt_ctx* ctx);
int ctx_device(rt_ctx* ctx);
rt_ctx* scene_ctx(rt_scene* scene);  // the context a scene was loaded on
int ctx_fail(rt_ctx* ctx, int code, const std::string& msg);  // records the message, returns code
// rt_render_device_aux without the C-ABI guard (throws rt::Error)
void ctx_render_aux(rt_ctx* ctx, rt_scene* scene, const rt_render_opts* o, void* dev_rgb, const rt_render_aux* aux,
                    rt_render_stats* stats);
}  // namespace rt
