// batch.cpp — batched rendering: many (scene, camera, resolution) frames in one captured
// launch sequence of persistent kernels (kernels/batch.hip), and its autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../kernels/batch.hpp"
#include "context.hpp"
#include "frame.hpp"

struct rt_batch {
    rt_ctx* ctx = nullptr;
    int n_jobs = 0;
    rt_batch_opts bo{};
    rt::DevBuf arena, views, frames, plan, tally;
    std::vector<int> group_start;        // jobs [group_start[g], group_start[g + 1]) form group g
    std::vector<int> group_depth;        // deepest job of each group
    // Each group runs on one of bo.concurrent branches (group_branch); a branch runs its groups
    // one after another on its stream (branch 0: the context's stream), in its own part of the
    // arena, with its own plan and chunk counter; the others fork from and join the context's
    // stream by events.
    std::vector<int> group_branch;
    std::vector<hipStream_t> side;       // streams of branches 1..concurrent-1
    std::vector<hipEvent_t> fork, join;  // one each per side stream
    size_t plan_stride = 0;              // int32 entries between the branches' plan buffers
    size_t plan_counters = 0;            // offset of the slice counters in a plan buffer
    std::vector<rt::FrameView> hframes;  // [2 * n_jobs]
    std::vector<rt::SceneView> hviews;
    std::vector<rt::FrameLayout> layouts;
    rt::BatchArgs args{};
    rt::BatchLaunch launch{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    size_t graph_nodes = 0;
    double ms_build = 0;
    uint64_t pixels = 0;
    ~rt_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (hipStream_t s : side) (void)hipStreamDestroy(s);
        for (hipEvent_t e : fork) (void)hipEventDestroy(e);
        for (hipEvent_t e : join) (void)hipEventDestroy(e);
    }
};

namespace rt {
namespace {

int compute_units(int device) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0) cus = 256;
    return cus;
}

void check_opts(const rt_batch_opts& bo) {
    if (bo.tile != 4 && bo.tile != 8 && bo.tile != 16) throw Error(RT_E_INVALID_ARG, "tile must be 4, 8 or 16");
    if (bo.chunk < kBatchBlock || bo.chunk % kBatchBlock != 0 || bo.chunk > (1 << 20))
        throw Error(RT_E_INVALID_ARG, "chunk must be a positive multiple of 64");
    if (bo.blocks_per_cu < 1 || bo.blocks_per_cu > 64) throw Error(RT_E_INVALID_ARG, "blocks_per_cu out of range");
    if (!(bo.group_gib >= 0.0)) throw Error(RT_E_INVALID_ARG, "group_gib must be >= 0");
    if (bo.concurrent < 1 || bo.concurrent > 8) throw Error(RT_E_INVALID_ARG, "concurrent must be 1 to 8");
}

// Every group's launch sequence on its branch, the groups of a branch one after another (they
// share that branch's frame buffers).
hipError_t launch_groups(rt_batch* b, hipStream_t st) {
    const size_t C = b->side.size() + 1;
    hipError_t e = hipSuccess;
    for (size_t c = 1; c < C && e == hipSuccess; ++c) {
        e = hipEventRecord(b->fork[c - 1], st);
        if (e == hipSuccess) e = hipStreamWaitEvent(b->side[c - 1], b->fork[c - 1], 0);
    }
    for (size_t g = 0; g + 1 < b->group_start.size() && e == hipSuccess; ++g) {
        const size_t c = (size_t)b->group_branch[g];
        const int j0 = b->group_start[g], j1 = b->group_start[g + 1];
        BatchArgs a = b->args;
        a.views += j0;
        a.frames += 2 * (size_t)j0;
        a.tally += 2 * (size_t)j0;
        a.base += c * b->plan_stride;
        a.counter = a.base + b->plan_counters;
        a.n_jobs = j1 - j0;
        BatchLaunch l = b->launch;
        l.max_depth = b->group_depth[g];
        e = launch_batch(a, l, c == 0 ? st : b->side[c - 1]);
    }
    // the side branches join the context's stream (also after an error, so a capture ends whole)
    for (size_t c = 1; c < C; ++c) {
        hipError_t e2 = hipEventRecord(b->join[c - 1], b->side[c - 1]);
        if (e2 == hipSuccess) e2 = hipStreamWaitEvent(st, b->join[c - 1], 0);
        if (e == hipSuccess) e = e2;
    }
    return e;
}

std::unique_ptr<rt_batch> build(rt_ctx* ctx, const rt_batch_job* jobs, int n, const rt_batch_opts& bo) {
    auto t0 = std::chrono::steady_clock::now();
    if (n <= 0 || !jobs) throw Error(RT_E_INVALID_ARG, "a batch needs at least one job");
    check_opts(bo);
    auto b = std::make_unique<rt_batch>();
    b->ctx = ctx;
    b->n_jobs = n;
    b->bo = bo;
    b->hviews.resize(n);
    b->hframes.resize(2 * (size_t)n);
    b->layouts.resize(n);
    std::vector<size_t> need(n);
    for (int j = 0; j < n; ++j) {
        const rt_batch_job& jb = jobs[j];
        if (!jb.scene || !jb.scene->dev || !jb.d_rgb_out) throw Error(RT_E_INVALID_ARG, "job without scene or output");
        if (jb.scene->ctx != ctx) throw Error(RT_E_INVALID_ARG, "job scene belongs to another context");
        const rt_render_opts& o = jb.opts;
        if (o.max_depth < 1) throw Error(RT_E_INVALID_ARG, "batched jobs need max_depth >= 1");
        if (o.outline_enabled && o.aa_mode != RT_AA_QUINCUNX)
            throw Error(RT_E_INVALID_ARG, "batched jobs do not take the outline pass");
        if (o.count_work || o.profile) throw Error(RT_E_INVALID_ARG, "count_work and profile are single-frame options");
        if (o.light_samples < 0 || o.light_samples > 65536) throw Error(RT_E_INVALID_ARG, "light_samples out of range");
        if (o.width <= 0 || o.height <= 0) throw Error(RT_E_INVALID_ARG, "invalid image size or samples");
        const SceneDevice& D = *jb.scene->dev;
        const rt_camera_desc& cd = jb.camera ? *jb.camera : D.host.camera;
        CameraDev cam = make_camera(cd, (double)o.width / (double)o.height);
        b->hviews[j] = make_view(D, cam, o.light_samples, cd.position);
        FrameView& F = b->hframes[2 * j];
        F = FrameView{};
        b->layouts[j] = frame_fill(F, b->hviews[j], &o, bo.tile);
        RayRec *q0 = nullptr, *q1 = nullptr;
        need[j] = frame_carve(F, b->hviews[j], b->layouts[j], nullptr, &q0, &q1, false);
        b->pixels += (uint64_t)o.width * o.height;
    }
    // Consecutive jobs form groups whose frame buffers fit group_gib (a job larger than that is
    // a group of its own). The groups run one after another in the same launch sequence and
    // share one arena, so the device holds the largest group's buffers, not every job's.
    // Consecutive jobs are usually views of the same scene, so a group's meshes stay in L2
    // (groups balanced across scenes measured slower: 1082 against 1211 Mrays/s at 16 GiB).
    const double budget = bo.group_gib > 0.0 ? bo.group_gib * 1073741824.0 : 1e300;
    std::vector<int> order(n);  // slot -> job: slots follow the job order
    std::vector<size_t> offs(n);
    std::vector<int> group_of(n);
    size_t slot = 0, in_group = 0;  // slot: the largest group's buffers, 256-byte aligned
    b->group_start.assign(1, 0);
    b->group_depth.assign(1, 0);
    for (int j = 0; j < n; ++j) {
        order[j] = j;
        if (in_group > 0 && (double)(in_group + need[j]) > budget) {
            b->group_start.push_back(j);
            b->group_depth.push_back(0);
            in_group = 0;
        }
        group_of[j] = (int)b->group_start.size() - 1;
        offs[j] = in_group;
        in_group += need[j];
        slot = std::max(slot, (in_group + 255) & ~(size_t)255);
        b->group_depth.back() = std::max(b->group_depth.back(), b->layouts[j].depth);
    }
    b->group_start.push_back(n);
    const int n_groups = (int)b->group_start.size() - 1;
    const int C = std::min(bo.concurrent, n_groups);  // branches with a group to run
    // each group, in order, to the branch with the least frame-buffer bytes so far (the bytes
    // follow the rays: a proxy for the work), so the branches end close together
    std::vector<double> load(C, 0.0);
    b->group_branch.assign(n_groups, 0);
    for (int g = 0; g < n_groups; ++g) {
        double bytes = 0;
        for (int j = b->group_start[g]; j < b->group_start[g + 1]; ++j) bytes += (double)need[j];
        const int c = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        b->group_branch[g] = c;
        load[c] += bytes;
    }
    const size_t arena = slot * (size_t)C;
    if ((double)arena > bo.max_gib * 1073741824.0)
        throw Error(RT_E_OOM, "batch needs " + std::to_string(arena >> 20) + " MiB of frame buffers, over max_gib");
    b->arena.reserve(arena);
    for (int c = 1; c < C; ++c) {
        hipStream_t s = nullptr;
        HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        b->side.push_back(s);
        hipEvent_t e = nullptr;
        HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        b->fork.push_back(e);
        HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        b->join.push_back(e);
    }
    for (int j = 0; j < n; ++j) {
        FrameView& F = b->hframes[2 * j];
        RayRec *q0 = nullptr, *q1 = nullptr;
        const size_t base = slot * (size_t)b->group_branch[group_of[j]];
        frame_carve(F, b->hviews[j], b->layouts[j], b->arena.as<char>() + base + offs[j], &q0, &q1, false);
        F.rgb = static_cast<uint8_t*>(jobs[j].d_rgb_out);
        F.rays = q0;
        F.rays_next = q1;
        FrameView& G2 = b->hframes[2 * j + 1];
        G2 = F;
        G2.rays = q1;
        G2.rays_next = q0;
    }
    // the device arrays in slot order
    std::vector<SceneView> dviews(n);
    std::vector<FrameView> dframes(2 * (size_t)n);
    for (int k = 0; k < n; ++k) {
        dviews[k] = b->hviews[order[k]];
        dframes[2 * (size_t)k] = b->hframes[2 * (size_t)order[k]];
        dframes[2 * (size_t)k + 1] = b->hframes[2 * (size_t)order[k] + 1];
    }
    hipStream_t st = ctx_stream(ctx);
    b->views.reserve(n * sizeof(SceneView));
    b->frames.reserve(2 * (size_t)n * sizeof(FrameView));
    // each branch: base[n + 1], then its slices' counters on 128-byte lines of their own
    b->plan_counters = ((size_t)n + 1 + 31) & ~(size_t)31;
    b->plan_stride = b->plan_counters + 32 * (size_t)kBatchSlices;
    b->plan.reserve(b->plan_stride * (size_t)C * sizeof(int32_t));
    b->tally.reserve(2 * (size_t)n * sizeof(unsigned long long));
    HIP_TRY(hipMemcpyAsync(b->views.p, dviews.data(), n * sizeof(SceneView), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(b->frames.p, dframes.data(), 2 * (size_t)n * sizeof(FrameView), hipMemcpyHostToDevice, st));
    b->args.views = b->views.as<SceneView>();
    b->args.frames = b->frames.as<FrameView>();
    b->args.base = b->plan.as<int32_t>();
    b->args.counter = b->plan.as<int32_t>() + b->plan_counters;
    b->args.n_jobs = n;
    b->args.chunk = bo.chunk;
    b->args.tally = b->tally.as<unsigned long long>();
    b->launch.grid = compute_units(ctx_device(ctx)) * bo.blocks_per_cu;
    HIP_TRY(hipStreamSynchronize(st));
    if (bo.use_graph) {
        HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipError_t le = launch_groups(b.get(), st);
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(st, &g);
        b->graph = g;
        HIP_TRY(le);
        HIP_TRY(ce);
        HIP_TRY(hipGraphGetNodes(b->graph, nullptr, &b->graph_nodes));
        HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    b->ms_build = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return b;
}

void run(rt_batch* b, rt_batch_stats* stats) {
    rt_ctx* ctx = b->ctx;
    hipStream_t st = ctx_stream(ctx);
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    struct Ev {
        hipEvent_t a, b;
        ~Ev() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } ev{e0, e1};
    HIP_TRY(hipEventRecord(e0, st));
    if (b->exec) HIP_TRY(hipGraphLaunch(b->exec, st));
    else HIP_TRY(launch_groups(b, st));
    HIP_TRY(hipEventRecord(e1, st));
    HIP_TRY(hipEventSynchronize(e1));
    if (!stats) return;
    memset(stats, 0, sizeof(*stats));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    stats->ms = ms;
    stats->ms_build = b->ms_build;
    stats->n_jobs = b->n_jobs;
    stats->graph_nodes = (int32_t)b->graph_nodes;
    stats->bytes = (int64_t)b->arena.bytes;
    stats->pixels = b->pixels;
    stats->groups = (int32_t)(b->group_start.size() - 1);
    std::vector<unsigned long long> tally(2 * (size_t)b->n_jobs);
    HIP_TRY(hipMemcpy(tally.data(), b->tally.p, tally.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    for (int j = 0; j < b->n_jobs; ++j) {
        stats->rays_primary += (uint64_t)b->layouts[j].n_paths;
        stats->rays_reflect += tally[2 * (size_t)j];
        stats->rays_shadow += tally[2 * (size_t)j + 1];
    }
}

template <class F>
int guard(rt_ctx* ctx, F fn) {
    try {
        fn();
        return RT_OK;
    } catch (const Error& e) {
        return ctx_fail(ctx, e.code, e.what());
    } catch (const std::bad_alloc&) {
        return ctx_fail(ctx, RT_E_OOM, "host allocation failed");
    } catch (const std::exception& e) {
        return ctx_fail(ctx, RT_E_INTERNAL, e.what());
    }
}

}  // namespace
}  // namespace rt

using namespace rt;

extern "C" {

void rt_batch_opts_default(rt_batch_opts* bo) {
    if (!bo) return;
    bo->tile = 8;
    bo->chunk = 256;
    bo->blocks_per_cu = 16;
    bo->use_graph = 1;
    bo->max_gib = 200.0;
    bo->group_gib = 8.0;
    bo->concurrent = 4;
}

int rt_batch_create(rt_ctx* ctx, const rt_batch_job* jobs, int32_t n_jobs, const rt_batch_opts* bo, rt_batch** out) {
    if (!ctx || !out) return ctx_fail(ctx, RT_E_INVALID_ARG, "null argument");
    *out = nullptr;
    rt_batch_opts o;
    rt_batch_opts_default(&o);
    if (bo) o = *bo;
    return guard(ctx, [&] { *out = build(ctx, jobs, n_jobs, o).release(); });
}

int rt_batch_run(rt_batch* b, rt_batch_stats* stats) {
    if (!b) return ctx_fail(nullptr, RT_E_INVALID_ARG, "null batch");
    return guard(b->ctx, [&] { run(b, stats); });
}

void rt_batch_destroy(rt_batch* b) {
    if (!b) return;
    (void)hipStreamSynchronize(ctx_stream(b->ctx));
    delete b;
}

int rt_batch_autotune(rt_ctx* ctx, const rt_batch_job* jobs, int32_t n_jobs, rt_batch_opts* bo, double* best_ms) {
    if (!ctx || !bo) return ctx_fail(ctx, RT_E_INVALID_ARG, "null argument");
    return guard(ctx, [&] {
        rt_batch_opts cur = *bo;
        check_opts(cur);
        // one timed replay after one untimed, best of two
        auto measure = [&](const rt_batch_opts& o) {
            std::unique_ptr<rt_batch> b = build(ctx, jobs, n_jobs, o);
            rt_batch_stats s{};
            run(b.get(), nullptr);
            double best = 1e300;
            for (int r = 0; r < 2; ++r) {
                run(b.get(), &s);
                best = std::min(best, s.ms);
            }
            return best;
        };
        double best = measure(cur);
        const int chunks[] = {64, 128, 256, 512, 1024, 2048};
        const int blocks[] = {4, 8, 16, 32};
        const int tiles[] = {4, 8, 16};
        for (int c : chunks) {
            if (c == cur.chunk) continue;
            rt_batch_opts t = cur;
            t.chunk = c;
            double ms = measure(t);
            if (ms < best) {
                best = ms;
                cur = t;
            }
        }
        for (int k : blocks) {
            if (k == cur.blocks_per_cu) continue;
            rt_batch_opts t = cur;
            t.blocks_per_cu = k;
            double ms = measure(t);
            if (ms < best) {
                best = ms;
                cur = t;
            }
        }
        for (int tl : tiles) {
            if (tl == cur.tile) continue;
            rt_batch_opts t = cur;
            t.tile = tl;
            double ms = measure(t);
            if (ms < best) {
                best = ms;
                cur = t;
            }
        }
        *bo = cur;
        if (best_ms) *best_ms = best;
    });
}

int rt_scene_auto_cameras(rt_scene* scene, rt_camera_desc out[4]) {
    if (!scene || !scene->dev || !out) return ctx_fail(scene ? scene->ctx : nullptr, RT_E_INVALID_ARG, "null argument");
    return guard(scene->ctx, [&] { auto_cameras(scene->dev->host, out); });
}

int rt_mesh_load_stl(const char* path, double* vertices, int64_t capacity_tris, int64_t* n_tris) {
    if (!path || !n_tris) return ctx_fail(nullptr, RT_E_INVALID_ARG, "null argument");
    return guard(nullptr, [&] {
        std::vector<double> v = load_stl_file(path);
        *n_tris = (int64_t)(v.size() / 9);
        if (vertices) std::copy(v.begin(), v.begin() + std::min<size_t>(v.size(), (size_t)capacity_tris * 9), vertices);
    });
}

int rt_auto_cameras_file(const char* scene_path, rt_camera_desc out[4]) {
    if (!scene_path || !out) return ctx_fail(nullptr, RT_E_INVALID_ARG, "null argument");
    return guard(nullptr, [&] { auto_cameras(load_scene_file(scene_path), out); });
}

}  // extern "C"
