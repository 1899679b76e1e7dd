"""Benchmark: Mrays/s of the per-pixel trace path at 1080p on one MI355X, and the
denoised-vs-plain frame time of the 1080p denoiser workload.

Workload (default `icosphere_1080p`): the procedural mesh scene
(81,920-triangle noisy icosphere with a host-built k-d tree, 3 transformed cubes, a
reflective grid plane, a disk area light with 16 shadow samples per hit, fog 0.05,
reflections to depth 4) rendered at 1920x1080 with quincunx anti-aliasing. A step is one
full frame (camera rays, closest-hit, shading, shadow rays, reflections, resolve,
quantise) rendered into a device buffer; the scene is resident in HBM before timing.

Rays = primary + reflection + shadow rays, counted exactly by the pipeline.
The "denoise" object: denoiser-workload frames at 1080p (icosphere scenes, 2x2 square light) rendered plain with 256 light
samples (the quality reference) and denoised (4 samples + the trained U-Net + tile fallback), with
PSNR / SSIM per image and the denoiser's BF16 MFMA roofline.
Prints one JSON line.
"""
import argparse
import json
import os
import re
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per unit of traversal work (layout in rtrace_amd/csrc/common/layout.hpp)
B_NODE, B_TRI, B_RETEST = 40, 48, 72  # node: 32 B box+links, +8 B f64 split (inner)
B_BVH_NODE = 128  # 4-wide mesh BVH node: four slot boxes and links
B_SPHERE, B_PLANE, B_CUBE = 32, 48, 304
B_RAY_CLOSEST = 64 + 16  # RayRec in, HitRec out
B_RAY_SHADOW = 48 + 24 + 1  # hit point + normal, light sample point in, visibility byte out


def kernel_bytes(work, rays, per_ray):
    return (work["nodes"] * B_NODE + work["bvh_nodes"] * B_BVH_NODE + work["tris"] * B_TRI + work["retests"] * B_RETEST +
            work["spheres"] * B_SPHERE + work["planes"] * B_PLANE + work["cubes"] * B_CUBE + rays * per_ray)


def make_workload(name, out_dir):
    from rtrace_amd import scenes
    if name.startswith("icosphere_square"):
        path = scenes.icosphere_scene(out_dir, subdiv=6, light="square", light_size=1.0, light_samples=16)
    elif name.startswith("icosphere"):
        path = scenes.icosphere_scene(out_dir, subdiv=6)
    elif name.startswith("spheres_ring"):
        path = scenes.spheres_ring(out_dir, 1000)
    elif name.startswith("sphere_800x600"):
        path = scenes.sphere_plane(out_dir)
    else:
        raise SystemExit(f"unknown workload {name}")
    return path, scenes.RENDER[name]


def configure(R, renderer, cfg, scene):
    renderer.max_depth = cfg["max_depth"]
    renderer.samples = cfg["samples"]
    renderer.anti_aliasing_mode = {"quincunx": R.AntiAliasingMode.Quincunx, "stochastic": R.AntiAliasingMode.Stochastic,
                                   "no-jitter": R.AntiAliasingMode.NoJitter}[cfg["aa"]]
    if cfg["aa"] != "quincunx":
        renderer.outline_config = scene.info()["outline"]


def cpu_baseline(path, cfg, budget_s=20.0):
    """The CPU f64 oracle (a restatement of the reference renderer) on the box's host
    cores, on a reduced-resolution render of the same scene and settings."""
    from oracle import pyoracle as O
    threads = min(16, os.cpu_count() or 1)
    w, h = 64, 36
    with open(path) as f:
        scene = json.load(f)
    outline = O.outline_from_scene(scene) if cfg["aa"] != "quincunx" else None
    spent = 0.0
    best = None
    while True:
        opts = O.make_opts(w, h, max_depth=cfg["max_depth"], samples=cfg["samples"], aa=cfg["aa"], threads=threads,
                           outline=outline)
        _, st, _ = O.render(path, opts)
        rays = st["rays_primary"] + st["rays_reflect"] + st["rays_shadow"]
        best = (rays / st["render_seconds"] / 1e6, w, h, st["render_seconds"], rays,
                {"kd_nodes": (st["kd_inner_visits"] + st["kd_leaf_visits"]) / rays, "tri_tests": st["tri_tests"] / rays})
        spent += st["render_seconds"]
        if st["render_seconds"] * 4 > budget_s - spent or w >= cfg["width"]:
            break
        w, h = w * 2, h * 2
    # one thread (the scalar baseline) on the smallest sample
    opts = O.make_opts(64, 36, max_depth=cfg["max_depth"], samples=cfg["samples"], aa=cfg["aa"], threads=1,
                       outline=outline)
    _, st1, _ = O.render(path, opts)
    rays1 = st1["rays_primary"] + st1["rays_reflect"] + st1["rays_shadow"]
    return {"value": best[0], "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"oracle render of the same scene at {best[1]}x{best[2]} ({best[4]} rays, {best[3]:.1f} s, "
                      f"{threads} threads)",
            # the reference's traversal (no early exit) per ray on this sample: an upper bound of
            # the traversal work, never the roofline's numerator
            "reference_traversal_per_ray": {k: round(v, 1) for k, v in best[5].items()},
            "one_thread": {"value": rays1 / st1["render_seconds"] / 1e6, "unit": "Mrays/s", "cores": 1,
                           "sample": f"64x36 ({rays1} rays, {st1['render_seconds']:.1f} s)"}}


COPY_BYTES = 1 << 30

BF16_PEAK_TFLOPS = 2500.0  # MI355X dense bf16 MFMA
# Denoiser multiply-accumulates per full-resolution pixel: sum over layers of
# 9 * cin * cout / 4^level (13 real input features, 3 outputs; layer table in
# include/rtrace_denoise.h)
DN_LAYERS = [(13, 32, 0), (32, 32, 0), (32, 64, 1), (64, 64, 1), (64, 128, 2), (128, 128, 2), (128, 128, 3),
             (128, 128, 3), (256, 128, 2), (128, 128, 2), (192, 64, 1), (64, 64, 1), (96, 32, 0), (32, 32, 0),
             (32, 3, 0)]
DN_MAC_PER_PX = sum(9 * ci * co / 4 ** lv for ci, co, lv in DN_LAYERS)


def denoise_section(seeds, threshold, low=4, high=256):
    """The denoiser workload at 1080p: the plain render with `high` light samples (the quality reference) against
    the denoised render (`low` samples + network + tile fallback), per image."""
    import rtrace_amd as R  # noqa: F401 - loads the library
    from rtrace_amd import denoise
    dn = denoise.Denoiser.load()
    rows = denoise.evaluate(dn, seeds, 1920, 1080, low=low, high=high, threshold=threshold, repeats=3)
    ms_dn_net = sorted(r["stage_ms"]["denoise"] for r in rows)[len(rows) // 2]
    flops = 2.0 * DN_MAC_PER_PX * 1920 * 1080
    ach = flops / (ms_dn_net * 1e-3) / 1e12
    plain = sum(r["ms_plain_high"] for r in rows) / len(rows)
    den = sum(r["ms_denoised"] for r in rows) / len(rows)
    return {
        "workload": "denoise_1080p (icosphere scenes, 2x2 square light), 1920x1080 quincunx, depth 4",
        "low_samples": low, "high_samples": high, "threshold": threshold, "tile": 16,
        "ms_per_frame_plain": round(plain, 3), "ms_per_frame_denoised": round(den, 3),
        "speedup": round(plain / den, 2),
        "psnr_min": min(r["psnr_denoised"] for r in rows), "psnr_low_min": min(r["psnr_low"] for r in rows),
        "per_image": rows,
        "roofline": {"bound": "mfma", "kernel": "denoiser (prep, 15 conv3x3, pools, compose, tile test)",
                     "achieved": round(ach, 1), "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(ach / BF16_PEAK_TFLOPS, 4), "flop_per_frame": flops,
                     "ms_per_frame": ms_dn_net},
    }


def traffic_child(workload):
    """Profiled child: one frame of the workload, then a plain copy of COPY_BYTES (the
    calibration kernel for the memory counters), in one process."""
    import torch

    import rtrace_amd as R
    path, cfg = make_workload(workload, os.path.join(tempfile.gettempdir(), "rtrace_amd_bench"))
    scene = R.Scene.from_json_file(path)
    r = R.Renderer(cfg["width"], cfg["height"])
    configure(R, r, cfg, scene)
    out = torch.empty(cfg["width"] * cfg["height"] * 3, dtype=torch.uint8, device="cuda")
    r.render_device(scene, out)
    src = torch.ones(COPY_BYTES, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()


def measure_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from two rocprofv3 --pmc passes (FETCH_SIZE, then
    WRITE_SIZE; KiB per dispatch), each scaled by bytes / counter of the copy kernel."""
    import csv
    import glob
    import shutil
    import subprocess

    prof = shutil.which("rocprofv3")
    if not prof:
        return None, "rocprofv3 not found"
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE", "TCC_REQ_sum TCC_HIT_sum"):
        d = tempfile.mkdtemp(prefix="rt_pmc_")
        cmd = [prof, "--pmc", *counter.split(), "--output-format", "csv", "-d", d, "-o", "run", "--", sys.executable,
               os.path.abspath(__file__), "--workload", workload, "--traffic-child"]
        try:
            subprocess.run(cmd, check=True, timeout=300, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        except Exception as e:  # noqa: BLE001 - report and carry on without traffic
            return None, f"{counter} pass failed: {e}"
        rows = []
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            rows += list(csv.DictReader(open(f)))
        if counter.startswith("TCC_"):
            # L2 requests and hits per launch; bytes per request from the copy (each copied
            # byte is one L2 read and one L2 write)
            def per_dispatch(name, match):
                vals = {}
                for r in rows:
                    if r["Counter_Name"] == name and match(r["Kernel_Name"]):
                        vals[r.get("Dispatch_Id", len(vals))] = vals.get(r.get("Dispatch_Id", len(vals)), 0.0) + float(r["Counter_Value"])
                return list(vals.values())
            kreq = per_dispatch("TCC_REQ_sum", lambda n: re.search(r"::%s[<(]" % kernel, n))
            khit = per_dispatch("TCC_HIT_sum", lambda n: re.search(r"::%s[<(]" % kernel, n))
            creq = [c for c in per_dispatch("TCC_REQ_sum", lambda n: "copy" in n.lower()) if c > 1000]
            shutil.rmtree(d, ignore_errors=True)
            if kreq and creq:
                bpr = 2.0 * COPY_BYTES / max(creq)
                req = sum(kreq) / len(kreq)
                res["l2"] = {"requests": int(req), "bytes": int(req * bpr), "bytes_per_request": round(bpr, 1),
                             "hit_rate": round(sum(khit) / max(sum(kreq), 1.0), 4) if khit else None}
            continue
        kern = [float(r["Counter_Value"]) for r in rows if re.search(r"::%s[<(]" % kernel, r["Kernel_Name"])]
        copy = [float(r["Counter_Value"]) for r in rows if "copyBuffer" in r["Kernel_Name"] or "copy" in r["Kernel_Name"].lower()]
        copy = [c for c in copy if c > 1000]  # the 256 MiB copy, not small staging copies
        shutil.rmtree(d, ignore_errors=True)
        if not kern or not copy:
            return None, f"{counter}: no samples"
        scale = COPY_BYTES / (max(copy) * 1024.0)
        res[counter] = (sum(kern) / len(kern)) * 1024.0 * scale
        res[counter + "_scale"] = scale
    return res, None


def _progress(msg):
    """A progress line on stderr (the JSON result line stays alone on stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def bench_batch(args):
    """Batch workload: 64 seeded scenes x 4 auto-camera views at their own resolutions,
    rendered as one batch (persistent kernels, one captured graph) with autotuned tile,
    chunk and occupancy. A step is one replay of the whole batch. Also times the same jobs
    rendered one by one (rt_render_device), the unbatched path."""
    import torch

    import rtrace_amd as R
    from rtrace_amd import scenes
    out_dir = os.path.join(tempfile.gettempdir(), "rtrace_amd_bench_c4")
    t0 = time.perf_counter()
    work = scenes.batch_scenes(out_dir, n_scenes=args.c4_scenes)
    jobs, tris = [], []
    for path, w, h, n_tri in work:
        scene = R.Scene.from_json_file(path)
        for cam in R.AutoCamera.generate_cameras(scene).cameras():
            r = R.Renderer(w, h)
            r.max_depth = 4
            jobs.append((scene, cam, r, torch.empty(h * w * 3, dtype=torch.uint8, device="cuda")))
            tris.append(n_tri)
        if len(tris) % 32 == 0:
            _progress(f"loaded {len(tris) // 4} scenes ({time.perf_counter() - t0:.1f} s)")
    t_load = time.perf_counter() - t0
    tune_jobs = jobs[::max(1, len(jobs) // 32)]
    opts, tune_ms = R.autotune_batch(tune_jobs)
    _progress(f"autotuned on {len(tune_jobs)} jobs: {opts.as_dict()}")
    b = R.Batch(jobs, opts)
    for _ in range(args.warmup):
        b.run()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_dev = []
    for _ in range(args.steps):
        st = b.run()
        ms_dev.append(st.ms)
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t1) * 1e3 / args.steps
    _progress(f"batch: {ms_step:.1f} ms per step")
    # the same frames one at a time
    seq_ms = None
    if not args.no_sequential:
        cam_scenes = {}
        singles = []
        for scene, cam, r, out in jobs:
            key = (scene._path, tuple(cam.position))
            if key not in cam_scenes:
                sc = json.load(open(scene._path))
                sc["camera"] = {k: v for k, v in cam.to_json().items() if v is not None}
                cam_scenes[key] = R.Scene.from_json_str(json.dumps(sc), os.path.dirname(scene._path))
            singles.append((cam_scenes[key], r, out))
        for s_, r, out in singles:  # warm
            r.render_device(s_, out)
        torch.cuda.synchronize()
        _progress(f"sequential warm-up done ({len(singles)} frames)")
        t2 = time.perf_counter()
        for s_, r, out in singles:
            r.render_device(s_, out)
        torch.cuda.synchronize()
        seq_ms = (time.perf_counter() - t2) * 1e3
    rays = st.rays_total
    line = {
        "metric": "Mrays/s, batched (64 scenes x 4 views)", "value": round(rays / (ms_step * 1e-3) / 1e6, 2), "unit": "Mrays/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32+f64",
        "data": "synthetic (seeded procedural scenes)",
        "config": {"workload": f"batch64: {len(work)} scenes x 4 auto-camera views, quincunx, depth 4, disk light "
                               f"8 samples", "jobs": len(jobs), "pixels": st.pixels,
                   "triangles_min": min(tris), "triangles_max": max(tris), "rays_per_step": rays,
                   "primary": st.rays_primary, "reflect": st.rays_reflect, "shadow": st.rays_shadow},
        "batch": {"opts": opts.as_dict(), "tune_ms_subset": round(tune_ms, 3), "tune_jobs": len(tune_jobs),
                  "graph_nodes": st.graph_nodes, "graph_launches_per_step": 1, "device_ms": round(min(ms_dev), 3),
                  "frame_buffer_gib": round(st.bytes / 2**30, 2), "build_ms": round(st.ms_build, 1),
                  "scene_load_s": round(t_load, 1)},
    }
    if seq_ms is not None:
        line["sequential"] = {"ms": round(seq_ms, 3), "mrays_s": round(rays / (seq_ms * 1e-3) / 1e6, 2),
                              "batch_speedup": round(seq_ms / ms_step, 3)}
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="icosphere_1080p")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 FETCH_SIZE/WRITE_SIZE passes")
    ap.add_argument("--traffic-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--no-denoise", action="store_true", help="skip the denoised-vs-plain section")
    ap.add_argument("--no-stage-events", action="store_true",
                    help="time the steps without per-stage hipEvents (stage times from as many profiled steps after)")
    ap.add_argument("--dn-seeds", default="100001,100002", help="denoiser scene seeds of the denoise section")
    ap.add_argument("--dn-threshold", type=float, default=0.06, help="tile fallback threshold")
    ap.add_argument("--c4-scenes", type=int, default=64, help="batch64: number of scenes")
    ap.add_argument("--no-sequential", action="store_true", help="batch64: skip the one-by-one comparison")
    args = ap.parse_args()
    if args.traffic_child:
        traffic_child(args.workload)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the trace path has no multi-GPU mode in this build (asked for {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch64":
        bench_batch(args)
        return

    import torch

    import rtrace_amd as R

    out_dir = os.path.join(tempfile.gettempdir(), "rtrace_amd_bench")
    path, cfg = make_workload(args.workload, out_dir)
    scene = R.Scene.from_json_file(path)
    W, H = cfg["width"], cfg["height"]
    r = R.Renderer(W, H)
    configure(R, r, cfg, scene)
    out = torch.empty(W * H * 3, dtype=torch.uint8, device="cuda")

    # untimed passes: exact work counters, then per-stage hipEvent timing
    r.count_work = True
    st_work = r.render_device(scene, out)
    r.count_work = False

    # By default the timed steps record hipEvents around every stage's launches on the
    # library's stream (profile), so the per-launch kernel times come from the timed region
    # itself. With --no-stage-events the timed steps replay the frame's hipGraph without
    # events (small frames, where ~6 event pairs per level weigh) and the stage times come
    # from as many profiled steps after them.
    stage_ms, stage_n = {}, {}

    def add_stages(st_):
        d_ = st_.as_dict()
        for k, v in d_["ms_stage"].items():
            stage_ms[k] = stage_ms.get(k, 0.0) + v
            stage_n[k] = stage_n.get(k, 0) + d_["launches_stage"][k]

    r.profile = not args.no_stage_events
    for _ in range(args.warmup):
        r.render_device(scene, out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev_ms = []
    for _ in range(args.steps):
        st = r.render_device(scene, out)
        dev_ms.append(st.ms_render)
        if r.profile:
            add_stages(st)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    if not r.profile:
        r.profile = True
        for _ in range(args.steps):
            add_stages(r.render_device(scene, out))
    r.profile = False
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    rays = st.rays_total
    value_all = rays / (ms_per_step * 1e-3) / 1e6

    # roofline of the dominant traversal kernel (bytes from the counters, time from events)
    prof = {"ms_stage": {k: v / args.steps for k, v in stage_ms.items()},
            "launches_stage": {k: v // args.steps for k, v in stage_n.items()}}
    wk = st_work.as_dict()["work"]
    # shadow queries whose light has no sample facing the hit point are not traced (their
    # answers cannot change the image): one visibility byte read each
    skipped = wk["shadow"]["skipped"]
    traced = rays - skipped
    # `value` counts the rays this build traces; the reference's own query count for the
    # frame (every World::hit call, skipped shadow queries included) gives value_all
    value = traced / (ms_per_step * 1e-3) / 1e6
    kernels = {
        "k_shadow": (kernel_bytes(wk["shadow"], st_work.rays_shadow - skipped, B_RAY_SHADOW) + skipped,
                     prof["ms_stage"]["shadow"], prof["launches_stage"]["shadow"]),
        "k_closest": (kernel_bytes(wk["closest"], st_work.rays_primary + st_work.rays_reflect, B_RAY_CLOSEST),
                      prof["ms_stage"]["closest"], prof["launches_stage"]["closest"]),
    }
    dom = max(kernels, key=lambda k: kernels[k][1])
    byt, ms, launches = kernels[dom]
    achieved = byt / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    line = {
        "metric": "Mrays/s at 1080p" if H == 1080 else f"Mrays/s at {W}x{H}", "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32+f64",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": args.workload, "width": W, "height": H, "aa": cfg["aa"], "samples": cfg["samples"],
                   "max_depth": cfg["max_depth"], "triangles": 81920 if args.workload.startswith("icosphere") else 0,
                   "rays_per_frame": rays, "primary": st.rays_primary, "reflect": st.rays_reflect,
                   "shadow": st.rays_shadow, "shadow_skipped": skipped, "rays_traced_per_frame": traced},
        "mrays_per_s_reference_queries": round(value_all, 2),
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_launch": int(byt / max(launches, 1)),
                     "ms_per_launch": round(ms / max(launches, 1), 4)},
        "stages_ms": {k: round(v, 3) for k, v in prof["ms_stage"].items()},
        "kernel_timing": ("hipEvents around each stage's launches in %d profiled steps after the timed ones"
                          % args.steps if args.no_stage_events else
                          "hipEvents around each stage's launches in the timed steps"),
        "device_ms_per_frame": round(sum(dev_ms) / len(dev_ms), 3),
        "work": wk,
        # this build's traversal per traced ray (k-d nodes and FP32 triangle tests, both stages)
        "work_per_ray": {"kd_nodes": round((wk["closest"]["nodes"] + wk["shadow"]["nodes"]) / max(traced, 1), 1),
                         "bvh_nodes": round((wk["closest"]["bvh_nodes"] + wk["shadow"]["bvh_nodes"]) / max(traced, 1), 1),
                         "tri_tests": round((wk["closest"]["tris"] + wk["shadow"]["tris"]) / max(traced, 1), 1)},
    }
    if not args.no_traffic:
        t, err = measure_traffic(args.workload, dom)
        if t:
            line["roofline"]["traffic"] = int(t["FETCH_SIZE"] + t["WRITE_SIZE"])
            line["roofline"]["traffic_note"] = (
                f"HBM bytes/launch = FETCH_SIZE*{t['FETCH_SIZE_scale']:.3f} + WRITE_SIZE*{t['WRITE_SIZE_scale']:.3f} "
                f"(KiB counters scaled by a {COPY_BYTES >> 20} MiB copy)")
            if "l2" in t:
                line["roofline"]["l2"] = dict(t["l2"], gbs=round(t["l2"]["bytes"] / (line["roofline"]["ms_per_launch"] * 1e-3) / 1e9, 1))
        else:
            line["roofline"]["traffic_note"] = err
    if not args.no_denoise:
        line["denoise"] = denoise_section([int(v) for v in args.dn_seeds.split(",")], args.dn_threshold)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(path, cfg)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
