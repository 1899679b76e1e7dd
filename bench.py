"""Benchmark: Mrays/s of the per-pixel trace path at 1080p on one MI355X, the benchmark
configurations that fit one GPU, and the denoised-vs-plain frame time of the 1080p
denoiser workload.

Headline workload (default `icosphere_1080p`): the procedural mesh scene (81,920-triangle
noisy icosphere with a host-built k-d tree, 3 transformed cubes, a reflective grid plane,
a disk area light with 16 shadow samples per hit, fog 0.05, reflections to depth 4)
rendered at 1920x1080 with quincunx anti-aliasing. A step is one full frame (camera rays,
closest-hit, shading, shadow rays, reflections, resolve, quantise) rendered into a device
buffer; the scene is resident in HBM before timing.

The `configs` object times the other single-GPU configurations the same way:
`sphere_800x600` (config 1), `icosphere_square_720p` (config 2 as stated: 1280x720,
1x1 square light, 16 stratified samples) and `spheres_ring_1080p` (config 3: 1,000
reflective spheres, stochastic 4 spp, depth 10, outline). The `denoise` object is config 5:
1080p frames rendered plain with 256 light samples (the quality reference) and denoised
(4 samples + the trained U-Net + tile fallback), with PSNR / SSIM and per-stage times per
image and the denoiser's BF16 MFMA roofline.

Rays = primary + reflection + shadow rays, counted exactly by the pipeline.
Prints one JSON line; progress goes to stderr.
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
N_SIMD = 256 * 4  # MI355X: 256 CUs x 4 SIMDs
N_XCC = 8  # GRBM_GUI_ACTIVE is counted per XCC; rocprofv3 sums the 8 instances
# algorithmic bytes per unit of traversal work (layout in rtrace_amd/csrc/common/layout.hpp)
B_NODE, B_TRI, B_RETEST = 40, 48, 72  # node: 32 B box+links, +8 B f64 split (inner)
B_BVH_NODE = 128  # 4-wide mesh BVH node: four slot boxes and links
B_SPHERE, B_PLANE, B_CUBE = 32, 48, 304
B_RAY_CLOSEST = 64 + 16  # RayRec in, HitRec out
B_RAY_SHADOW = 48 + 24 + 1  # hit point + normal, light sample point in, visibility byte out
TRIANGLES = {"icosphere": 81920}


def _progress(msg):
    """A progress line on stderr (the JSON result line stays alone on stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


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


def host_threads():
    """The host cores this process may use: its CPU affinity, bounded by OMP_NUM_THREADS when the
    environment sets it (a one-GPU box grants a share of its host, 16 threads)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def cpu_baseline(path, cfg, budget_s=20.0):
    """The CPU f64 oracle (a restatement of the reference renderer) on the box's host cores,
    on a reduced-resolution render of the same scene and settings; and, as the checker, the
    parity of the HIP render at that resolution against it (primary-ray hit ids, pixels)."""
    import numpy as np

    import rtrace_amd as R
    from oracle import pyoracle as O
    threads = host_threads()
    w, h = 64, 36
    with open(path) as f:
        scene = json.load(f)
    outline = O.outline_from_scene(scene) if cfg["aa"] != "quincunx" else None
    spent = 0.0
    best = None
    while True:
        opts = O.make_opts(w, h, max_depth=cfg["max_depth"], samples=cfg["samples"], aa=cfg["aa"], threads=threads,
                           outline=outline)
        img, st, hits = O.render(path, opts, want_hits=True)
        rays = st["rays_primary"] + st["rays_reflect"] + st["rays_shadow"]
        best = (rays / st["render_seconds"] / 1e6, w, h, st["render_seconds"], rays,
                {"kd_nodes": (st["kd_inner_visits"] + st["kd_leaf_visits"]) / rays, "tri_tests": st["tri_tests"] / rays},
                img, hits, st)
        spent += st["render_seconds"]
        if st["render_seconds"] * 4 > budget_s - spent or w >= cfg["width"]:
            break
        w, h = w * 2, h * 2
    # one thread (the scalar baseline) on the smallest sample
    opts = O.make_opts(64, 36, max_depth=cfg["max_depth"], samples=cfg["samples"], aa=cfg["aa"], threads=1,
                       outline=outline)
    _, st1, _ = O.render(path, opts)
    rays1 = st1["rays_primary"] + st1["rays_reflect"] + st1["rays_shadow"]

    # parity of the HIP path on the same sample (the oracle is the checker here)
    bw, bh, oimg, ohits, ost = best[1], best[2], best[6], best[7], best[8]
    r = R.Renderer(bw, bh)
    configure(R, r, cfg, R.Scene.from_json_file(path))
    gimg = r.render(R.Scene.from_json_file(path))
    n_prim = bw * bh + (bw + 1) * (bh + 1) if cfg["aa"] == "quincunx" else bw * bh * cfg["samples"]
    gh = R.last_primary_hits(n_prim)
    mism = int(np.count_nonzero((gh[0] != ohits[0]) | (gh[1] != ohits[1])))
    d = np.abs(gimg.astype(int) - oimg.astype(int)).max(-1)
    gst = r.last_stats
    parity = {"sample": f"{bw}x{bh} render of the benched scene and settings", "primary_rays": n_prim,
              "hit_id_mismatches": mism, "pixels": int(d.size), "pixels_exact": int(np.count_nonzero(d == 0)),
              "pixels_within_1lsb_frac": round(float(np.mean(d <= 1)), 6), "max_abs_diff": int(d.max()),
              "ray_counts_equal": (gst.rays_primary, gst.rays_reflect, gst.rays_shadow) ==
                                  (ost["rays_primary"], ost["rays_reflect"], ost["rays_shadow"])}
    base = {"value": best[0], "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"oracle render of the same scene at {best[1]}x{best[2]} ({best[4]} rays, {best[3]:.1f} s, "
                      f"{threads} threads = this process's host CPU share)",
            # the reference's traversal (no early exit) per ray on this sample: an upper bound of
            # the traversal work, never the roofline's numerator
            "reference_traversal_per_ray": {k: round(v, 1) for k, v in best[5].items()},
            "one_thread": {"value": rays1 / st1["render_seconds"] / 1e6, "unit": "Mrays/s", "cores": 1,
                           "sample": f"64x36 ({rays1} rays, {st1['render_seconds']:.1f} s)"}}
    return base, parity


COPY_BYTES = 1 << 30

BF16_PEAK_TFLOPS = 2500.0  # MI355X dense bf16 MFMA
# Denoiser multiply-accumulates per full-resolution pixel: sum over layers of
# 9 * cin * cout / 4^level (13 real input features, 3 outputs; layer table in
# include/rtrace_denoise.h)
DN_LAYERS = [(13, 32, 0), (32, 32, 0), (32, 64, 1), (64, 64, 1), (64, 128, 2), (128, 128, 2), (128, 128, 3),
             (128, 128, 3), (256, 128, 2), (128, 128, 2), (192, 64, 1), (64, 64, 1), (96, 32, 0), (32, 32, 0),
             (32, 3, 0)]
DN_MAC_PER_PX = sum(9 * ci * co / 4 ** lv for ci, co, lv in DN_LAYERS)


def denoise_section(seeds, threshold, low=4, high=256, variant_samples=(2, 1)):
    """The denoiser workload at 1080p: the plain render with `high` light samples (the quality reference) against
    the denoised render (`low` samples + network + tile fallback), per image; then the same with the networks
    trained on renders with `variant_samples` light samples per hit."""
    import rtrace_amd as R  # noqa: F401 - loads the library
    from rtrace_amd import denoise
    dn = denoise.Denoiser.load()
    rows = denoise.evaluate(dn, seeds, 1920, 1080, low=low, high=high, threshold=threshold, repeats=3,
                            log=lambda m: _progress("denoise " + m[:160]))
    # the network timed on its own (the denoised frame runs the tile fallback beside it)
    ms_dn_net = sorted(r["ms_network"] for r in rows)[len(rows) // 2]
    flops = 2.0 * DN_MAC_PER_PX * 1920 * 1080
    ach = flops / (ms_dn_net * 1e-3) / 1e12
    plain = sum(r["ms_plain_high"] for r in rows) / len(rows)
    den = sum(r["ms_denoised"] for r in rows) / len(rows)
    # cheaper network inputs: the same frames from renders with fewer light samples per hit, each
    # with the network trained on them and the fallback threshold at the same share of the
    # variance estimate's maximum
    variants = []
    for n in variant_samples:
        dv = denoise.Denoiser.load(denoise.WEIGHTS_BY_SAMPLES[n])
        thr = denoise.fallback_threshold(n, threshold)
        vr = denoise.evaluate(dv, seeds, 1920, 1080, low=n, high=high, threshold=thr, repeats=3,
                              log=lambda m, n=n: _progress(f"denoise variant {n}: " + m[:140]))
        vp = sum(r["ms_plain_high"] for r in vr) / len(vr)
        vd = sum(r["ms_denoised"] for r in vr) / len(vr)
        variants.append({
            "low_samples": n, "threshold": thr, "ms_per_frame_plain": round(vp, 3),
            "ms_per_frame_denoised": round(vd, 3), "speedup": round(vp / vd, 2),
            "psnr_min": min(r["psnr_denoised"] for r in vr), "ssim_min": min(r["ssim_denoised"] for r in vr),
            "psnr_low_min": min(r["psnr_low"] for r in vr),
            "per_image": [{k: r[k] for k in ("seed", "ms_denoised", "speedup", "psnr_low", "psnr_denoised",
                                             "ssim_denoised", "tiles_fallback", "stage_ms", "render_stages_ms")}
                          for r in vr]})
        del dv
    return {
        "workload": "denoise_1080p (icosphere scenes, 2x2 square light), 1920x1080 quincunx, depth 4",
        "low_samples": low, "high_samples": high, "threshold": threshold, "tile": 16,
        "ms_per_frame_plain": round(plain, 3), "ms_per_frame_denoised": round(den, 3),
        "speedup": round(plain / den, 2),
        "psnr_min": min(r["psnr_denoised"] for r in rows), "psnr_low_min": min(r["psnr_low"] for r in rows),
        "per_image": rows,
        "variants": variants,
        "roofline": {"bound": "mfma", "kernel": "denoiser (prep, 15 conv3x3, pools, compose)",
                     "achieved": round(ach, 1), "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(ach / BF16_PEAK_TFLOPS, 4), "flop_per_frame": flops,
                     "ms_per_frame": ms_dn_net},
    }


def counters_child(workload):
    """Profiled child: one frame of the workload, one denoiser forward pass on 1080p features,
    then a plain copy of COPY_BYTES (the calibration kernel for the memory counters), in one
    process."""
    import torch

    import rtrace_amd as R
    from rtrace_amd import denoise
    path, cfg = make_workload(workload, os.path.join(tempfile.gettempdir(), "rtrace_amd_bench"))
    scene = R.Scene.from_json_file(path)
    r = R.Renderer(cfg["width"], cfg["height"])
    configure(R, r, cfg, scene)
    out = torch.empty(cfg["width"] * cfg["height"] * 3, dtype=torch.uint8, device="cuda")
    r.render_device(scene, out)
    dn = denoise.Denoiser.load()
    feat = torch.rand(1080, 1920, 13, dtype=torch.float32, device="cuda")
    dn.denoise(feat)
    torch.cuda.synchronize()
    src = torch.ones(COPY_BYTES, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()


# One rocprofv3 --pmc pass per entry (a pass holds at most 8 SQ, 4 TCC, 2 GRBM counters).
PMC_PASSES = {
    "fetch": "FETCH_SIZE",
    "write": "WRITE_SIZE",
    "l2": "TCC_REQ_sum TCC_HIT_sum",
    "sq": "SQ_WAVE_CYCLES SQ_WAIT_ANY SQ_ACTIVE_INST_VALU SQ_THREAD_CYCLES_VALU SQ_BUSY_CYCLES "
          "SQ_VALU_MFMA_BUSY_CYCLES SQ_INSTS_VALU_MFMA_MOPS_BF16 GRBM_GUI_ACTIVE",
}


def run_counters(workload):
    """Runs the PMC passes over counters_child; returns {pass: [(kernel, dispatch, counter, value)]}."""
    import csv
    import glob
    import shutil
    import subprocess

    prof = shutil.which("rocprofv3")
    if not prof:
        return None, "rocprofv3 not found"
    res = {}
    for name, counters in PMC_PASSES.items():
        d = tempfile.mkdtemp(prefix="rt_pmc_")
        cmd = [prof, "--pmc", *counters.split(), "--output-format", "csv", "-d", d, "-o", "run", "--", sys.executable,
               os.path.abspath(__file__), "--workload", workload, "--counters-child"]
        try:
            subprocess.run(cmd, check=True, timeout=300, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        except Exception as e:  # noqa: BLE001 - report and carry on without counters
            return None, f"{name} pass failed: {e}"
        rows = []
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(f) as fh:
                rows += list(csv.DictReader(fh))
        shutil.rmtree(d, ignore_errors=True)
        res[name] = [(r["Kernel_Name"], r.get("Dispatch_Id", ""), r["Counter_Name"], float(r["Counter_Value"]))
                     for r in rows]
        _progress(f"pmc pass {name}: {len(rows)} rows")
    return res, None


def _per_dispatch(rows, counter, match):
    vals = {}
    for kname, disp, cname, v in rows:
        if cname == counter and match(kname):
            vals[disp] = vals.get(disp, 0.0) + v
    return list(vals.values())


def _kern(name):
    return lambda n: re.search(r"::%s[<(]" % name, n) is not None


def _is_copy(n):
    return "copy" in n.lower()


def counter_summary(pmc, kernel, ms_per_launch):
    """HBM and L2 traffic per launch of `kernel` (KiB counters scaled by the copy kernel's
    bytes per count) and what bounds it: VALU lane activity, VALU busy share and the share
    of wave time spent waiting."""
    out = {}
    copy_f = [c for c in _per_dispatch(pmc["fetch"], "FETCH_SIZE", _is_copy) if c > 1000]
    copy_w = [c for c in _per_dispatch(pmc["write"], "WRITE_SIZE", _is_copy) if c > 1000]
    kf = _per_dispatch(pmc["fetch"], "FETCH_SIZE", _kern(kernel))
    kw = _per_dispatch(pmc["write"], "WRITE_SIZE", _kern(kernel))
    if kf and kw and copy_f and copy_w:
        sf = COPY_BYTES / (max(copy_f) * 1024.0)
        sw = COPY_BYTES / (max(copy_w) * 1024.0)
        hbm = (sum(kf) / len(kf)) * 1024.0 * sf + (sum(kw) / len(kw)) * 1024.0 * sw
        out["hbm_bytes"] = int(hbm)
        out["hbm_gbs"] = round(hbm / (ms_per_launch * 1e-3) / 1e9, 1)
        out["hbm_frac"] = round(out["hbm_gbs"] / HBM_PEAK_GBS, 4)
        out["hbm_note"] = f"FETCH_SIZE*{sf:.3f} + WRITE_SIZE*{sw:.3f} (KiB counters scaled by a {COPY_BYTES >> 20} MiB copy)"
    kreq = _per_dispatch(pmc["l2"], "TCC_REQ_sum", _kern(kernel))
    khit = _per_dispatch(pmc["l2"], "TCC_HIT_sum", _kern(kernel))
    creq = [c for c in _per_dispatch(pmc["l2"], "TCC_REQ_sum", _is_copy) if c > 1000]
    if kreq and creq:
        bpr = 2.0 * COPY_BYTES / max(creq)  # each copied byte is one L2 read and one L2 write
        req = sum(kreq) / len(kreq)
        out["l2_bytes"] = int(req * bpr)
        out["l2_gbs"] = round(req * bpr / (ms_per_launch * 1e-3) / 1e9, 1)
        out["l2_hit_rate"] = round(sum(khit) / max(sum(kreq), 1.0), 4) if khit else None
    sq = pmc["sq"]
    g = {c: _per_dispatch(sq, c, _kern(kernel)) for c in PMC_PASSES["sq"].split()}
    if all(g.values()):
        m = {c: sum(v) / len(v) for c, v in g.items()}
        out["valu_lanes_active"] = round(m["SQ_THREAD_CYCLES_VALU"] / max(m["SQ_ACTIVE_INST_VALU"] * 64.0, 1.0), 4)
        gui = max(m["GRBM_GUI_ACTIVE"] / N_XCC, 1.0)  # GPU cycles of the dispatch
        out["valu_busy"] = round(m["SQ_ACTIVE_INST_VALU"] * 4.0 / N_SIMD / gui, 4)
        out["wave_wait_frac"] = round(m["SQ_WAIT_ANY"] / max(m["SQ_WAVE_CYCLES"], 1.0), 4)
    # the limiter, from the measurements
    if out.get("hbm_frac", 0) >= 0.6:
        out["limiter"] = "hbm bandwidth"
    elif out.get("valu_busy", 0) >= 0.6:
        out["limiter"] = (f"valu issue: the SIMDs issue VALU {out['valu_busy'] * 100:.0f}% of the cycles with "
                          f"{out.get('valu_lanes_active', 0) * 64:.0f} of 64 lanes active (divergent traversal); "
                          f"HBM at {out.get('hbm_frac', 0) * 100:.1f}% of peak")
    elif out.get("wave_wait_frac", 0) >= 0.5:
        out["limiter"] = "memory latency (waves mostly waiting on dependent loads; HBM and VALU far from peak)"
    elif "valu_busy" in out:
        out["limiter"] = "mixed (no unit near its peak)"
    out["counters"] = ("valu_busy = 4 SQ_ACTIVE_INST_VALU / (1024 SIMDs x GRBM_GUI_ACTIVE / 8 XCCs); valu_lanes_active = "
                       "SQ_THREAD_CYCLES_VALU / (64 SQ_ACTIVE_INST_VALU); wave_wait_frac = SQ_WAIT_ANY / SQ_WAVE_CYCLES")
    return out


def mfma_summary(pmc):
    """MFMA busy share of the denoiser's convolutions (every conv kernel: SQ_VALU_MFMA_BUSY_CYCLES
    per SIMD over the dispatches' GPU cycles), and the bf16 MFMA work they did."""
    sq = pmc["sq"]
    match = lambda n: re.search(r"::k_conv3x3(_tile|_c32|_rows)?[<(]", n) is not None  # noqa: E731
    busy = _per_dispatch(sq, "SQ_VALU_MFMA_BUSY_CYCLES", match)
    gui = _per_dispatch(sq, "GRBM_GUI_ACTIVE", match)
    mops = _per_dispatch(sq, "SQ_INSTS_VALU_MFMA_MOPS_BF16", match)
    if not busy or not gui:
        return None
    return {"kernel": "dn::k_conv3x3_c32, _rows and _tile (the 15 layers)", "launches": len(busy),
            "mfma_busy": round(sum(busy) / max(sum(gui) / N_XCC * N_SIMD, 1.0), 4),
            "bf16_mops_per_forward": round(sum(mops), 1) if mops else None,
            "counter": "SQ_VALU_MFMA_BUSY_CYCLES / (1024 SIMDs x GRBM_GUI_ACTIVE / 8 XCCs), summed over the launches"}


def measure_frame(R, torch, workload, steps, warmup, stage_events=True, count=True):
    """Times `steps` frames of a workload after `warmup` ones. With stage_events the timed frames
    record hipEvents around each stage's launches (the per-launch kernel times come from the timed
    region); otherwise the timed frames replay the frame's hipGraph and the stage times come from as
    many profiled frames after them."""
    out_dir = os.path.join(tempfile.gettempdir(), "rtrace_amd_bench")
    path, cfg = make_workload(workload, out_dir)
    scene = R.Scene.from_json_file(path)
    W, H = cfg["width"], cfg["height"]
    r = R.Renderer(W, H)
    configure(R, r, cfg, scene)
    out = torch.empty(W * H * 3, dtype=torch.uint8, device="cuda")
    st_work = None
    if count:  # untimed: exact work counters
        r.count_work = True
        st_work = r.render_device(scene, out)
        r.count_work = False
    stage_ms, stage_n = {}, {}

    def add_stages(st_):
        d_ = st_.as_dict()
        for k, v in d_["ms_stage"].items():
            stage_ms[k] = stage_ms.get(k, 0.0) + v
            stage_n[k] = stage_n.get(k, 0) + d_["launches_stage"][k]

    r.profile = stage_events
    for _ in range(warmup):
        r.render_device(scene, out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev_ms = []
    for _ in range(steps):
        st = r.render_device(scene, out)
        dev_ms.append(st.ms_render)
        if r.profile:
            add_stages(st)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    if not r.profile:
        r.profile = True
        for _ in range(steps):
            add_stages(r.render_device(scene, out))
    r.profile = False
    ms = (t1 - t0) * 1e3 / steps
    rays = st.rays_total
    res = {"workload": workload, "width": W, "height": H, "aa": cfg["aa"], "samples": cfg["samples"],
           "max_depth": cfg["max_depth"], "steps": steps, "warmup": warmup, "ms_per_step": round(ms, 3),
           "device_ms_per_frame": round(sum(dev_ms) / len(dev_ms), 3),
           "mrays_per_s": round(rays / (ms * 1e-3) / 1e6, 2), "rays_per_frame": rays,
           "primary": st.rays_primary, "reflect": st.rays_reflect, "shadow": st.rays_shadow,
           "stages_ms": {k: round(v / steps, 3) for k, v in stage_ms.items()},
           "launches_stage": {k: v // steps for k, v in stage_n.items()},
           "kernel_timing": ("hipEvents around each stage's launches in the timed steps" if stage_events else
                             f"graph replays timed; stage times from {steps} profiled steps after them")}
    if st_work is not None:
        wk = st_work.as_dict()["work"]
        skipped = wk["shadow"]["skipped"]
        res["work"] = wk
        res["shadow_skipped"] = skipped
        res["rays_traced_per_frame"] = rays - skipped
        # `value` counts the rays this build traces; the reference's own query count for the
        # frame (every World::hit call, skipped shadow queries included) gives mrays_per_s
        res["mrays_per_s_traced"] = round((rays - skipped) / (ms * 1e-3) / 1e6, 2)
        res["_st_work"] = st_work
    return res, path, cfg


def bench_batch(args, steps=None, warmup=None, emit=True):
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
    opts0 = R.default_batch_opts()
    if args.c4_group_gib is not None:
        opts0.group_gib = args.c4_group_gib
    if args.c4_concurrent is not None:
        opts0.concurrent = args.c4_concurrent
    # the subset is tuned with job groups scaled to its share of the jobs, so that it runs about as
    # many groups, as many of them in flight, as the whole batch
    tune_gib = opts0.group_gib * len(tune_jobs) / len(jobs)
    tune0 = R.BatchOpts.from_buffer_copy(opts0)
    tune0.group_gib = tune_gib
    tuned, tune_ms = R.autotune_batch(tune_jobs, tune0)
    opts = R.BatchOpts.from_buffer_copy(tuned)
    opts.group_gib = opts0.group_gib
    _progress(f"autotuned on {len(tune_jobs)} jobs: {opts.as_dict()}")
    b = R.Batch(jobs, opts)
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    for _ in range(warmup):
        b.run()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_dev = []
    for _ in range(steps):
        st = b.run()
        ms_dev.append(st.ms)
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t1) * 1e3 / steps
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
        "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32+f64",
        "data": "synthetic (seeded procedural scenes)",
        "config": {"workload": f"batch64: {len(work)} scenes x 4 auto-camera views, quincunx, depth 4, disk light "
                               f"8 samples", "jobs": len(jobs), "pixels": st.pixels,
                   "triangles_min": min(tris), "triangles_max": max(tris), "rays_per_step": rays,
                   "primary": st.rays_primary, "reflect": st.rays_reflect, "shadow": st.rays_shadow},
        "batch": {"opts": opts.as_dict(), "tune_ms_subset": round(tune_ms, 3), "tune_jobs": len(tune_jobs),
                  "tune_group_gib": round(tune_gib, 4),
                  "graph_nodes": st.graph_nodes, "graph_launches_per_step": 1, "device_ms": round(min(ms_dev), 3),
                  "frame_buffer_gib": round(st.bytes / 2**30, 2), "job_groups": st.groups,
                  "groups_in_flight": opts.concurrent,
                  "build_ms": round(st.ms_build, 1),
                  "scene_load_s": round(t_load, 1)},
    }
    if seq_ms is not None:
        line["sequential"] = {"ms": round(seq_ms, 3), "mrays_s": round(rays / (seq_ms * 1e-3) / 1e6, 2),
                              "batch_speedup": round(seq_ms / ms_step, 3)}
    del b, jobs
    if emit:
        print(json.dumps(line))
    return line


SUB_CONFIGS = [("config 1", "sphere_800x600", False), ("config 2 as stated", "icosphere_square_720p", True),
               ("config 3", "spheres_ring_1080p", True)]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="icosphere_1080p")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 counter passes")
    ap.add_argument("--counters-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--no-denoise", action="store_true", help="skip the denoised-vs-plain section")
    ap.add_argument("--no-configs", action="store_true", help="skip the other configurations")
    ap.add_argument("--no-stage-events", action="store_true",
                    help="time the steps without per-stage hipEvents (stage times from as many profiled steps after)")
    ap.add_argument("--dn-seeds", default="100001,100002", help="denoiser scene seeds of the denoise section")
    ap.add_argument("--dn-threshold", type=float, default=0.06, help="tile fallback threshold")
    ap.add_argument("--c4-scenes", type=int, default=64, help="batch64: number of scenes")
    ap.add_argument("--no-sequential", action="store_true", help="batch64: skip the one-by-one comparison")
    ap.add_argument("--no-batch", action="store_true", help="skip the batch64 entry of the default run's configs")
    ap.add_argument("--c4-group-gib", type=float, default=None,
                    help="batch64: GiB of frame buffers per job group (default: the library's)")
    ap.add_argument("--c4-concurrent", type=int, default=None,
                    help="batch64: job groups in flight at once (default: the library's)")
    args = ap.parse_args()
    if args.counters_child:
        counters_child(args.workload)
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

    head, path, cfg = measure_frame(R, torch, args.workload, args.steps, args.warmup,
                                    stage_events=not args.no_stage_events)
    _progress(f"{args.workload}: {head['ms_per_step']} ms/frame, {head['mrays_per_s_traced']} Mrays/s traced")
    W, H = cfg["width"], cfg["height"]
    st_work = head.pop("_st_work")
    wk = head["work"]
    skipped = head["shadow_skipped"]
    ms_stage = head["stages_ms"]
    launches = head["launches_stage"]
    # roofline of the dominant traversal kernel (algorithmic bytes from the counters, time from events)
    kernels = {
        "k_shadow": (kernel_bytes(wk["shadow"], st_work.rays_shadow - skipped, B_RAY_SHADOW) + skipped,
                     ms_stage["shadow"], launches["shadow"]),
        "k_closest": (kernel_bytes(wk["closest"], st_work.rays_primary + st_work.rays_reflect, B_RAY_CLOSEST),
                      ms_stage["closest"], launches["closest"]),
    }
    dom = max(kernels, key=lambda k: kernels[k][1])
    byt, ms, nl = kernels[dom]
    ms_launch = ms / max(nl, 1)
    achieved = byt / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    traced = head["rays_traced_per_frame"]
    line = {
        "metric": "Mrays/s at 1080p" if H == 1080 else f"Mrays/s at {W}x{H}", "value": head["mrays_per_s_traced"],
        "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32+f64", "data": "synthetic (seeded procedural scene)",
        "config": {"workload": args.workload, "width": W, "height": H, "aa": cfg["aa"], "samples": cfg["samples"],
                   "max_depth": cfg["max_depth"], "triangles": TRIANGLES.get(args.workload.split("_")[0], 0),
                   "rays_per_frame": head["rays_per_frame"], "primary": head["primary"], "reflect": head["reflect"],
                   "shadow": head["shadow"], "shadow_skipped": skipped, "rays_traced_per_frame": traced},
        "mrays_per_s_reference_queries": head["mrays_per_s"],
        # `bound` names the roofline the figure is priced on (the memory roofline: traversal is
        # pointer chasing; no MFMA in tracing); `limiter` says what the counters show bounds it
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "achieved_is": "algorithmic bytes (every node, triangle record and ray record the walk "
                                    "reads or writes, at their layout sizes) per launch / launch time",
                     "bytes_per_launch": int(byt / max(nl, 1)), "ms_per_launch": round(ms_launch, 4)},
        "stages_ms": ms_stage,
        "kernel_timing": head["kernel_timing"],
        "device_ms_per_frame": head["device_ms_per_frame"],
        "work": wk,
        # this build's traversal per traced ray (k-d nodes and FP32 triangle tests, both stages)
        "work_per_ray": {"kd_nodes": round((wk["closest"]["nodes"] + wk["shadow"]["nodes"]) / max(traced, 1), 1),
                         "bvh_nodes": round((wk["closest"]["bvh_nodes"] + wk["shadow"]["bvh_nodes"]) / max(traced, 1), 1),
                         "tri_tests": round((wk["closest"]["tris"] + wk["shadow"]["tris"]) / max(traced, 1), 1)},
    }
    if not args.no_configs:
        subs = {}
        for label, wl, count in SUB_CONFIGS:
            res, _, _ = measure_frame(R, torch, wl, max(5, args.steps // 2), max(2, args.warmup // 2),
                                      stage_events=False, count=count)
            res.pop("_st_work", None)
            res.pop("work", None)
            res["config"] = label
            subs[wl] = res
            _progress(f"{wl}: {res['ms_per_step']} ms/frame, {res['mrays_per_s']} Mrays/s")
        if not args.no_batch:
            # config 4 in the same run: the 64-scene batch, a few steps (its own `--workload batch64`
            # run times more)
            bl = bench_batch(args, steps=3, warmup=1, emit=False)
            subs["batch64"] = {"config": "config 4", "workload": bl["config"]["workload"], "mrays_per_s": bl["value"],
                               "ms_per_step": bl["ms_per_step"], "steps": bl["steps"], "warmup": bl["warmup"],
                               "jobs": bl["config"]["jobs"], "rays_per_step": bl["config"]["rays_per_step"],
                               "frame_buffer_gib": bl["batch"]["frame_buffer_gib"],
                               "job_groups": bl["batch"]["job_groups"], "opts": bl["batch"]["opts"],
                               "sequential": bl.get("sequential")}
            _progress(f"batch64: {bl['ms_per_step']} ms/step, {bl['value']} Mrays/s")
        line["configs"] = subs
    if not args.no_traffic:
        pmc, err = run_counters(args.workload)
        if pmc:
            cs = counter_summary(pmc, dom, ms_launch)
            line["roofline"]["traffic"] = cs.get("hbm_bytes")
            line["roofline"]["measured"] = cs
            if "limiter" in cs:
                line["roofline"]["limiter"] = cs["limiter"]
            mf = mfma_summary(pmc)
            if mf:
                line["denoise_mfma_counters"] = mf
        else:
            line["roofline"]["traffic_note"] = err
    if not args.no_denoise:
        line["denoise"] = denoise_section([int(v) for v in args.dn_seeds.split(",")], args.dn_threshold)
        if "denoise_mfma_counters" in line:
            line["denoise"]["roofline"]["mfma_counters"] = line.pop("denoise_mfma_counters")
    if not args.no_cpu_baseline:
        line["cpu_baseline"], line["parity"] = cpu_baseline(path, cfg)
        _progress(f"cpu baseline {line['cpu_baseline']['value']:.3f} Mrays/s; parity {line['parity']}")
    print(json.dumps(line))


if __name__ == "__main__":
    main()
