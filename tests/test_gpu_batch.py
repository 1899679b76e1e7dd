"""GPU tests of batched rendering (rt_batch_*): a batch of scenes x auto-camera views x
resolutions gives exactly the images and ray counts of rendering each job alone, with and
without the captured graph, for every raygen tile size, over repeated replays; one
auto-camera view is also checked against the oracle."""
import json
import os

import numpy as np
import pytest

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def R():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    import rtrace_amd

    rtrace_amd.lib()
    return rtrace_amd


@pytest.fixture(scope="module")
def workload(tmp_path_factory):
    from rtrace_amd import scenes
    d = str(tmp_path_factory.mktemp("c4"))
    return scenes.batch_scenes(d, n_scenes=3, seed=11)


SIZES = [(96, 72), (64, 64), (128, 72)]
MODES = [("quincunx", 1, 0), ("stochastic", 2, 5), ("no-jitter", 1, 0)]


def _jobs(R, workload):
    import torch
    jobs, singles = [], []
    for i, (path, _, _, _) in enumerate(workload):
        scene = R.Scene.from_json_file(path)
        cams = R.AutoCamera.generate_cameras(scene).cameras() + [None]
        for v, cam in enumerate(cams):
            w, h = SIZES[(i + v) % 3]
            aa, spp, ls = MODES[(i + 2 * v) % 3]
            r = R.Renderer(w, h)
            r.max_depth = 3 + (v % 2)
            r.samples = spp
            r.light_samples = ls
            r.anti_aliasing_mode = O.AA_BY_NAME[aa]
            out = torch.zeros(h * w * 3, dtype=torch.uint8, device="cuda")
            jobs.append((scene, cam, r, out))
    return jobs


def _single(R, scene, cam, r):
    """The same frame through rt_render_device with the camera written into the scene."""
    import torch
    out = torch.zeros(r.height * r.width * 3, dtype=torch.uint8, device="cuda")
    st = r.render_device(scene if cam is None else _with_camera(R, scene, cam), out)
    return out, st


_cam_scenes = {}


def _with_camera(R, scene, cam):
    key = (id(scene), tuple(cam.position), cam.kind)
    if key not in _cam_scenes:
        sc = json.loads(open(scene._path).read())
        sc["camera"] = {k: v for k, v in cam.to_json().items() if v is not None}
        _cam_scenes[key] = R.Scene.from_json_str(json.dumps(sc), os.path.dirname(scene._path))
    return _cam_scenes[key]


@pytest.mark.parametrize("use_graph,tile,group_gib,concurrent", [
    (1, 8, 0.0, 1), (0, 8, 0.0, 1), (1, 4, 0.0, 1), (1, 16, 0.0, 1), (1, 8, 0.004, 1), (0, 8, 1e-9, 1),
    (1, 8, 0.004, 3), (0, 8, 1e-9, 2), (1, 8, 1e-9, 4)])
def test_batch_equals_single_frames(R, workload, use_graph, tile, group_gib, concurrent):
    """group_gib > 0 runs the jobs in consecutive groups (1e-9: every job its own group);
    `concurrent` of them at once on their own stream branches and frame buffers, the groups of
    a branch sharing one set. The images and ray counts stay the same."""
    jobs = _jobs(R, workload)
    o = R.default_batch_opts()
    o.use_graph, o.tile, o.chunk, o.blocks_per_cu, o.group_gib = use_graph, tile, 256, 2, group_gib
    o.concurrent = concurrent
    b = R.Batch(jobs, o)
    st = b.run()
    assert st.n_jobs == len(jobs) and (st.graph_nodes > 0) == bool(use_graph)
    if group_gib == 1e-9:
        assert st.groups == len(jobs)
    elif group_gib == 0.0:
        assert st.groups == 1
    else:
        assert st.groups > 1
    if group_gib > 0.0:  # the branches' buffers are smaller than all jobs' together
        o1 = R.default_batch_opts()
        o1.use_graph, o1.tile, o1.chunk, o1.blocks_per_cu, o1.group_gib, o1.concurrent = 0, tile, 256, 2, 0.0, 1
        one = R.Batch(jobs, o1)
        assert st.bytes < one.run().bytes
        one.close()
        b.run()  # the outputs again from the grouped batch
    rays = [0, 0, 0]
    for scene, cam, r, out in jobs:
        ref, s1 = _single(R, scene, cam, r)
        assert np.array_equal(out.cpu().numpy(), ref.cpu().numpy())
        rays = [rays[0] + s1.rays_primary, rays[1] + s1.rays_reflect, rays[2] + s1.rays_shadow]
    assert [st.rays_primary, st.rays_reflect, st.rays_shadow] == rays
    b.close()


def test_batch_edge_jobs_equal_single_frames(R, workload, tmp_path):
    """Jobs of 1x1, one row, one column and odd sizes in every sampling mode, on a scene with no
    objects, one with no lights and a mesh scene, batched together (one group, and one group per
    job on 3 branches): each image and ray count equals the frame rendered alone."""
    import torch
    mesh_path = workload[0][0]
    with open(mesh_path) as f:
        base = json.load(f)
    paths = [mesh_path]
    for name, sc in (("no_objects", dict(base, objects=[])), ("no_lights", dict(base, lights=[]))):
        p = os.path.join(os.path.dirname(mesh_path), name + ".json")
        with open(p, "w") as f:
            json.dump(sc, f)
        paths.append(p)
    jobs = []
    for i, path in enumerate(paths):
        scene = R.Scene.from_json_file(path)
        for k, (w, h) in enumerate([(1, 1), (1, 9), (9, 1), (17, 5)]):
            aa, spp, ls = MODES[(i + k) % 3]
            r = R.Renderer(w, h)
            r.max_depth = 3
            r.samples = spp
            r.light_samples = ls
            r.anti_aliasing_mode = O.AA_BY_NAME[aa]
            jobs.append((scene, None, r, torch.zeros(h * w * 3, dtype=torch.uint8, device="cuda")))
    for group_gib, concurrent in ((0.0, 1), (1e-9, 3)):
        o = R.default_batch_opts()
        o.group_gib, o.concurrent = group_gib, concurrent
        b = R.Batch(jobs, o)
        st = b.run()
        rays = [0, 0, 0]
        for scene, cam, r, out in jobs:
            ref, s1 = _single(R, scene, cam, r)
            assert np.array_equal(out.cpu().numpy(), ref.cpu().numpy()), (r.width, r.height)
            rays = [rays[0] + s1.rays_primary, rays[1] + s1.rays_reflect, rays[2] + s1.rays_shadow]
        assert [st.rays_primary, st.rays_reflect, st.rays_shadow] == rays
        b.close()


@pytest.mark.parametrize("concurrent", [1, 3])
def test_graph_replays_are_identical(R, workload, concurrent):
    """Replays of the captured graph, one job group or every job its own group on 3 branches."""
    import torch
    jobs = _jobs(R, workload)
    o = R.default_batch_opts()
    o.group_gib, o.concurrent = (0.0, 1) if concurrent == 1 else (1e-9, concurrent)
    b = R.Batch(jobs, o)
    b.run()
    first = [o.clone() for *_, o in jobs]
    for _ in range(3):
        for *_, o in jobs:
            o.zero_()
        torch.cuda.synchronize()
        b.run()
        torch.cuda.synchronize()  # eager work between replays
        for f, (*_, o) in zip(first, jobs):
            assert torch.equal(f, o)
    b.close()


def test_auto_camera_view_matches_the_oracle(R, workload, tmp_path):
    path = workload[0][0]
    cam = R.AutoCamera.generate_cameras(path).perspective
    sc = json.loads(open(path).read())
    sc["camera"] = {k: v for k, v in cam.to_json().items() if v is not None}
    p2 = os.path.join(os.path.dirname(path), "persp_view.json")
    with open(p2, "w") as f:
        json.dump(sc, f)
    import torch
    scene = R.Scene.from_json_file(path)
    r = R.Renderer(120, 90)
    r.max_depth = 3
    out = torch.zeros(90 * 120 * 3, dtype=torch.uint8, device="cuda")
    R.Batch([(scene, cam, r, out)]).run()
    ref, _, _ = O.render(p2, O.make_opts(120, 90, max_depth=3))
    d = np.abs(out.cpu().numpy().reshape(90, 120, 3).astype(int) - ref.astype(int)).max(-1)
    assert (d <= 1).mean() >= 0.999 and d.max() <= 1  # the stated contract
    assert d.max() == 0  # what the build achieves: bit-exact


def test_autotune_picks_a_candidate(R, workload):
    jobs = _jobs(R, workload)[:6]
    o, ms = R.autotune_batch(jobs)
    assert ms > 0 and o.tile in (4, 8, 16) and o.chunk % 64 == 0 and 1 <= o.blocks_per_cu <= 64


def test_batch_rejects_bad_jobs(R, workload):
    import torch
    scene = R.Scene.from_json_file(workload[0][0])
    r = R.Renderer(32, 32)
    r.max_depth = 0
    out = torch.zeros(32 * 32 * 3, dtype=torch.uint8, device="cuda")
    with pytest.raises(R.RtraceError, match="max_depth"):
        R.Batch([(scene, None, r, out)])
    r.max_depth = 2
    o = R.default_batch_opts()
    o.tile = 5
    with pytest.raises(R.RtraceError, match="tile"):
        R.Batch([(scene, None, r, out)], o)
    o = R.default_batch_opts()
    o.concurrent = 0
    with pytest.raises(R.RtraceError, match="concurrent"):
        R.Batch([(scene, None, r, out)], o)
