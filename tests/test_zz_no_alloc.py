"""rt_render_device_noalloc (frame buffers in the library's fixed scratch array, no
device allocation) against rt_render_device, through the C ABI, on an MI355X.

Bar: bit-exact images, primary-ray hit records and ray counts, on two calls in a row
(the second with a smaller frame, which finds the first call's contents in the shared
array), plus the refusal of a frame whose buffers exceed the array.
"""
import ctypes
import os

import numpy as np
import pytest

from conftest import REF_SCENES

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def R():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    import rtrace_amd

    rtrace_amd.lib()
    return rtrace_amd


def _frame(R, scene, r, noalloc):
    import torch

    out = torch.full((r.height * r.width * 3,), 7, dtype=torch.uint8, device="cuda")
    st = r.render_device_noalloc(scene, out) if noalloc else r.render_device(scene, out)
    torch.cuda.synchronize()
    n = r.width * r.height + (r.width + 1) * (r.height + 1) if r.anti_aliasing_mode == R.AntiAliasingMode.Quincunx \
        else r.width * r.height * r.samples
    hits = R.last_primary_hits(n)
    return out.cpu().numpy().copy(), (st.rays_primary, st.rays_reflect, st.rays_shadow), hits


@pytest.mark.parametrize("sizes,aa", [(((160, 120), (96, 72)), "quincunx"), (((128, 96), (64, 40)), "stochastic")])
def test_noalloc_matches_render_device_twice(R, sizes, aa):
    # disk area light (16 rejection-sampled points per hit), plane, spheres: every frame buffer in use
    scene = R.Scene.from_json_file(os.path.join(REF_SCENES, "diffuse_light_demo.json"))
    for w, h in sizes:  # the second call is smaller: it must not depend on what the first left behind
        r = R.Renderer(w, h)
        if aa == "stochastic":
            r.anti_aliasing_mode = R.AntiAliasingMode.Stochastic
            r.samples = 4
        print(f"noalloc vs render_device: {aa} {w}x{h}", flush=True)
        img_a, rays_a, hits_a = _frame(R, scene, r, noalloc=True)
        img_b, rays_b, hits_b = _frame(R, scene, r, noalloc=False)
        assert rays_a == rays_b
        assert np.array_equal(img_a, img_b), f"{w}x{h}: {np.count_nonzero(img_a != img_b)} bytes differ"
        for a, b in zip(hits_a, hits_b):
            assert np.array_equal(a, b)


def test_noalloc_refuses_frames_beyond_the_scratch_array(R):
    import torch

    scene = R.Scene.from_json_file(os.path.join(REF_SCENES, "diffuse_light_demo.json"))
    r = R.Renderer(1920, 1080)
    out = torch.empty(1920 * 1080 * 3, dtype=torch.uint8, device="cuda")
    with pytest.raises(R.RtraceError) as e:
        r.render_device_noalloc(scene, out)
    assert e.value.code == -1 and "bytes of scratch" in str(e.value)
    assert str(R.NOALLOC_SCRATCH_BYTES) in str(e.value)
    r = R.Renderer(64, 48)
    r.max_depth = 65
    with pytest.raises(R.RtraceError) as e:
        r.render_device_noalloc(scene, torch.empty(64 * 48 * 3, dtype=torch.uint8, device="cuda"))
    assert e.value.code == -1 and "max_depth" in str(e.value)
    # a refused call leaves the entry point usable
    r.max_depth = 10
    small = torch.empty(64 * 48 * 3, dtype=torch.uint8, device="cuda")
    r.render_device_noalloc(scene, small)
    ref = torch.empty_like(small)
    r.render_device(scene, ref)
    assert torch.equal(small, ref)
