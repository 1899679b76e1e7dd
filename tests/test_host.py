"""CPU-side checks of the product: the C-ABI library loads and exports every entry point
include/rtrace_hip.h declares (no compute call without a GPU), the CLI sizing glue,
and the procedural scene generators."""
import hashlib
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "rtrace_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?(?:int|void|char)\s*\*?\s*(rt_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("rt_ctx_create", "rt_render", "rt_scene_upload", "rt_intersect", "rt_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import ctypes

    import rtrace_amd

    lib = rtrace_amd.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_library_is_built_for_gfx950():
    import subprocess

    so = os.path.join(ROOT, "rtrace_amd", "librtrace_hip.so")
    out = subprocess.run(["/opt/rocm/llvm/bin/llvm-readelf", "-S", so], capture_output=True, text=True).stdout
    assert ".hip_fatbin" in out
    blob = open(so, "rb").read()
    assert b"gfx950" in blob


def test_cli_size_matches_reference_formula():
    from rtrace_amd import cli_size

    # the doc images' sizes (herebyfile.mjs -s 500) for 1:1, 4:3, 2:1, 5:3, 3:2 cameras
    assert cli_size(10, 10, 500) == (354, 354)
    assert cli_size(8, 6, 500) == (400, 300)
    assert cli_size(12, 6, 500) == (447, 224)
    assert cli_size(10, 6, 500) == (429, 257)
    assert cli_size(16, 9, 1469) == (1280, 720)
    assert cli_size(8, 6, 1000) == (800, 600)


def test_scene_generators_are_deterministic(tmp_path):
    from rtrace_amd import scenes

    a = scenes.icosphere_scene(str(tmp_path / "a"), subdiv=3)
    b = scenes.icosphere_scene(str(tmp_path / "b"), subdiv=3)
    assert open(a).read() == open(b).read()
    sa = open(os.path.join(os.path.dirname(a), "icosphere_s3.stl"), "rb").read()
    sb = open(os.path.join(os.path.dirname(b), "icosphere_s3.stl"), "rb").read()
    assert sa == sb
    assert int.from_bytes(sa[80:84], "little") == 20 * 4 ** 3


def test_icosphere_noise_bounds():
    from rtrace_amd import scenes

    t = scenes.icosphere(3, 1.5, 0.05, seed=2)
    r = np.linalg.norm(t.reshape(-1, 3), axis=1)
    assert r.min() >= 1.5 * 0.95 - 1e-12 and r.max() <= 1.5 * 1.05 + 1e-12
    assert r.std() > 0.01  # noise was applied


def test_splitmix64_known_answer():
    """SplitMix64 seeded with 0 (Steele, Lea, Flood 2014; also Java's SplittableRandom)."""
    from rtrace_amd.scenes import SplitMix64

    g = SplitMix64(0)
    assert [g.next_u64() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_batch_jobs_keep_scenes_and_outputs_alive():
    """A batch's captured launch holds raw device pointers of each job's scene and output
    (rt_batch_create): the Python job list must keep both objects alive with the batch."""
    import ctypes

    import rtrace_amd as R

    class Scene:
        _h = ctypes.c_void_p(0x1000)

    class Out:
        def data_ptr(self):
            return 0x2000

    scene, out = Scene(), Out()
    arr, keep = R._make_jobs([(scene, None, R.Renderer(8, 8), out)])
    assert any(k is scene for k in keep) and any(k is out for k in keep)
    assert arr[0].scene == 0x1000 and arr[0].d_rgb_out == 0x2000


def test_shipped_denoiser_parameter_sets():
    """One float32 parameter set per input sample count, each of the network's full size, and
    distinct; the fallback threshold keeps its share of the variance estimate's maximum."""
    import numpy as np
    from oracle import denoise_ref as D
    from rtrace_amd import denoise
    sets = {n: np.load(p, allow_pickle=False) for n, p in denoise.WEIGHTS_BY_SAMPLES.items()}
    assert sorted(sets) == [1, 2, 4]
    for n, w in sets.items():
        assert w.dtype == np.float32 and w.shape == (D.N_PARAMS,), n
        assert np.isfinite(w).all()
    assert not np.array_equal(sets[4], sets[2]) and not np.array_equal(sets[2], sets[1])
    assert denoise.fallback_threshold(4) == 0.06
    assert abs(denoise.fallback_threshold(2) - 0.12) < 1e-15
    assert abs(denoise.fallback_threshold(1) - 0.24) < 1e-15
