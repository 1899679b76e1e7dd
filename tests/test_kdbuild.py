"""Host k-d build (KdTree::new / build_recursive, src/mesh.rs:75-197): the concurrent
subtree build produces the serial build's arrays exactly (nodes, triangle records, f64
node bounds, roots, leaf statistics) on seeded random meshes, including a flat mesh with
a degenerate triangle. The check program is compiled with g++ from the library's own
builder source."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_parallel_kd_build_matches_serial(tmp_path):
    exe = str(tmp_path / "kdbuild_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-pthread", "-o", exe,
                    os.path.join(ROOT, "tests", "cpp", "kdbuild_check.cpp"),
                    os.path.join(ROOT, "rtrace_amd", "csrc", "host", "kdbuild.cpp")], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip() == "ok 6"


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_mesh_bvh_invariants(tmp_path):
    """Mesh BVH (this build's traversal structure): every triangle in exactly one leaf with
    its exact record, bounded leaves and depth, and outward-rounded f32 child boxes that
    hold every triangle below them, on random, flat, tiny and equal-centroid meshes; the
    4-wide walked form holds the same leaves within half the depth."""
    exe = str(tmp_path / "bvh_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-o", exe, os.path.join(ROOT, "tests", "cpp", "bvh_check.cpp"),
                    os.path.join(ROOT, "rtrace_amd", "csrc", "host", "mesh_bvh.cpp")], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip() == "ok 21"


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_object_bvh_depth_fits_the_traversal_stack(tmp_path):
    """Object BVH over spheres (scenes of 16 or more spheres, e.g. the 1,000-sphere ring):
    its walk (obvh_visit) pushes one entry per level into a 16-entry per-lane stack, so the
    build bounds the depth by growing leaves past 65,536 spheres (300k spheres: depth <= 16,
    every sphere in exactly one leaf, boxes nested)."""
    exe = str(tmp_path / "obvh_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-o", exe, os.path.join(ROOT, "tests", "cpp", "obvh_check.cpp"),
                    os.path.join(ROOT, "rtrace_amd", "csrc", "host", "object_bvh.cpp")], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip() == "ok 16"


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.exists("/opt/rocm/include/hip/hip_runtime.h"),
                    reason="needs g++ and the HIP headers")
def test_mesh_walk_matches_brute_force_on_the_host(tmp_path):
    """The device mesh walk (mesh_trace, compiled for the host) against brute force over all
    triangles with the exact f64 test, for shadow (any-hit) and closest-hit rays from near
    the mesh and from 1e3 to 3e6 mesh radii away (re-based FP32 origin), on a 4,992-triangle
    noisy sphere and a 3,000-triangle random soup: the same answers for all 32,000 rays,
    and far shadow rays test about as many triangles as near ones (the walk from the far
    origin itself tested 868 per ray on the sphere)."""
    exe = str(tmp_path / "walk_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    "-pthread", "-o", exe, os.path.join(ROOT, "tests", "cpp", "walk_check.cpp"),
                    os.path.join(ROOT, "rtrace_amd", "csrc", "host", "kdbuild.cpp"),
                    os.path.join(ROOT, "rtrace_amd", "csrc", "host", "mesh_bvh.cpp")], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    print(out.stdout)
    assert out.stdout.strip().splitlines()[-1] == "ok 32000"


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.exists("/opt/rocm/include/hip/hip_runtime.h"),
                    reason="needs g++ and the HIP headers")
def test_shadow_task_order_is_a_permutation(tmp_path):
    """The shadow stage walks its queue light by light within blocks of 64 hits
    (shadow_task, compiled for the host): every task is visited exactly once for mixes of
    point, disk and square lights and hit counts around the block size, and a point light's
    items in a full block are 64 different hits' rays toward it."""
    exe = str(tmp_path / "shadow_order_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    "-o", exe, os.path.join(ROOT, "tests", "cpp", "shadow_order_check.cpp")], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip() == "ok 48"


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.exists("/opt/rocm/include/hip/hip_runtime.h"),
                    reason="needs g++ and the HIP headers")
def test_plane_prefilter_is_conservative(tmp_path):
    """plane_may_hit (the FP32 prefilter before the f64 plane test, compiled for the host)
    never rejects a ray the exact test accepts: random planes and rays up to 1e5 units out,
    rays nearly parallel to the plane around the 1e-8 threshold, and rays whose t sits ulps
    from tmin or tmax; and it settles at least 90 % of the shadow rays that leave a ground
    plane lit from above."""
    exe = str(tmp_path / "plane_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    "-o", exe, os.path.join(ROOT, "tests", "cpp", "plane_check.cpp")], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok ")
