"""The device k-d build (rt_kd_build with on_device = 1, kernels/kdbuild_gpu.hip) against the
host build (KdTree::new / build_recursive, src/mesh.rs:75-197, restated in
host/kdbuild.cpp): node array, triangle records, f64 node bounds, roots and leaf statistics
byte for byte, on the reference's STL files and on generated meshes that exercise the
build's rules: ties in the centroid sort (duplicated triangles), the empty-side rule,
depth-16 leaves with hundreds of copied triangles (fan-triangulated caps), zero-area and
signed-zero coordinates, and meshes at and around the 10-triangle leaf size."""
import os

import numpy as np
import pytest

from conftest import REF_SCENES

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def R():
    import rtrace_amd as R
    return R


def _same(R, tris, label):
    h = R.kd_build(tris, on_device=False)
    d = R.kd_build(tris, on_device=True)
    for k in ("root_kd", "root_bf", "leaves", "max_leaf"):
        assert h[k] == d[k], (label, k, h[k], d[k])
    for k in ("nodes", "recs", "bounds"):
        assert h[k].shape == d[k].shape, (label, k, h[k].shape, d[k].shape)
        hb = np.ascontiguousarray(h[k]).view(np.uint8).reshape(len(h[k]), 48)
        db = np.ascontiguousarray(d[k]).view(np.uint8).reshape(len(d[k]), 48)
        bad = np.nonzero(np.any(hb != db, axis=1))[0]
        assert len(bad) == 0, (label, k, "first differing entry", int(bad[0]) if len(bad) else None)
    return h


@pytest.mark.parametrize("name", ["plus.stl", "sphere.stl", "icosahedron.stl"])
def test_reference_meshes(R, name):
    tris = R.load_stl(os.path.join(REF_SCENES, name))
    _same(R, tris, name)


@pytest.mark.parametrize("sd", [4, 6])
def test_noisy_icospheres(R, sd):
    from rtrace_amd import scenes
    tris = np.asarray(scenes.icosphere(sd, 1.5, 0.05, seed=7), np.float64)
    h = _same(R, tris, f"icosphere {sd}")
    assert h["leaves"] > 100


def test_extruded_polygon_fan_caps(R):
    from rtrace_amd import scenes
    tris = np.asarray(scenes.extruded_polygon(120000, 1.0, 0.8, seed=11), np.float64)
    h = _same(R, tris, "extruded polygon")
    assert h["max_leaf"] > 100  # the caps' slivers pile up in leaves at depth 16


def test_ties_and_degenerate_triangles(R):
    rng = np.random.default_rng(3)
    base = rng.uniform(-1, 1, (4000, 3, 3))
    tris = np.concatenate([base, base[::2], base[:500]])  # duplicates: equal centroids, stable order
    tris[::7, 2] = tris[::7, 0]  # zero-area triangles
    tris[::11, :, 2] = 0.0
    tris[1::11, :, 2] = -0.0  # signed zeros on the split axis
    _same(R, tris.copy(), "ties")


def test_flat_grid(R):
    # an axis-aligned grid: many triangles share coordinates and centroids on two axes
    xs, ys = np.meshgrid(np.arange(120.0), np.arange(120.0))
    a = np.stack([xs, ys, np.zeros_like(xs)], -1).reshape(-1, 3)
    b = a + [1, 0, 0]
    c = a + [0, 1, 0]
    d = a + [1, 1, 0]
    tris = np.concatenate([np.stack([a, b, c], 1), np.stack([b, d, c], 1)])
    _same(R, tris, "flat grid")


@pytest.mark.parametrize("n", [0, 5, 10, 11, 12, 23, 1000])
def test_small_meshes(R, n):
    rng = np.random.default_rng(n)
    _same(R, rng.normal(size=(n, 3, 3)), f"n={n}")


def test_triangle_soup(R):
    rng = np.random.default_rng(9)
    c = rng.uniform(-5, 5, (60000, 1, 3))
    _same(R, c + rng.normal(scale=0.05, size=(60000, 3, 3)), "soup")
