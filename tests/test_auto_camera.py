"""AutoCamera (src/auto_camera.rs:9-130) on the host: the reference's own unit tests
(src/auto_camera.rs:143-260), the library against the oracle's restatement on every
reference scene, and the closed forms the views must satisfy."""
import glob
import json
import math
import os

import numpy as np
import pytest

import rtrace_amd as R
from conftest import REF_SCENES
from oracle import pyoracle as O

SETTINGS = {"ambient_illumination": {"color": "#FFFFFF", "intensity": 0.1}}
MAT = {"color": "#FF0000", "ambient": 0.1, "diffuse": 0.7, "specular": 0.2, "shininess": 10}


def _scene(tmp_path, objects, name="s.json"):
    sc = {"camera": {"kind": "ortho", "position": [0, -5, 0], "target": [0, 0, 0], "up": [0, 0, 1], "width": 4,
                     "height": 4},
          "objects": objects, "lights": [{"position": [2, 2, 2], "color": "#FFFFFF", "intensity": 1.0}],
          "scene_settings": SETTINGS}
    p = tmp_path / name
    p.write_text(json.dumps(sc))
    return str(p)


def test_sphere_reference_unit_test(tmp_path):
    r = R.AutoCamera.generate_cameras(_scene(tmp_path, [{"kind": "sphere", "center": [0, 0, 0], "radius": 1.0,
                                                         "material": MAT}]))
    for c in r.cameras():
        assert list(c.target) == [0.0, 0.0, 0.0]
    assert r.left.position[1] < 0 and r.front.position[0] < 0 and r.top.position[2] > 0
    assert all(v > 0 for v in r.perspective.position)
    assert r.perspective.has_fov and r.perspective.fov == 50.0
    assert r.to_cameras_json()["left"]["kind"] == "ortho" and r.to_cameras_json()["perspective"]["kind"] == "perspective"


def test_cube_reference_unit_test(tmp_path):
    r = R.AutoCamera.generate_cameras(_scene(tmp_path, [{"kind": "cube", "center": [1, 1, 1], "size": [2, 2, 2],
                                                         "material": MAT}]))
    for c in r.cameras():
        assert list(c.target) == [1.0, 1.0, 1.0]


@pytest.mark.parametrize("objects", [[], [{"kind": "plane", "point": [0, 0, 0], "normal": [0, 0, 1], "material": MAT}]],
                         ids=["empty", "only-planes"])
def test_no_finite_objects_is_an_error(tmp_path, objects):
    with pytest.raises(R.RtraceError, match="no finite objects"):
        R.AutoCamera.generate_cameras(_scene(tmp_path, objects))


def _oracle_cams(path):
    import ctypes
    L = O.lib()
    L.oracle_auto_cameras.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
    out = np.zeros(4 * 14)
    err = ctypes.create_string_buffer(256)
    rc = L.oracle_auto_cameras(path.encode(), out.ctypes.data, err, 256)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return out.reshape(4, 14)


def _as_row(c):
    return [c.kind] + list(c.position) + list(c.target) + list(c.up) + [c.width, c.height,
                                                                        c.fov if c.has_fov else 0.0, 0.0]


def _finite_scenes():
    out = []
    for p in sorted(glob.glob(os.path.join(REF_SCENES, "*.json"))):
        with open(p) as f:
            sc = json.load(f)
        if any(o["kind"] != "plane" for o in sc["objects"]):
            out.append(p)
    return out


@pytest.mark.parametrize("path", _finite_scenes(), ids=lambda p: os.path.basename(p))
def test_library_equals_oracle_on_reference_scenes(path):
    got = np.array([_as_row(c) for c in R.AutoCamera.generate_cameras(path).cameras()])
    assert np.array_equal(got, _oracle_cams(path))


def test_transformed_objects_use_the_reference_bounds(tmp_path):
    """Rotated cube and mesh: the cube's box is center + M (corner, 1) (src/ray.rs:228-262),
    the mesh's the box of its untransformed box's corners through M (src/scene.rs:599-648)."""
    import shutil
    shutil.copy(os.path.join(REF_SCENES, "icosahedron.stl"), tmp_path / "ico.stl")
    objs = [{"kind": "cube", "center": [0.5, 0, 0], "size": [1, 2, 1], "transform": ["rotate(0, 0, 30)", "translate(1, 0, 0)"],
             "material": MAT},
            {"kind": "mesh", "filename": "ico.stl", "transform": ["rotate(45, 0, 0)", "scale(2, 1, 1)"],
             "material": MAT}]
    path = _scene(tmp_path, objs)
    got = np.array([_as_row(c) for c in R.AutoCamera.generate_cameras(path).cameras()])
    assert np.array_equal(got, _oracle_cams(path))


def test_views_frame_the_bounds_with_the_stated_margin(tmp_path):
    """Closed forms: ortho viewport = 1.15 x the largest extent at twice that distance; the
    perspective camera sits max_dim / tan(25 deg) from the centre, 35 deg above the x-y
    plane, at 45 deg in x-y."""
    r = R.AutoCamera.generate_cameras(_scene(tmp_path, [
        {"kind": "sphere", "center": [1, 2, 3], "radius": 0.5, "material": MAT},
        {"kind": "cube", "center": [-1, 0, 0], "size": [1, 3, 1], "material": MAT}]))
    lo, hi = np.array([-1.5, -1.5, -0.5]), np.array([1.5, 2.5, 3.5])
    c, vw = (lo + hi) / 2, (hi - lo).max() * 1.15
    assert np.allclose(r.left.target, c) and math.isclose(r.left.width, vw) and math.isclose(r.top.height, vw)
    assert np.allclose(np.array(r.left.position) - c, [0, -2 * vw, 0])
    assert np.allclose(np.array(r.front.position) - c, [-2 * vw, 0, 0])
    assert np.allclose(np.array(r.top.position) - c, [0, 0, 2 * vw]) and list(r.top.up) == [0, 1, 0]
    d = np.array(r.perspective.position) - c
    assert math.isclose(np.linalg.norm(d), vw / math.tan(math.radians(25)), rel_tol=1e-12)
    assert math.isclose(math.degrees(math.asin(d[2] / np.linalg.norm(d))), 35.0, rel_tol=1e-12)
    assert math.isclose(d[0], d[1], rel_tol=1e-12)
