"""The Node addon (rtrace_amd/bin/rtrace_hip.node): the reference binding's exports
(index.d.ts:7-21) and, on a GPU, renders equal to the reference's own renders of the same
scenes at the same sizing (bindings/node/src/lib.rs:31-45: H = D / sqrt(R^2 + 1), W = R H,
both rounded)."""
import json
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import REF_IMAGES, REF_SCENES, ROOT

ADDON = os.path.join(ROOT, "rtrace_amd", "bin", "rtrace_hip.node")
NODE = shutil.which("node")
pytestmark = pytest.mark.skipif(not NODE or not os.path.exists(ADDON), reason="node or the addon is missing")

EXPORTS = ["helloWorld", "greetWithName", "renderScene", "renderSceneThreaded", "renderSceneBruteForce",
           "renderSceneFromFile", "renderSceneFromFileThreaded", "renderSceneFromFileBruteForce"]


def node(script):
    r = subprocess.run([NODE, "-e", f"const m = require({json.dumps(ADDON)});\n" + script], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return r.stdout.strip()


def test_exports_and_greetings():
    out = json.loads(node("console.log(JSON.stringify([Object.keys(m), m.helloWorld(), m.greetWithName('rtrace')]))"))
    assert out == [EXPORTS, "hello world", "hello world, rtrace"]


def test_bad_arguments_throw_invalid_arg():
    out = node("try { m.renderSceneFromFile(1, 2) } catch (e) { console.log(e.code) }")
    assert out == "InvalidArg"


def _size(path, diag):
    with open(path) as f:
        cam = json.load(f)["camera"]
    r = cam["width"] / cam["height"]
    h = diag / math.sqrt(r * r + 1.0)
    return int(round(r * h)), int(round(h))


def _gold(name):
    from PIL import Image
    return np.asarray(Image.open(os.path.join(REF_IMAGES, name + ".png")).convert("RGB"))


@pytest.mark.gpu
@pytest.mark.parametrize("fn,suffix", [("renderSceneFromFile", " (multi-threaded)"),
                                       ("renderSceneFromFileBruteForce", " (brute force)"),
                                       ("renderSceneFromFileThreaded", " with 3 threads")])
def test_file_render_equals_reference(tmp_path, fn, suffix):
    """The binding renders with Renderer::new (quincunx, depth 10, seed 0) at the CLI's sizing,
    so at diagonal 500 its k-d renders are the reference's own doc render of the scene
    (object-mesh.png, CLI -s 500); the brute-force one is checked against the oracle's
    brute-force render (src/ray.rs:650-671)."""
    from PIL import Image
    from oracle import pyoracle as O
    scene = os.path.join(REF_SCENES, "object-mesh.json")
    out = str(tmp_path / "o.png")
    extra = ", 3" if "Threaded" in fn else ""
    msg = node(f"console.log(m.{fn}({json.dumps(scene)}, {json.dumps(out)}, 500{extra}))")
    W, H = _size(scene, 500)
    assert msg == f"Successfully rendered {W}×{H} image (diagonal 500) to '{out}'" + suffix
    img = np.asarray(Image.open(out).convert("RGB"))
    if "Brute" in fn:
        ref, _, _ = O.render(scene, O.make_opts(W, H, use_kdtree=False))
    else:
        ref = _gold("object-mesh")
    assert img.shape == ref.shape and np.array_equal(img, ref)


@pytest.mark.gpu
def test_json_render_and_parse_error(tmp_path):
    """renderScene from JSON text at diagonal 500 is the reference's lighting-multiple.png."""
    from PIL import Image
    scene = os.path.join(REF_SCENES, "lighting-multiple.json")
    text = open(scene).read()
    out = str(tmp_path / "j.png")
    msg = node(f"console.log(m.renderScene({json.dumps(text)}, {json.dumps(out)}, 500))")
    W, H = _size(scene, 500)
    assert msg == f"Successfully rendered {W}×{H} image (diagonal 500) to '{out}' (multi-threaded)"
    img = np.asarray(Image.open(out).convert("RGB"))
    ref = _gold("lighting-multiple")
    assert img.shape == ref.shape and np.array_equal(img, ref)
    err = node("try { m.renderScene('{', 'x.png') } catch (e) { console.log(e.code + '|' + e.message) }")
    assert err.startswith("InvalidArg|Failed to parse scene JSON: ")
