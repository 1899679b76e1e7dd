"""The C++ CLI (rtrace_amd/bin/rtrace) as a product: the reference CLI's validation order,
messages and exit codes (cli/src/main.rs:36-141, clap's u32 parsing of --size and
--samples), and on a GPU the PNGs it writes for the reference's own doc scenes, which the
reference rendered with the same CLI at -s 500 (tests/golden/ref/images)."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import REF_IMAGES, REF_SCENES, ROOT
from test_oracle import DOC_CASES, scene_path

CLI = os.path.join(ROOT, "rtrace_amd", "bin", "rtrace")
pytestmark = pytest.mark.skipif(not os.path.exists(CLI), reason="the CLI is not built")
SPHERE = os.path.join(REF_SCENES, "object-sphere.json")


def run(*args):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=300)


# ---------------------------------------------------------------- no device needed
def test_missing_input_file(tmp_path):
    r = run("-i", str(tmp_path / "nope.json"), "-o", str(tmp_path / "o.png"), "--backend", "hip")
    assert r.returncode == 1 and r.stderr == f"Error: Input file '{tmp_path / 'nope.json'}' does not exist\n"


def test_bad_anti_aliasing_mode(tmp_path):
    r = run("-i", SPHERE, "-o", str(tmp_path / "o.png"), "--anti-aliasing", "supersample", "--backend", "hip")
    assert r.returncode == 1
    assert r.stderr == ("Error: Invalid anti-aliasing mode 'supersample'. Valid options are: quincunx, stochastic, "
                        "no-jitter\n")


def test_zero_samples(tmp_path):
    r = run("-i", SPHERE, "-o", str(tmp_path / "o.png"), "--samples", "0", "--backend", "hip")
    assert r.returncode == 1 and r.stderr == "Error: Samples must be greater than 0\n"


@pytest.mark.parametrize("flag,value,why", [("-s", "-5", "invalid digit found in string"),
                                            ("--size", "4294967296", "number too large to fit in target type"),
                                            ("--samples", "-1", "invalid digit found in string"),
                                            ("--samples", "", "cannot parse integer from empty string"),
                                            ("--max-depth", "x", "invalid digit found in string")])
def test_unsigned_flags_parse_as_clap_does(tmp_path, flag, value, why):
    r = run("-i", SPHERE, "-o", str(tmp_path / "o.png"), flag, value)
    assert r.returncode == 2
    name = {"-s": "--size <SIZE>", "--size": "--size <SIZE>", "--samples": "--samples <SAMPLES>",
            "--max-depth": "--max-depth <MAX_DEPTH>"}[flag]
    assert r.stderr.startswith(f"error: invalid value '{value}' for '{name}': {why}\n")


def test_default_backend_is_the_reference_renderer(tmp_path):
    r = run("-i", SPHERE, "-o", str(tmp_path / "o.png"))
    assert r.returncode == 1 and "use --backend hip" in r.stderr and not os.path.exists(tmp_path / "o.png")
    r = run("-i", SPHERE, "-o", str(tmp_path / "o.png"), "--backend", "cuda")
    assert r.returncode == 1 and r.stderr == "Error: Invalid backend 'cuda'. Valid options are: cpu, hip\n"


def test_required_arguments_and_help():
    r = run("-i", SPHERE)
    assert r.returncode == 2 and "--output <OUTPUT>" in r.stderr
    r = run("--help")
    assert r.returncode == 0 and "--backend" in r.stdout


# ---------------------------------------------------------------- renders on the GPU
def _expected_stdout(path, kw, out, W, H):
    with open(path) as f:
        sc = json.load(f)
    cam = sc["camera"]
    aa = kw.get("anti_aliasing", "quincunx")
    lines = [f"Loaded scene with {len(sc['objects'])} objects and {len(sc.get('lights', []))} lights",
             f"Using camera aspect ratio {cam['width'] / cam['height']:.3f} to compute {W}×{H} pixels from diagonal 500"]
    if (sc.get("scene_settings") or {}).get("outline") is not None:
        lines.append("Outline detection enabled from scene configuration")
        if aa == "quincunx":
            lines.append("Warning: Quincunx anti-aliasing is not compatible with outline detection. Switching to "
                         "no-jitter mode.")
            aa = "no-jitter"
    lines.append(f"Rendering {W}×{H} image (diagonal 500) with {aa} anti-aliasing ({kw.get('samples', 1)} samples)...")
    return lines


@pytest.mark.gpu
@pytest.mark.parametrize("name,scene,kw", DOC_CASES, ids=[c[0] for c in DOC_CASES])
def test_cli_writes_the_reference_doc_render(name, scene, kw, tmp_path):
    from PIL import Image

    path = scene_path(scene, str(tmp_path))
    out = str(tmp_path / f"{name}.png")
    args = ["-i", path, "-o", out, "-s", "500", "--backend", "hip"]
    if "samples" in kw:
        args += ["--samples", str(kw["samples"])]
    if "anti_aliasing" in kw:
        args += ["--anti-aliasing", kw["anti_aliasing"]]
    r = run(*args)
    assert r.returncode == 0, r.stderr
    gold = np.asarray(Image.open(os.path.join(REF_IMAGES, name + ".png")).convert("RGB"))
    img = np.asarray(Image.open(out).convert("RGB"))
    H, W = gold.shape[:2]
    lines = r.stdout.splitlines()
    assert lines[:len(_expected_stdout(path, kw, out, W, H))] == _expected_stdout(path, kw, out, W, H)
    assert lines[-2:] == [f"Image saved to: {out}", f"Successfully rendered to '{out}'"]
    assert lines[-3].startswith("Total rendering time: ")
    assert img.shape == gold.shape and np.array_equal(img, gold), \
        f"{np.count_nonzero((img != gold).any(-1))} pixels differ"
