"""Pins for the oracle's area-light sampling, including the light extension.

The reference has one area-light kind: a disk facing the hit point, 16 rejection-sampled
points from StdRng seeded per hit (src/lighting.rs:57-217). The extension keeps that and
adds a square light (a jittered grid of `samples` cells in the same light-facing basis,
jitter from SplitMix64 seeded with the same light seed) and a per-light / per-render
sample count. The extension has no reference counterpart, so it is pinned here by what
the mathematics fixes: the published SplitMix64 sequence, one sample per grid cell in
the plane of the light, equality with the reference disk sampler, a vanishing light
reducing to a point light, and shadow-ray counts linear in the sample count.
"""
import json

import numpy as np
import pytest

from oracle import pyoracle as O

PHI = 0x9E3779B97F4A7C15
M64 = (1 << 64) - 1


def light_seed(path_seed, p):
    """calculate_diffuse_light_contribution seeding (src/lighting.rs:161-165), saturating casts."""
    def sat(x):
        x = x * 1000.0
        if not (x > 0):
            return 0
        return min(int(x), M64)
    s = (path_seed * PHI) & M64
    s = (s + sat(p[0]) * 0x85EBCA6B) & M64
    s = (s + sat(p[1]) * 0xC2B2AE35) & M64
    s = (s + sat(p[2]) * 0x6C8E9CF5) & M64
    return s


def basis(p, lp):
    """Light-facing basis of sample_disk_light_point (src/lighting.rs:74-82)."""
    d = np.asarray(lp, float) - np.asarray(p, float)
    d = d / np.linalg.norm(d)
    up = np.array([1.0, 0, 0]) if abs(d[0]) < 0.9 else np.array([0, 1.0, 0])
    u = np.cross(up, d)
    u /= np.linalg.norm(u)
    v = np.cross(d, u)
    v /= np.linalg.norm(v)
    return d, u, v


def test_splitmix64_published_sequence():
    """Outputs 1..3 of SplitMix64 seeded with 0 (Steele, Lea, Flood 2014)."""
    assert [O.splitmix64_nth(0, k) for k in (1, 2, 3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                                           0x06C45D188009454F]


@pytest.mark.parametrize("n,cols", [(1, 1), (2, 1), (4, 2), (7, 1), (8, 2), (12, 3), (16, 4), (255, 15), (256, 16)])
def test_jitter_grid_columns(n, cols):
    assert O.jitter_cols(n) == cols


@pytest.mark.parametrize("n", [1, 4, 7, 8, 16, 256])
def test_square_samples_one_per_cell_in_light_plane(n):
    rng = np.random.default_rng(n)
    for trial in range(20):
        p = rng.uniform(-5, 5, 3)
        lp = rng.uniform(-5, 5, 3) + np.array([0, 0, 8.0])
        size = float(rng.uniform(0.2, 3.0))
        pts = O.area_samples(int(rng.integers(0, 2**63)), p, lp, True, size, n)
        d, u, v = basis(p, lp)
        rel = pts - lp
        assert np.all(np.abs(rel @ d) <= 1e-12 * (1 + np.abs(lp).max()) * 8)
        x, y = rel @ u, rel @ v
        assert np.all(np.abs(x) <= size / 2 + 1e-12) and np.all(np.abs(y) <= size / 2 + 1e-12)
        cols = O.jitter_cols(n)
        rows = n // cols
        cx = np.floor((x / size + 0.5) * cols).astype(int)
        cy = np.floor((y / size + 0.5) * rows).astype(int)
        s = np.arange(n)
        # a cell boundary within rounding of a jitter of exactly 0 is the only tolerance
        ok_x = (cx == s % cols) | (np.abs((x / size + 0.5) * cols - s % cols) < 1e-9)
        ok_y = (cy == s // cols) | (np.abs((y / size + 0.5) * rows - s // cols) < 1e-9)
        assert ok_x.all() and ok_y.all()


def test_square_jitter_is_splitmix_of_light_seed():
    p, lp, size, n, ps = [0.3, -0.2, 0.1], [1.0, 2.0, 6.0], 1.5, 16, 12345
    pts = O.area_samples(ps, p, lp, True, size, n)
    _, u, v = basis(p, lp)
    ls = light_seed(ps, p)
    x, y = (pts - lp) @ u, (pts - lp) @ v
    for s in range(n):
        jx = (O.splitmix64_nth(ls, 2 * s + 1) >> 11) * 2.0**-53
        jy = (O.splitmix64_nth(ls, 2 * s + 2) >> 11) * 2.0**-53
        assert abs(x[s] - ((s % 4 + jx) / 4 - 0.5) * size) < 1e-12
        assert abs(y[s] - ((s // 4 + jy) / 4 - 0.5) * size) < 1e-12


@pytest.mark.parametrize("n", [16, 8, 3])
def test_disk_samples_are_the_reference_rejection_sampler(n):
    """Disk sample s = light + x u + y v with (x, y) the s-th accepted sample_disk_point
    draw of StdRng seeded with the light seed (the sampler itself is pinned through the
    RFC 8439 and PCG32 vectors in test_oracle.py)."""
    p, lp, diam, ps = [0.5, 0.25, -1.0], [-2.0, 3.0, 7.0], 1.2, 99
    pts = O.area_samples(ps, p, lp, False, diam, n)
    _, u, v = basis(p, lp)
    xy = O.sample_disk(light_seed(ps, p), diam / 2, n)
    want = lp + xy[:, :1] * u + xy[:, 1:] * v
    assert np.abs(pts - want).max() < 1e-12


def _scene(tmp_path, light):
    sc = {
        "camera": {"kind": "perspective", "position": [0, -6, 3], "target": [0, 0, 0], "up": [0, 0, 1],
                   "width": 4, "height": 3, "fov": 50},
        "objects": [
            {"kind": "sphere", "center": [0, 0, 0.2], "radius": 1.0,
             "material": {"color": "#CC4422", "ambient": 0.1, "diffuse": 0.8, "specular": 0.5, "shininess": 24}},
            {"kind": "plane", "point": [0, 0, -1], "normal": [0, 0, 1],
             "material": {"color": "#DDDDDD", "ambient": 0.1, "diffuse": 0.9, "specular": 0.1, "shininess": 8}},
        ],
        "lights": [dict({"position": [2.5, -2.0, 5.0], "color": "#FFFFFF", "intensity": 1.0}, **light)],
        "scene_settings": {"ambient_illumination": {"color": "#FFFFFF", "intensity": 0.15}},
    }
    path = tmp_path / ("s%d.json" % abs(hash(json.dumps(light))))
    path.write_text(json.dumps(sc))
    return str(path)


@pytest.mark.parametrize("light", [{"shape": "square", "size": 1e-7, "samples": 8},
                                   {"diameter": 1e-7, "samples": 5},
                                   {"shape": "square", "size": 1e-7}])
def test_vanishing_area_light_is_the_point_light(tmp_path, light):
    """As the light shrinks every sample sits on the point light and is visible exactly
    when it is, so (sum/n) * (visible/n) becomes the point-light term."""
    opts = O.make_opts(96, 72, threads=4)
    a, _, _ = O.render(_scene(tmp_path, {}), opts)
    b, _, _ = O.render(_scene(tmp_path, light), opts)
    d = np.abs(a.astype(int) - b.astype(int)).max(-1)
    assert (d <= 1).mean() >= 0.999, d.max()


def test_shadow_ray_count_is_linear_in_samples(tmp_path):
    path = _scene(tmp_path, {"shape": "square", "size": 1.0})
    r = {}
    for n in (3, 4, 5, 16):
        _, st, _ = O.render(path, O.make_opts(40, 30, threads=4, light_samples=n))
        r[n] = st["rays_shadow"]
    per_hit = r[4] - r[3]
    assert per_hit > 0 and r[5] - r[4] == per_hit and r[16] == r[3] + 13 * per_hit
    _, st, _ = O.render(path, O.make_opts(40, 30, threads=4))
    assert st["rays_shadow"] == r[16]  # default count is the reference's 16


@pytest.mark.parametrize("light,msg", [({"shape": "triangle", "size": 1.0}, "unknown variant"),
                                       ({"shape": "square"}, "missing field `size`"),
                                       ({"diameter": 1.0, "samples": 0}, "samples")])
def test_light_extension_errors(tmp_path, light, msg):
    with pytest.raises(RuntimeError, match=msg):
        O.render(_scene(tmp_path, light), O.make_opts(8, 6, threads=1))
