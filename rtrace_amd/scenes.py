"""Seeded procedural scenes for the benchmark configurations.

Every generator is deterministic (SplitMix64 streams) and writes a Scene JSON in
the reference's schema (src/scene.rs:34-446) plus binary STL meshes (f32, as the
reference's loader reads them, src/mesh.rs:919-1003), so the oracle and the HIP
path consume identical vertices.

  sphere_plane     800x600 ortho sphere over a reflective grid plane, point light, quincunx.
  icosphere_scene  ~82k-triangle noisy icosphere + 3 transformed cubes + grid plane,
                   perspective 50 deg, disk area light (16 samples), fog 0.05, depth 4.
  spheres_ring     1000 reflective spheres on a ring over a grid plane, 4 point lights,
                   stochastic 4 spp, depth 10, outline on.
  batch_scenes     64 seeded scenes (displaced icospheres / extruded polygons, 1k-500k
                   triangles, 256x256 to 1920x1080) for the batched launch, 4 auto-camera views each.
  denoise_scene    seeded icosphere-style scenes with a 2x2 square light: the denoiser's
                   training data (low / high light-sample pairs) and benchmark frames.
"""
import json
import math
import os
import struct

import numpy as np

MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014): the seeded stream of every generator here."""

    def __init__(self, seed):
        self.state = seed & MASK64

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo, hi):
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / 9007199254740992.0))


def material(color, ambient=0.1, diffuse=0.7, specular=0.3, shininess=32, reflectivity=None, texture=None):
    m = {"color": color, "ambient": ambient, "diffuse": diffuse, "specular": specular, "shininess": shininess}
    if reflectivity is not None:
        m["reflectivity"] = reflectivity
    if texture is not None:
        m["texture"] = texture
    return m


def hex_color(rgb):
    return "#%02X%02X%02X" % tuple(int(max(0, min(255, round(c * 255)))) for c in rgb)


def write_stl(path, tris):
    """Binary STL of an (n, 3, 3) float array (normal field zero, never read by the path)."""
    tris = np.asarray(tris, dtype=np.float32).reshape(-1, 3, 3)
    rec = np.zeros(len(tris), dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]))
    rec["v"] = tris
    with open(path, "wb") as f:
        f.write(b"rtrace_amd procedural mesh".ljust(80, b" "))
        f.write(struct.pack("<I", len(tris)))
        f.write(rec.tobytes())


_UNIT_ICO = {}


def _unit_icosphere(subdiv):
    """Unit-sphere vertices and faces of the icosahedron subdivided `subdiv` times."""
    if subdiv in _UNIT_ICO:
        return _UNIT_ICO[subdiv]
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
             (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
             (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    _UNIT_ICO[subdiv] = (np.array(verts), np.array(faces, dtype=np.int64))
    return _UNIT_ICO[subdiv]


def icosphere(subdiv, radius, noise, seed):
    """Icosahedron subdivided `subdiv` times (20 * 4^subdiv faces), vertices pushed to
    radius * (1 + U(-noise, noise)) per vertex (shared vertices get one draw)."""
    V, F = _unit_icosphere(subdiv)
    rng = SplitMix64(seed)
    scale = np.array([radius * (1.0 + rng.uniform(-noise, noise)) for _ in range(len(V))])
    return (V * scale[:, None])[F]


def _camera_persp(pos, target, fov, aspect=(16, 9)):
    return {"kind": "perspective", "position": list(pos), "target": list(target), "up": [0, 0, 1],
            "width": aspect[0], "height": aspect[1], "fov": fov}


def _settings(background="#101822", fog=None, outline=None):
    s = {"ambient_illumination": {"color": "#FFFFFF", "intensity": 0.1}, "background_color": background}
    if fog:
        s["fog"] = fog
    if outline:
        s["outline"] = outline
    return s


def sphere_plane(out_dir):
    """800x600 (Renderer::new(800, 600)); the 6x6 ortho view widens to 8x6 (src/camera.rs:79)."""
    sc = {
        "camera": {"kind": "ortho", "position": [0, -10, 5], "target": [0, 0, 0], "up": [0, 0, 1],
                   "width": 6, "height": 6},
        "objects": [
            {"kind": "sphere", "center": [0, 0, 0], "radius": 1.5,
             "material": material("#FF4444", 0.1, 0.8, 0.4, 32)},
            {"kind": "plane", "point": [0, 0, -1.5], "normal": [0, 0, 1],
             "material": material("#CCCCCC", 0.1, 0.7, 0.2, 16, reflectivity=0.3,
                                  texture={"type": "grid", "line_color": "#333333", "line_width": 0.05,
                                           "cell_size": 1.0})},
        ],
        "lights": [{"position": [3, -3, 5], "color": "#FFFFFF", "intensity": 1.0}],
        "scene_settings": _settings("#001122"),
    }
    return _write(out_dir, "sphere_800x600", sc)


def icosphere_scene(out_dir, subdiv=6, light="disk", light_size=1.0, light_samples=16):
    """Procedural mesh scene. subdiv=6 gives 81,920 triangles (~80k).

    light="disk": the reference's area light (diameter `light_size`, 16 rejection samples,
    byte-comparable with the reference); light="square": the extension's square light of
    side `light_size` with `light_samples` jittered-grid samples."""
    os.makedirs(out_dir, exist_ok=True)
    rng = SplitMix64(2)
    tris = icosphere(subdiv, 1.5, 0.05, seed=2)
    mesh_name = f"icosphere_s{subdiv}.stl"
    write_stl(os.path.join(out_dir, mesh_name), tris)
    objects = [{"kind": "mesh", "filename": mesh_name,
                "material": material("#8FB8DE", 0.1, 0.7, 0.4, 48, reflectivity=0.25)}]
    for i in range(3):
        ang = 2 * math.pi * i / 3 + rng.uniform(-0.3, 0.3)
        r = rng.uniform(2.6, 3.4)
        tf = [f"rotate({rng.uniform(0, 90):.4f}, {rng.uniform(0, 90):.4f}, {rng.uniform(0, 90):.4f})",
              f"scale({rng.uniform(0.8, 1.3):.4f}, {rng.uniform(0.8, 1.3):.4f}, {rng.uniform(0.8, 1.3):.4f})",
              f"translate({r * math.cos(ang):.4f}, {r * math.sin(ang):.4f}, {rng.uniform(-0.8, 0.4):.4f})"]
        col = hex_color((rng.uniform(0.3, 1), rng.uniform(0.3, 1), rng.uniform(0.3, 1)))
        objects.append({"kind": "cube", "center": [0, 0, 0], "size": [1, 1, 1], "transform": tf,
                        "material": material(col, 0.1, 0.7, 0.3, 32, reflectivity=0.4)})
    objects.append({"kind": "plane", "point": [0, 0, -1.8], "normal": [0, 0, 1],
                    "material": material("#D8D8D0", 0.1, 0.7, 0.1, 8, reflectivity=0.2,
                                         texture={"type": "grid", "line_color": "#404040", "line_width": 0.04,
                                                  "cell_size": 1.0})})
    if light == "disk":
        lt = {"diameter": light_size}
        name = f"icosphere_s{subdiv}"
    elif light == "square":
        lt = {"shape": "square", "size": light_size, "samples": light_samples}
        name = f"icosphere_s{subdiv}_square{light_size:g}_n{light_samples}"
    else:
        raise ValueError(f"unknown light {light!r}")
    lights = [dict({"position": [3, -4, 6], "color": "#FFFFFF", "intensity": 1.0}, **lt)]
    sc = {"camera": _camera_persp([0, -7.5, 3.5], [0, 0, 0], 50), "objects": objects, "lights": lights,
          "scene_settings": _settings("#202838", fog={"color": "#B0C4DE", "density": 0.05, "start": 2.0,
                                                       "end": 20.0})}
    return _write(out_dir, name, sc)


def spheres_ring(out_dir, n_spheres=1000):
    """1000 reflective spheres on a ring (annulus 14..20 around the origin, resting on
    the plane z = 0), 4 point lights; render with stochastic 4 spp, depth 10."""
    rng = SplitMix64(3)
    objects = []
    for _ in range(n_spheres):
        r = rng.uniform(0.1, 0.6)
        a = rng.uniform(0, 2 * math.pi)
        d = math.sqrt(rng.uniform(14.0 ** 2, 20.0 ** 2))
        col = hex_color((rng.uniform(0.2, 1), rng.uniform(0.2, 1), rng.uniform(0.2, 1)))
        objects.append({"kind": "sphere", "center": [round(d * math.cos(a), 5), round(d * math.sin(a), 5), round(r, 5)],
                        "radius": round(r, 5),
                        "material": material(col, 0.1, 0.7, 0.4, 32, reflectivity=round(rng.uniform(0.2, 0.9), 4))})
    objects.append({"kind": "plane", "point": [0, 0, 0], "normal": [0, 0, 1],
                    "material": material("#C8C8C8", 0.1, 0.7, 0.1, 8,
                                         texture={"type": "grid", "line_color": "#505050", "line_width": 0.05,
                                                  "cell_size": 2.0})})
    lights = [{"position": [x, y, 12], "color": "#FFFFFF", "intensity": 0.45}
              for x, y in [(15, 15), (-15, 15), (-15, -15), (15, -15)]]
    outline = {"enabled": True, "depth_weight": 1.0, "normal_weight": 1.0, "threshold": 0.1, "color": "#000000",
               "thickness": 1.0, "use_8_neighbors": False}
    sc = {"camera": _camera_persp([0, -34, 16], [0, 0, 0], 50), "objects": objects, "lights": lights,
          "scene_settings": _settings("#0A0A14", outline=outline)}
    return _write(out_dir, f"spheres_ring_n{n_spheres}", sc)


def denoise_scene(out_dir, seed, subdiv=None):
    """Icosphere-style scene with a 2x2 square light (the denoiser's data and benchmark):
    a noisy icosphere (subdivision 3-5 by seed unless given), 3 transformed cubes, 0-3
    spheres, a grid plane, random light position, fog on half of the seeds, camera on a
    random side. Render with quincunx, depth 4; the light's sample count is set per render."""
    os.makedirs(out_dir, exist_ok=True)
    rng = SplitMix64(0x5EED0000 + seed)
    sd = 3 + seed % 3 if subdiv is None else subdiv
    radius = rng.uniform(1.2, 1.8)
    mesh_name = f"denoise_{seed}_s{sd}.stl"
    write_stl(os.path.join(out_dir, mesh_name), icosphere(sd, radius, rng.uniform(0.02, 0.08), seed=seed + 17))
    rc = lambda lo=0.3: hex_color((rng.uniform(lo, 1), rng.uniform(lo, 1), rng.uniform(lo, 1)))  # noqa: E731
    objects = [{"kind": "mesh", "filename": mesh_name,
                "material": material(rc(), 0.1, rng.uniform(0.5, 0.8), rng.uniform(0.1, 0.5),
                                     int(rng.uniform(8, 64)), reflectivity=rng.uniform(0.0, 0.3))}]
    for i in range(3):
        ang = 2 * math.pi * i / 3 + rng.uniform(-0.5, 0.5)
        r = rng.uniform(2.4, 3.6)
        tf = [f"rotate({rng.uniform(0, 90):.4f}, {rng.uniform(0, 90):.4f}, {rng.uniform(0, 90):.4f})",
              f"scale({rng.uniform(0.6, 1.4):.4f}, {rng.uniform(0.6, 1.4):.4f}, {rng.uniform(0.6, 1.4):.4f})",
              f"translate({r * math.cos(ang):.4f}, {r * math.sin(ang):.4f}, {rng.uniform(-1.0, 0.4):.4f})"]
        objects.append({"kind": "cube", "center": [0, 0, 0], "size": [1, 1, 1], "transform": tf,
                        "material": material(rc(), 0.1, 0.7, 0.3, 32, reflectivity=rng.uniform(0.0, 0.4))})
    for _ in range(int(rng.uniform(0, 4))):
        ang, r = rng.uniform(0, 2 * math.pi), rng.uniform(1.8, 4.0)
        objects.append({"kind": "sphere", "center": [r * math.cos(ang), r * math.sin(ang), rng.uniform(-1.3, 0.5)],
                        "radius": rng.uniform(0.2, 0.6), "material": material(rc(), 0.1, 0.7, 0.4, 32)})
    plane_z = rng.uniform(-2.0, -1.5)
    objects.append({"kind": "plane", "point": [0, 0, plane_z], "normal": [0, 0, 1],
                    "material": material(rc(0.6), 0.1, 0.7, 0.1, 8, reflectivity=rng.uniform(0.0, 0.3),
                                         texture={"type": "grid", "line_color": "#404040",
                                                  "line_width": rng.uniform(0.02, 0.06),
                                                  "cell_size": rng.uniform(0.6, 1.5)})})
    az, el, dist = rng.uniform(0, 2 * math.pi), rng.uniform(0.6, 1.2), rng.uniform(6.0, 9.0)
    light = {"position": [dist * math.cos(el) * math.cos(az), dist * math.cos(el) * math.sin(az), dist * math.sin(el)],
             "color": "#FFFFFF", "intensity": rng.uniform(0.8, 1.2), "shape": "square", "size": 2.0, "samples": 16}
    cam_az, cam_d = rng.uniform(0, 2 * math.pi), rng.uniform(7.0, 9.0)
    cam = _camera_persp([cam_d * math.cos(cam_az), cam_d * math.sin(cam_az), rng.uniform(2.5, 5.0)], [0, 0, -0.3],
                        50)
    fog = None
    if rng.uniform(0, 1) < 0.5:
        fog = {"color": rc(0.5), "density": rng.uniform(0.02, 0.08), "start": 2.0, "end": 20.0}
    sc = {"camera": cam, "objects": objects, "lights": [light], "scene_settings": _settings(rc(0.0), fog=fog)}
    return _write(out_dir, f"denoise_{seed}_s{sd}", sc)


def extruded_polygon(target_tris, radius, height, seed):
    """A random star-shaped polygon (angles jittered inside their sectors, radii in
    [0.55, 1] x radius) extruded along z into `segments` bands, with fan-triangulated caps:
    2 * n * segments + 2 * n triangles, n and segments chosen to land near target_tris."""
    rng = SplitMix64(seed)
    seg = max(1, int(math.sqrt(target_tris / 16.0)))
    n = max(3, int(round(target_tris / (2.0 * (seg + 1)))))
    k = np.arange(n)
    ang = 2 * math.pi * (k + np.array([rng.uniform(-0.4, 0.4) for _ in range(n)])) / n
    r = radius * np.array([0.55 + 0.45 * rng.uniform(0, 1) for _ in range(n)])
    px, py = r * np.cos(ang), r * np.sin(ang)
    zs = np.linspace(-height / 2, height / 2, seg + 1)
    i0, i1 = k, (k + 1) % n
    z0, z1 = zs[:-1, None], zs[1:, None]
    a = np.stack(np.broadcast_arrays(px[i0], py[i0], z0), -1)  # (seg, n, 3)
    b = np.stack(np.broadcast_arrays(px[i1], py[i1], z0), -1)
    c = np.stack(np.broadcast_arrays(px[i1], py[i1], z1), -1)
    d = np.stack(np.broadcast_arrays(px[i0], py[i0], z1), -1)
    sides = np.concatenate([np.stack([a, b, c], -2), np.stack([a, c, d], -2)], 0).reshape(-1, 3, 3)
    caps = []
    for z in (-height / 2, height / 2):
        ctr = np.broadcast_to(np.array([0.0, 0.0, z]), (n, 3))
        caps.append(np.stack([ctr, np.stack([px[i0], py[i0], np.full(n, z)], -1),
                              np.stack([px[i1], py[i1], np.full(n, z)], -1)], 1))
    return np.concatenate([sides] + caps, 0)


def batch_scenes(out_dir, n_scenes=64, seed=4):
    """The batched workload: `n_scenes` seeded scenes, each a displaced icosphere or an
    extruded random polygon whose triangle count is log-uniform in [1k, 500k] (icospheres
    take the nearest subdivision level), seen with a transform, over a grid plane with
    1-3 spheres and a disk light of 8 shadow samples; each with its own resolution,
    log-uniform in pixel count between 256x256 and 1920x1080 (aspect 1:1, 4:3 or 16:9).
    Returns [(scene path, width, height, triangles)]; render each from its 4 auto-camera
    views with quincunx, depth 4."""
    os.makedirs(out_dir, exist_ok=True)
    rng = SplitMix64(seed)
    out = []
    for i in range(n_scenes):
        target = math.exp(rng.uniform(math.log(1000.0), math.log(500000.0)))
        mesh_name = f"batch_{seed}_{i}.stl"
        if rng.uniform(0, 1) < 0.5:
            sd = int(min(7, max(3, round(math.log(target / 20.0, 4)))))
            tris = icosphere(sd, 1.0, 0.12, seed=1000 * seed + i)
        else:
            tris = extruded_polygon(target, 1.0, rng.uniform(0.4, 1.6), seed=1000 * seed + i)
        write_stl(os.path.join(out_dir, mesh_name), tris)
        tf = [f"scale({rng.uniform(0.8, 1.6):.4f}, {rng.uniform(0.8, 1.6):.4f}, {rng.uniform(0.8, 1.6):.4f})",
              f"rotate({rng.uniform(0, 180):.4f}, {rng.uniform(0, 180):.4f}, {rng.uniform(0, 180):.4f})",
              f"translate(0, 0, {rng.uniform(-0.2, 0.4):.4f})"]
        rc = lambda lo=0.3: hex_color((rng.uniform(lo, 1), rng.uniform(lo, 1), rng.uniform(lo, 1)))  # noqa: E731
        objects = [{"kind": "mesh", "filename": mesh_name, "transform": tf,
                    "material": material(rc(), 0.1, 0.7, 0.4, 32, reflectivity=rng.uniform(0.0, 0.3))}]
        for _ in range(1 + int(rng.uniform(0, 3))):
            ang, rr = rng.uniform(0, 2 * math.pi), rng.uniform(2.0, 3.5)
            objects.append({"kind": "sphere", "center": [rr * math.cos(ang), rr * math.sin(ang), -1.0],
                            "radius": rng.uniform(0.3, 0.6), "material": material(rc(), 0.1, 0.7, 0.5, 48)})
        objects.append({"kind": "plane", "point": [0, 0, -1.6], "normal": [0, 0, 1],
                        "material": material("#C8C8C0", 0.1, 0.7, 0.1, 8, reflectivity=0.2,
                                             texture={"type": "grid", "line_color": "#404040", "line_width": 0.04,
                                                      "cell_size": 1.0})})
        az, el = rng.uniform(0, 2 * math.pi), rng.uniform(0.7, 1.2)
        lights = [{"position": [7 * math.cos(el) * math.cos(az), 7 * math.cos(el) * math.sin(az), 7 * math.sin(el)],
                   "color": "#FFFFFF", "intensity": 1.0, "diameter": rng.uniform(0.8, 2.0), "samples": 8}]
        sc = {"camera": _camera_persp([0, -7, 3], [0, 0, 0], 50), "objects": objects, "lights": lights,
              "scene_settings": _settings(rc(0.0))}
        px = math.exp(rng.uniform(math.log(256 * 256), math.log(1920 * 1080)))
        aspect = (1.0, 4.0 / 3.0, 16.0 / 9.0)[int(rng.uniform(0, 3))]
        w = int(round(math.sqrt(px * aspect)))
        w = max(256, min(1920, w))
        h = max(256, min(1080, int(round(w / aspect))))
        out.append((_write(out_dir, f"batch_{seed}_{i}", sc), w, h, len(tris)))
    return out


def _write(out_dir, name, scene):
    os.makedirs(out_dir, exist_ok=True)
    path = os.path.join(out_dir, name + ".json")
    with open(path, "w") as f:
        json.dump(scene, f, indent=1)
    return path


# Render settings of each configuration (Renderer fields); the 1080p variants are the
# headline "Mrays/s at 1080p" measurement of the same scene.
RENDER = {
    "sphere_800x600": dict(width=800, height=600, aa="quincunx", samples=1, max_depth=10),
    "icosphere_720p": dict(width=1280, height=720, aa="quincunx", samples=1, max_depth=4),
    "icosphere_1080p": dict(width=1920, height=1080, aa="quincunx", samples=1, max_depth=4),
    "icosphere_square_1080p": dict(width=1920, height=1080, aa="quincunx", samples=1, max_depth=4),
    "icosphere_square_720p": dict(width=1280, height=720, aa="quincunx", samples=1, max_depth=4),
    "spheres_ring_1080p": dict(width=1920, height=1080, aa="stochastic", samples=4, max_depth=10),
    "denoise_1080p": dict(width=1920, height=1080, aa="quincunx", samples=1, max_depth=4),
}
