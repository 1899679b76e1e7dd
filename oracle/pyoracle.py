"""ctypes front-end to the CPU f64 oracle (librt_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the rtrace_amd product path.

`cli_size` and `cli_options` restate the CLI glue of the reference
(cli/src/main.rs:79-124) so that a test renders a scene exactly the way
`rtrace -i scene.json -s N` does.
"""
import ctypes
import json
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librt_oracle.so")

AA_NO_JITTER, AA_QUINCUNX, AA_STOCHASTIC = 0, 1, 2
AA_BY_NAME = {"no-jitter": AA_NO_JITTER, "quincunx": AA_QUINCUNX, "stochastic": AA_STOCHASTIC}


class OracleOpts(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int32), ("height", ctypes.c_int32), ("max_depth", ctypes.c_int32),
        ("use_kdtree", ctypes.c_int32), ("threads", ctypes.c_int32), ("samples", ctypes.c_int32),
        ("aa_mode", ctypes.c_int32), ("has_seed", ctypes.c_int32), ("seed", ctypes.c_uint64),
        ("outline_enabled", ctypes.c_int32), ("outline_use_8_neighbors", ctypes.c_int32),
        ("outline_depth_weight", ctypes.c_double), ("outline_normal_weight", ctypes.c_double),
        ("outline_threshold", ctypes.c_double), ("outline_color", ctypes.c_double * 3),
        ("outline_thickness", ctypes.c_double), ("light_samples", ctypes.c_int32),
    ]


class OracleStats(ctypes.Structure):
    _fields_ = [
        ("rays_primary", ctypes.c_uint64), ("rays_reflect", ctypes.c_uint64), ("rays_shadow", ctypes.c_uint64),
        ("kd_inner_visits", ctypes.c_uint64), ("kd_leaf_visits", ctypes.c_uint64), ("tri_tests", ctypes.c_uint64),
        ("sphere_tests", ctypes.c_uint64), ("render_seconds", ctypes.c_double), ("threads_used", ctypes.c_int32),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_render.restype = ctypes.c_int
        L.oracle_render.argtypes = [ctypes.c_char_p, ctypes.POINTER(OracleOpts), ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.POINTER(OracleStats), ctypes.c_char_p, ctypes.c_int]
        L.oracle_render_aux.restype = ctypes.c_int
        L.oracle_render_aux.argtypes = [ctypes.c_char_p, ctypes.POINTER(OracleOpts), ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.POINTER(OracleStats), ctypes.c_char_p, ctypes.c_int]
        L.oracle_chacha_block.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        L.oracle_seed_key.argtypes = [ctypes.c_uint64, ctypes.c_void_p]
        L.oracle_stdrng_u64.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p]
        L.oracle_stdrng_f64.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p]
        L.oracle_stdrng_construction.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_pcg32_f64.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_int, ctypes.c_void_p]
        L.oracle_sample_disk.argtypes = [ctypes.c_uint64, ctypes.c_double, ctypes.c_int, ctypes.c_void_p]
        L.oracle_area_samples.argtypes = [ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                          ctypes.c_double, ctypes.c_int, ctypes.c_void_p]
        L.oracle_splitmix64_nth.restype = ctypes.c_uint64
        L.oracle_splitmix64_nth.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.oracle_jitter_cols.restype = ctypes.c_int
        L.oracle_jitter_cols.argtypes = [ctypes.c_int]
        L.oracle_parse_transforms.argtypes = [ctypes.c_char_p, ctypes.c_void_p]
        L.oracle_hex_to_color.argtypes = [ctypes.c_char_p, ctypes.c_void_p]
        L.oracle_kd_stats.argtypes = [ctypes.c_char_p] + [ctypes.POINTER(ctypes.c_int32)] * 4 + [ctypes.c_char_p, ctypes.c_int]
        L.oracle_kd_tree.argtypes = [ctypes.c_void_p, ctypes.c_int64] + [ctypes.c_void_p] * 7 + [ctypes.c_int64,
                                     ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64),
                                     ctypes.POINTER(ctypes.c_int64)]
        L.oracle_mesh_hit.restype = ctypes.c_int
        L.oracle_mesh_hit.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
        L.oracle_load_stl.restype = ctypes.c_int
        L.oracle_load_stl.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int64,
                                      ctypes.POINTER(ctypes.c_int64), ctypes.c_char_p, ctypes.c_int]
        L.oracle_outline.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.POINTER(OracleOpts), ctypes.c_void_p]
        _lib = L
    return _lib


def load_stl(path):
    """Mesh::from_stl_file (src/mesh.rs:751-1003) -> (n, 3, 3) float64 triangles; raises
    ValueError carrying the reference's error message."""
    p = os.fspath(path).encode()
    n = ctypes.c_int64(0)
    err = ctypes.create_string_buffer(512)
    if lib().oracle_load_stl(p, None, 0, ctypes.byref(n), err, 512) != 0:
        raise ValueError(err.value.decode())
    out = np.zeros((n.value, 3, 3), np.float64)
    if n.value and lib().oracle_load_stl(p, out.ctypes.data, n.value, ctypes.byref(n), err, 512) != 0:
        raise ValueError(err.value.decode())
    return out


def rust_round(x):
    """f64::round: half away from zero."""
    f = math.floor(abs(x))
    r = f + (1.0 if abs(x) - f >= 0.5 else 0.0)
    return math.copysign(r, x)


def cli_size(scene, diagonal):
    """cli/src/main.rs:79-89 (same arithmetic as bindings/node/src/lib.rs:34-44)."""
    cam = scene["camera"]
    r = cam["width"] / cam["height"]
    h = diagonal / math.sqrt(r * r + 1.0)
    w = r * h
    return int(rust_round(w)), int(rust_round(h))


def outline_from_scene(scene):
    """Scene::get_outline_config (src/scene.rs:515-534) with serde defaults (:384-403)."""
    o = (scene.get("scene_settings") or {}).get("outline")
    if not o or not o.get("enabled"):
        return None
    col = o.get("color", "#000000")
    c = hex_to_color(col)
    if c is None:
        raise ValueError("Invalid hex color format")
    return dict(depth_weight=o.get("depth_weight", 1.0), normal_weight=o.get("normal_weight", 1.0),
                threshold=o.get("threshold", 0.1), color=c, thickness=o.get("thickness", 1.0),
                use_8_neighbors=bool(o.get("use_8_neighbors", False)))


def hex_to_color(s):
    out = (ctypes.c_double * 3)()
    if lib().oracle_hex_to_color(s.encode(), out) != 0:
        return None
    return tuple(out)


def make_opts(width, height, max_depth=10, use_kdtree=True, threads=0, samples=1, aa="quincunx",
              seed=0, outline=None, light_samples=0):
    o = OracleOpts()
    o.light_samples = light_samples
    o.width, o.height, o.max_depth = width, height, max_depth
    o.use_kdtree, o.threads, o.samples = int(use_kdtree), threads, samples
    o.aa_mode = AA_BY_NAME[aa] if isinstance(aa, str) else aa
    o.has_seed = 0 if seed is None else 1
    o.seed = 0 if seed is None else seed
    if outline:
        o.outline_enabled = 1
        o.outline_use_8_neighbors = int(outline["use_8_neighbors"])
        o.outline_depth_weight = outline["depth_weight"]
        o.outline_normal_weight = outline["normal_weight"]
        o.outline_threshold = outline["threshold"]
        for i in range(3):
            o.outline_color[i] = outline["color"][i]
        o.outline_thickness = outline["thickness"]
    return o


def cli_options(scene_path, size=1000, max_depth=10, samples=None, anti_aliasing="quincunx", threads=0):
    """The Renderer the reference CLI builds for `-i scene -s size ...` (cli/src/main.rs:55-124)."""
    with open(scene_path) as f:
        scene = json.load(f)
    w, h = cli_size(scene, size)
    outline = outline_from_scene(scene)
    aa = anti_aliasing
    if outline is not None and aa == "quincunx":
        aa = "no-jitter"
    return make_opts(w, h, max_depth=max_depth, samples=samples or 1, aa=aa, seed=0, outline=outline,
                     threads=threads)


def render(scene_path, opts, want_hits=False):
    """Returns (rgb HxWx3 uint8, stats dict, hits or None)."""
    L = lib()
    W, H = opts.width, opts.height
    rgb = np.zeros((H, W, 3), dtype=np.uint8)
    n_prim = W * H + (W + 1) * (H + 1) if opts.aa_mode == AA_QUINCUNX else W * H * max(opts.samples, 1)
    hits = None
    ptrs = [None, None, None]
    if want_hits:
        hits = (np.zeros(n_prim, np.int32), np.zeros(n_prim, np.int32), np.zeros(n_prim, np.float64))
        ptrs = [h.ctypes.data for h in hits]
    st = OracleStats()
    err = ctypes.create_string_buffer(512)
    rc = L.oracle_render(scene_path.encode(), ctypes.byref(opts), rgb.ctypes.data, ptrs[0], ptrs[1], ptrs[2],
                         ctypes.byref(st), err, 512)
    if rc != 0:
        raise RuntimeError(f"oracle_render failed ({rc}): {err.value.decode()}")
    stats = {k: getattr(st, k) for k, _ in OracleStats._fields_}
    return rgb, stats, hits


FEAT_COUNT = 13


def render_aux(scene_path, opts):
    """Returns (rgb HxWx3 uint8, features HxWx13 f32, radiance HxWx3 f32, stats dict)."""
    W, H = opts.width, opts.height
    rgb = np.zeros((H, W, 3), dtype=np.uint8)
    feat = np.zeros((H, W, FEAT_COUNT), dtype=np.float32)
    rad = np.zeros((H, W, 3), dtype=np.float32)
    st = OracleStats()
    err = ctypes.create_string_buffer(512)
    rc = lib().oracle_render_aux(scene_path.encode(), ctypes.byref(opts), rgb.ctypes.data, feat.ctypes.data,
                                 rad.ctypes.data, ctypes.byref(st), err, 512)
    if rc != 0:
        raise RuntimeError(f"oracle_render_aux failed ({rc}): {err.value.decode()}")
    return rgb, feat, rad, {k: getattr(st, k) for k, _ in OracleStats._fields_}


def chacha_block(words16, double_rounds):
    a = np.asarray(words16, dtype=np.uint32)
    out = np.zeros(16, dtype=np.uint32)
    lib().oracle_chacha_block(a.ctypes.data, double_rounds, out.ctypes.data)
    return out


def seed_key(seed):
    out = np.zeros(8, dtype=np.uint32)
    lib().oracle_seed_key(seed, out.ctypes.data)
    return out


def stdrng_u64(seed, n):
    out = np.zeros(n, dtype=np.uint64)
    lib().oracle_stdrng_u64(seed, n, out.ctypes.data)
    return out


def stdrng_f64(seed, n):
    out = np.zeros(n, dtype=np.float64)
    lib().oracle_stdrng_f64(seed, n, out.ctypes.data)
    return out


def stdrng_construction(seed32):
    """Two draws of rand's StdRng construction sequence (from_seed, next_u64, from_rng, next_u64)."""
    a = np.frombuffer(bytes(seed32), dtype=np.uint8).copy()
    assert a.size == 32
    out = np.zeros(2, dtype=np.uint64)
    lib().oracle_stdrng_construction(a.ctypes.data, out.ctypes.data)
    return out


def pcg32_f64(state, stream, n, low=None, high=None):
    """n f64 draws from Pcg32::new(state, stream): gen::<f64>() without bounds, else gen_range(low..high)."""
    out = np.zeros(n, dtype=np.float64)
    kind = 0 if low is None else 1
    lib().oracle_pcg32_f64(state, stream, kind, 0.0 if low is None else low, 0.0 if high is None else high,
                           n, out.ctypes.data)
    return out


def sample_disk(seed, radius, n):
    out = np.zeros(2 * n, dtype=np.float64)
    lib().oracle_sample_disk(seed, radius, n, out.ctypes.data)
    return out.reshape(n, 2)


def area_samples(path_seed, p, light_pos, square, extent, n):
    """n shadow-sample points (n x 3) of a disk (square=False, extent=diameter) or
    square (extent=side) light for a hit at p."""
    out = np.zeros(3 * n, dtype=np.float64)
    pa = np.asarray(p, dtype=np.float64)
    la = np.asarray(light_pos, dtype=np.float64)
    lib().oracle_area_samples(path_seed, pa.ctypes.data, la.ctypes.data, int(square), extent, n, out.ctypes.data)
    return out.reshape(n, 3)


def splitmix64_nth(seed, k):
    return lib().oracle_splitmix64_nth(seed, k)


def jitter_cols(n):
    return lib().oracle_jitter_cols(n)


def parse_transforms(strings):
    out = np.zeros(16, dtype=np.float64)
    if lib().oracle_parse_transforms("\n".join(strings).encode(), out.ctypes.data) != 0:
        return None
    return out.reshape(4, 4).T  # column-major storage -> row/col matrix


def kd_stats(stl_path):
    vals = [ctypes.c_int32() for _ in range(4)]
    err = ctypes.create_string_buffer(256)
    rc = lib().oracle_kd_stats(stl_path.encode(), *[ctypes.byref(v) for v in vals], err, 256)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return dict(zip(["tris", "leaves", "max_leaf", "nodes"], [v.value for v in vals]))


def kd_tree(triangles):
    """KdTree::new (src/mesh.rs:75-197) of an (n, 3, 3) float64 triangle array: nodes in
    depth-first pre-order with leaf flag, axis, split, right child, leaf triangle lists
    and f64 bounds."""
    v = np.ascontiguousarray(triangles, np.float64).reshape(-1, 9)
    nn, ni = ctypes.c_int64(), ctypes.c_int64()
    z = ctypes.c_void_p()
    lib().oracle_kd_tree(v.ctypes.data, len(v), z, z, z, z, z, z, z, 0, z, 0, ctypes.byref(nn), ctypes.byref(ni))
    n = nn.value
    leaf, axis, right, first, count = (np.zeros(n, np.int32) for _ in range(5))
    split = np.zeros(n, np.float64)
    bounds = np.zeros((n, 6), np.float64)
    ids = np.zeros(max(ni.value, 1), np.int32)
    lib().oracle_kd_tree(v.ctypes.data, len(v), leaf.ctypes.data, axis.ctypes.data, split.ctypes.data,
                         right.ctypes.data, first.ctypes.data, count.ctypes.data, bounds.ctypes.data, n,
                         ids.ctypes.data, ni.value, ctypes.byref(nn), ctypes.byref(ni))
    return {"leaf": leaf, "axis": axis, "split": split, "right": right, "first": first, "count": count,
            "bounds": bounds, "tri_ids": ids[:ni.value]}


def mesh_hit(stl_path, use_kdtree, origin, direction, t_min=0.001, t_max=float("inf")):
    o = np.asarray(origin, np.float64)
    d = np.asarray(direction, np.float64)
    t = ctypes.c_double()
    tri = lib().oracle_mesh_hit(stl_path.encode(), int(use_kdtree), o.ctypes.data, d.ctypes.data, t_min, t_max,
                                ctypes.byref(t))
    return tri, (t.value if tri >= 0 else None)


def outline(depth, normal, color, opts):
    """depth (H,W) with NaN for background, normal (H,W,3), color (H,W,3) float64 -> new color."""
    H, W = depth.shape
    d = np.ascontiguousarray(depth, np.float64)
    n = np.ascontiguousarray(normal, np.float64)
    c = np.ascontiguousarray(color, np.float64).copy()
    lib().oracle_outline(W, H, d.ctypes.data, n.ctypes.data, ctypes.byref(opts), c.ctypes.data)
    return c
