"""rtrace_amd — MI355X (gfx950) backend for rtrace's per-pixel trace path.

Python mirror of the reference's render API over the C ABI in include/rtrace_hip.h:

    scene = Scene.from_json_file("scene.json")          # src/scene.rs:462-470
    r = Renderer(800, 600)                               # src/renderer.rs:50-62
    r.anti_aliasing_mode = AntiAliasingMode.Stochastic
    image = r.render(scene)                              # (H, W, 3) uint8, src/renderer.rs:120
    r.render_to_file(scene, "out.png")                   # src/renderer.rs:914-923

There is no CPU fallback: every render runs the HIP kernels in librtrace_hip.so and
fails loudly when the library or a gfx950 device is missing.
"""
import ctypes
import os

import numpy as np

__all__ = ["AntiAliasingMode", "OutlineConfig", "Renderer", "Scene", "RenderStats", "RtraceError", "lib",
           "intersect", "last_primary_hits", "cli_size", "load_stl", "kd_build"]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librtrace_hip.so")


class RtraceError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


class AntiAliasingMode:
    """src/renderer.rs:17-24"""
    NoJitter = 0
    Quincunx = 1
    Stochastic = 2


class _Opts(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int32), ("height", ctypes.c_int32), ("max_depth", ctypes.c_int32),
        ("use_kdtree", ctypes.c_int32), ("samples", ctypes.c_int32), ("aa_mode", ctypes.c_int32),
        ("has_seed", ctypes.c_int32), ("seed", ctypes.c_uint64),
        ("outline_enabled", ctypes.c_int32), ("outline_use_8_neighbors", ctypes.c_int32),
        ("outline_depth_weight", ctypes.c_double), ("outline_normal_weight", ctypes.c_double),
        ("outline_threshold", ctypes.c_double), ("outline_color", ctypes.c_double * 3),
        ("outline_thickness", ctypes.c_double), ("count_work", ctypes.c_int32), ("profile", ctypes.c_int32),
        ("light_samples", ctypes.c_int32),
    ]


class CameraDesc(ctypes.Structure):
    """rt_camera_desc: a Camera config (src/scene.rs:34-46)."""
    _fields_ = [("kind", ctypes.c_int32), ("position", ctypes.c_double * 3), ("target", ctypes.c_double * 3),
                ("up", ctypes.c_double * 3), ("width", ctypes.c_double), ("height", ctypes.c_double),
                ("has_fov", ctypes.c_int32), ("fov", ctypes.c_double), ("has_grid", ctypes.c_int32),
                ("grid_pitch", ctypes.c_double), ("grid_thickness", ctypes.c_double), ("grid_color", ctypes.c_double * 3)]

    def to_json(self):
        """The serde form of Camera (src/scene.rs:34-46), as AutoCameraResult::to_cameras_json writes it."""
        return {"kind": "perspective" if self.kind == 1 else "ortho", "position": list(self.position),
                "target": list(self.target), "up": list(self.up), "width": self.width, "height": self.height,
                "fov": self.fov if self.has_fov else None, "grid_pitch": None, "grid_color": None,
                "grid_thickness": None}


class _Aux(ctypes.Structure):
    _fields_ = [("features", ctypes.c_void_p), ("radiance", ctypes.c_void_p), ("tile_mask", ctypes.c_void_p),
                ("mask_tile", ctypes.c_int32)]


NOALLOC_SCRATCH_BYTES = 128 << 20  # RT_NOALLOC_SCRATCH_BYTES (include/rtrace_hip.h)
FEAT_COUNT = 13  # rt_render_aux feature channels (include/rtrace_hip.h)


class _BatchJob(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_void_p), ("camera", ctypes.c_void_p), ("opts", _Opts), ("d_rgb_out", ctypes.c_void_p)]


class BatchOpts(ctypes.Structure):
    """rt_batch_opts: raygen tile, chunk (items per take), persistent workgroups per CU, graph, memory cap,
    job-group size (GiB of frame buffers per group), groups in flight at once (each with its own buffers)."""
    _fields_ = [("tile", ctypes.c_int32), ("chunk", ctypes.c_int32), ("blocks_per_cu", ctypes.c_int32),
                ("use_graph", ctypes.c_int32), ("max_gib", ctypes.c_double), ("group_gib", ctypes.c_double),
                ("concurrent", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchStats(ctypes.Structure):
    _fields_ = [("ms", ctypes.c_double), ("ms_build", ctypes.c_double), ("n_jobs", ctypes.c_int32),
                ("graph_nodes", ctypes.c_int32), ("groups", ctypes.c_int32), ("bytes", ctypes.c_int64),
                ("pixels", ctypes.c_uint64),
                ("rays_primary", ctypes.c_uint64), ("rays_reflect", ctypes.c_uint64), ("rays_shadow", ctypes.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}

    @property
    def rays_total(self):
        return self.rays_primary + self.rays_reflect + self.rays_shadow

STAGES = ["raygen", "closest", "shade", "shadow", "finish", "post"]
KERNEL_CLOSEST, KERNEL_SHADOW = 0, 1


class Work(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in ("nodes", "tris", "retests", "spheres", "planes", "cubes", "skipped", "bvh_nodes")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class RenderStats(ctypes.Structure):
    _fields_ = [
        ("rays_primary", ctypes.c_uint64), ("rays_reflect", ctypes.c_uint64), ("rays_shadow", ctypes.c_uint64),
        ("work", Work * 2), ("ms_render", ctypes.c_double), ("ms_upload", ctypes.c_double),
        ("ms_stage", ctypes.c_double * len(STAGES)), ("launches_stage", ctypes.c_int32 * len(STAGES)), ("levels", ctypes.c_int32),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k in ("rays_primary", "rays_reflect", "rays_shadow", "ms_render", "ms_upload",
                                            "levels")}
        d["work"] = {"closest": self.work[0].as_dict(), "shadow": self.work[1].as_dict()}
        d["ms_stage"] = dict(zip(STAGES, list(self.ms_stage)))
        d["launches_stage"] = dict(zip(STAGES, list(self.launches_stage)))
        return d

    @property
    def rays_total(self):
        return self.rays_primary + self.rays_reflect + self.rays_shadow


_lib = None
_ctx = None


def lib():
    """Load librtrace_hip.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RtraceError(-5, f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
        # PyTorch-ROCm ships its own HIP runtime under the soname the library links against:
        # when it is installed, load it first so both share that runtime (loaded the other way
        # round, torch later finds no GPU).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        L.rt_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(P)]
        L.rt_ctx_destroy.argtypes = [P]
        L.rt_last_error.argtypes = [P]
        L.rt_last_error.restype = ctypes.c_char_p
        L.rt_scene_load_json_file.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(P)]
        L.rt_scene_load_json_str.argtypes = [P, ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(P)]
        L.rt_scene_upload.argtypes = [P, P, ctypes.POINTER(P)]
        L.rt_scene_free.argtypes = [P]
        L.rt_scene_info.argtypes = [P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(_Opts)]
        L.rt_render_opts_default.argtypes = [ctypes.POINTER(_Opts), ctypes.c_int32, ctypes.c_int32]
        L.rt_render.argtypes = [P, P, ctypes.POINTER(_Opts), P, ctypes.POINTER(RenderStats)]
        L.rt_render_device.argtypes = [P, P, ctypes.POINTER(_Opts), P, ctypes.POINTER(RenderStats)]
        L.rt_render_device_noalloc.argtypes = [P, P, ctypes.POINTER(_Opts), P, ctypes.POINTER(RenderStats)]
        L.rt_render_device_aux.argtypes = [P, P, ctypes.POINTER(_Opts), P, ctypes.POINTER(_Aux),
                                           ctypes.POINTER(RenderStats)]
        L.rt_last_primary_hits.argtypes = [P, P, P, P, ctypes.c_size_t]
        L.rt_intersect.argtypes = [P, P, P, P, ctypes.c_size_t, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                   P, P, P]
        L.rt_scene_auto_cameras.argtypes = [P, P]
        L.rt_auto_cameras_file.argtypes = [ctypes.c_char_p, P]
        L.rt_mesh_load_stl.argtypes = [ctypes.c_char_p, P, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
        L.rt_kd_build.argtypes = [P, P, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(_KdArrays)]
        L.rt_batch_opts_default.argtypes = [ctypes.POINTER(BatchOpts)]
        L.rt_batch_create.argtypes = [P, P, ctypes.c_int32, ctypes.POINTER(BatchOpts), ctypes.POINTER(P)]
        L.rt_batch_run.argtypes = [P, ctypes.POINTER(BatchStats)]
        L.rt_batch_run.restype = ctypes.c_int
        L.rt_batch_destroy.argtypes = [P]
        L.rt_batch_autotune.argtypes = [P, P, ctypes.c_int32, ctypes.POINTER(BatchOpts), ctypes.POINTER(ctypes.c_double)]
        _lib = L
    return _lib


def _context():
    global _ctx
    if _ctx is None:
        h = ctypes.c_void_p()
        rc = lib().rt_ctx_create(0, ctypes.byref(h))
        if rc != 0:
            raise RtraceError(rc, lib().rt_last_error(None).decode())
        _ctx = h
    return _ctx


def _check(rc):
    if rc != 0:
        raise RtraceError(rc, lib().rt_last_error(_context()).decode())


class Scene:
    """A scene uploaded to the device (Scene + World assembly, src/scene.rs / src/renderer.rs:137-317)."""

    def __init__(self, handle, path=None):
        self._h = handle
        self._path = path  # the JSON file it came from, if any

    @classmethod
    def from_json_file(cls, path):
        h = ctypes.c_void_p()
        _check(lib().rt_scene_load_json_file(_context(), os.fspath(path).encode(), ctypes.byref(h)))
        return cls(h, os.fspath(path))

    @classmethod
    def from_json_str(cls, text, base_dir=None):
        h = ctypes.c_void_p()
        _check(lib().rt_scene_load_json_str(_context(), text.encode(), base_dir.encode() if base_dir else None,
                                            ctypes.byref(h)))
        return cls(h)

    def info(self):
        w, h, en = ctypes.c_double(), ctypes.c_double(), ctypes.c_int32()
        o = _Opts()
        _check(lib().rt_scene_info(self._h, ctypes.byref(w), ctypes.byref(h), ctypes.byref(en), ctypes.byref(o)))
        outline = None
        if en.value:
            outline = OutlineConfig(o.outline_depth_weight, o.outline_normal_weight, o.outline_threshold,
                                    tuple(o.outline_color), bool(o.outline_use_8_neighbors), o.outline_thickness)
        return {"camera_width": w.value, "camera_height": h.value, "outline": outline}

    def close(self):
        if self._h:
            lib().rt_scene_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class OutlineConfig:
    """src/outline.rs:5-31"""

    def __init__(self, depth_weight=1.0, normal_weight=1.0, threshold=0.1, edge_color=(0.0, 0.0, 0.0),
                 use_8_neighbors=False, line_thickness=1.0):
        self.depth_weight = depth_weight
        self.normal_weight = normal_weight
        self.threshold = threshold
        self.edge_color = tuple(edge_color)
        self.use_8_neighbors = use_8_neighbors
        self.line_thickness = line_thickness


class Renderer:
    """Fields and defaults of `Renderer` (src/renderer.rs:37-118). thread_count is
    accepted for API parity; the GPU output is independent of it, as the reference's
    is of its thread count."""

    def __init__(self, width, height, use_kdtree=True, thread_count=None):
        self.width = width
        self.height = height
        self.max_depth = 10
        self.use_kdtree = use_kdtree
        self.thread_count = thread_count
        self.samples = 1
        self.anti_aliasing_mode = AntiAliasingMode.Quincunx
        self.seed = 0
        self.outline_config = None
        self.count_work = False
        self.profile = False
        self.light_samples = 0  # extension: > 0 overrides every area light's shadow samples
        self.last_stats = None

    new = classmethod(lambda cls, w, h: cls(w, h))
    new_brute_force = classmethod(lambda cls, w, h: cls(w, h, use_kdtree=False))
    new_with_threads = classmethod(lambda cls, w, h, n: cls(w, h, thread_count=n))
    new_with_options = classmethod(lambda cls, w, h, use_kdtree, thread_count: cls(w, h, use_kdtree, thread_count))

    def with_outline_detection(self, config):
        self.outline_config = config
        return self

    def _opts(self):
        o = _Opts()
        lib().rt_render_opts_default(ctypes.byref(o), self.width, self.height)
        o.max_depth = self.max_depth
        o.use_kdtree = int(bool(self.use_kdtree))
        o.samples = self.samples
        o.aa_mode = self.anti_aliasing_mode
        o.has_seed = 0 if self.seed is None else 1
        o.seed = 0 if self.seed is None else self.seed
        oc = self.outline_config
        if oc is not None:
            o.outline_enabled = 1
            o.outline_use_8_neighbors = int(bool(oc.use_8_neighbors))
            o.outline_depth_weight = oc.depth_weight
            o.outline_normal_weight = oc.normal_weight
            o.outline_threshold = oc.threshold
            for i in range(3):
                o.outline_color[i] = oc.edge_color[i]
            o.outline_thickness = oc.line_thickness
        o.count_work = int(bool(self.count_work))
        o.profile = int(bool(self.profile))
        o.light_samples = int(self.light_samples)
        return o

    def render(self, scene):
        """-> (height, width, 3) uint8, row-major RGB (create_image_from_data, src/renderer.rs:899-912)."""
        o = self._opts()
        img = np.zeros((self.height, self.width, 3), dtype=np.uint8)
        st = RenderStats()
        _check(lib().rt_render(_context(), scene._h, ctypes.byref(o), img.ctypes.data, ctypes.byref(st)))
        self.last_stats = st
        return img

    def render_device(self, scene, out_tensor):
        """Render into a device buffer (e.g. a torch.uint8 CUDA tensor of H*W*3); blocks until done."""
        o = self._opts()
        st = RenderStats()
        _check(lib().rt_render_device(_context(), scene._h, ctypes.byref(o), ctypes.c_void_p(out_tensor.data_ptr()),
                                      ctypes.byref(st)))
        self.last_stats = st
        return st

    def render_device_noalloc(self, scene, out_tensor):
        """render_device with the frame's buffers in the library's fixed scratch array
        (rt_render_device_noalloc: no device allocation, eager launches; frames whose
        buffers exceed NOALLOC_SCRATCH_BYTES are refused)."""
        o = self._opts()
        st = RenderStats()
        _check(lib().rt_render_device_noalloc(_context(), scene._h, ctypes.byref(o),
                                              ctypes.c_void_p(out_tensor.data_ptr()), ctypes.byref(st)))
        self.last_stats = st
        return st

    def render_device_aux(self, scene, out_tensor, features=None, radiance=None, tile_mask=None, mask_tile=16):
        """render_device plus per-pixel features (H*W*13 f32), radiance (H*W*3 f32) and an
        optional per-tile mask (uint8, ceil(H/T) x ceil(W/T)): only masked tiles are traced
        and written into out_tensor. Tensors are device tensors."""
        o = self._opts()
        a = _Aux()
        a.features = features.data_ptr() if features is not None else None
        a.radiance = radiance.data_ptr() if radiance is not None else None
        a.tile_mask = tile_mask.data_ptr() if tile_mask is not None else None
        a.mask_tile = mask_tile
        st = RenderStats()
        _check(lib().rt_render_device_aux(_context(), scene._h, ctypes.byref(o), ctypes.c_void_p(out_tensor.data_ptr()),
                                          ctypes.byref(a), ctypes.byref(st)))
        self.last_stats = st
        return st

    def render_to_file(self, scene, path):
        from PIL import Image
        Image.fromarray(self.render(scene)).save(path)


def last_primary_hits(n):
    """Primary-ray (object, triangle, t) of the last render, canonical order."""
    obj = np.zeros(n, np.int32)
    tri = np.zeros(n, np.int32)
    t = np.zeros(n, np.float64)
    _check(lib().rt_last_primary_hits(_context(), obj.ctypes.data, tri.ctypes.data, t.ctypes.data, n))
    return obj, tri, t


def intersect(scene, origins, dirs, t_min=0.001, t_max=float("inf"), use_kdtree=True):
    """World::hit for a batch of rays (src/ray.rs:699-711) -> (obj, tri, t)."""
    o = np.ascontiguousarray(origins, np.float64).reshape(-1, 3)
    d = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
    n = len(o)
    obj = np.zeros(n, np.int32)
    tri = np.zeros(n, np.int32)
    t = np.zeros(n, np.float64)
    _check(lib().rt_intersect(_context(), scene._h, o.ctypes.data, d.ctypes.data, n, t_min, t_max, int(use_kdtree),
                              obj.ctypes.data, tri.ctypes.data, t.ctypes.data))
    return obj, tri, t


def load_stl(path):
    """Mesh::from_stl_file (src/mesh.rs:751-1003) on the host: the triangles as an
    (n, 3, 3) float64 array (vertex-major, xyz). Raises RtraceError with the reference's
    message for a malformed file."""
    p = os.fspath(path).encode()
    n = ctypes.c_int64(0)
    _check_global(lib().rt_mesh_load_stl(p, None, 0, ctypes.byref(n)))
    out = np.zeros((n.value, 3, 3), np.float64)
    if n.value:
        _check_global(lib().rt_mesh_load_stl(p, out.ctypes.data, n.value, ctypes.byref(n)))
    return out


class _KdArrays(ctypes.Structure):
    _fields_ = [("nodes", ctypes.c_void_p), ("nodes_cap", ctypes.c_int64), ("recs", ctypes.c_void_p),
                ("recs_cap", ctypes.c_int64), ("bounds", ctypes.c_void_p), ("bounds_cap", ctypes.c_int64),
                ("n_nodes", ctypes.c_int64), ("n_recs", ctypes.c_int64), ("root_kd", ctypes.c_int64),
                ("root_bf", ctypes.c_int64), ("leaves", ctypes.c_int64), ("max_leaf", ctypes.c_int64)]


def kd_build(triangles, on_device=False):
    """rt_kd_build: the flattened k-d tree of KdTree::new (src/mesh.rs:75-197) for an (n, 3, 3)
    float64 triangle array, built on the host or on the device. Returns a dict with `nodes`
    (n_nodes x 48 bytes), `recs` (n_recs x 48 bytes), `bounds` (n_nodes x 6 float64) and the
    roots and leaf statistics."""
    v = np.ascontiguousarray(triangles, np.float64).reshape(-1, 9)
    a = _KdArrays()
    # a host build needs no device (no context); a device build runs on the context's GPU
    ctx, check = (_context(), _check) if on_device else (None, _check_global)
    check(lib().rt_kd_build(ctx, v.ctypes.data, len(v), int(bool(on_device)), ctypes.byref(a)))
    nodes = np.zeros((a.n_nodes, 48), np.uint8)
    recs = np.zeros((a.n_recs, 48), np.uint8)
    bounds = np.zeros((a.n_nodes, 6), np.float64)
    a.nodes, a.nodes_cap = nodes.ctypes.data, a.n_nodes
    a.recs, a.recs_cap = recs.ctypes.data, a.n_recs
    a.bounds, a.bounds_cap = bounds.ctypes.data, a.n_nodes
    check(lib().rt_kd_build(ctx, v.ctypes.data, len(v), int(bool(on_device)), ctypes.byref(a)))
    return {"nodes": nodes, "recs": recs, "bounds": bounds, "root_kd": a.root_kd, "root_bf": a.root_bf,
            "leaves": a.leaves, "max_leaf": a.max_leaf}


def _check_global(rc):
    if rc != 0:
        raise RtraceError(rc, lib().rt_last_error(None).decode())


def cli_size(camera_width, camera_height, diagonal):
    """cli/src/main.rs:79-89: H = round(D / sqrt(R^2 + 1)), W = round(R * H), round half away from zero."""
    import math
    r = camera_width / camera_height
    h = diagonal / math.sqrt(r * r + 1.0)
    w = r * h

    def rnd(x):
        f = math.floor(abs(x))
        return int(math.copysign(f + (1.0 if abs(x) - f >= 0.5 else 0.0), x))

    return rnd(w), rnd(h)


# ---------------------------------------------------------------- auto cameras
CAMERA_NAMES = ("left", "front", "top", "perspective")


class AutoCameraResult:
    """The four cameras of AutoCamera::generate_cameras (src/auto_camera.rs:9-130)."""

    def __init__(self, cams):
        self.left, self.front, self.top, self.perspective = cams

    def cameras(self):
        return [self.left, self.front, self.top, self.perspective]

    def to_cameras_json(self):
        return {n: c.to_json() for n, c in zip(CAMERA_NAMES, self.cameras())}


class AutoCamera:
    @staticmethod
    def generate_cameras(scene):
        """`scene`: a loaded Scene, or a Scene JSON path (host only, no device needed)."""
        out = (CameraDesc * 4)()
        if isinstance(scene, (str, os.PathLike)):
            _check_global(lib().rt_auto_cameras_file(os.fspath(scene).encode(), out))
        else:
            _check(lib().rt_scene_auto_cameras(scene._h, out))
        return AutoCameraResult([out[i] for i in range(4)])


# ---------------------------------------------------------------- batches
class Batch:
    """Many frames in one captured launch of persistent kernels (rt_batch_*).

    jobs: list of (scene, camera or None, renderer, out) where `renderer` supplies the
    Renderer fields of that frame and `out` is a device uint8 tensor of H*W*3. The batch
    keeps a reference to every scene and output tensor: its captured launch holds their
    device pointers until close()."""

    def __init__(self, jobs, opts=None):
        self._jobs, self._keep = _make_jobs(jobs)
        o = opts if opts is not None else default_batch_opts()
        self.opts = o
        self._h = ctypes.c_void_p()
        _check(lib().rt_batch_create(_context(), self._jobs, len(jobs), ctypes.byref(o), ctypes.byref(self._h)))

    def run(self):
        st = BatchStats()
        _check(lib().rt_batch_run(self._h, ctypes.byref(st)))
        return st

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().rt_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown: the library may already be gone
            pass


def default_batch_opts():
    o = BatchOpts()
    lib().rt_batch_opts_default(ctypes.byref(o))
    return o


def _make_jobs(jobs):
    """The C job array, and what must outlive it: the camera copies it points at, and every
    scene and output tensor whose device memory the captured launch reads and writes."""
    arr = (_BatchJob * len(jobs))()
    keep = []
    for i, (scene, cam, renderer, out) in enumerate(jobs):
        keep.append(scene)
        keep.append(out)
        arr[i].scene = scene._h
        if cam is not None:
            c = CameraDesc()
            ctypes.memmove(ctypes.byref(c), ctypes.byref(cam), ctypes.sizeof(CameraDesc))
            keep.append(c)
            arr[i].camera = ctypes.addressof(c)
        arr[i].opts = renderer._opts()
        arr[i].d_rgb_out = out.data_ptr()
    return arr, keep


def autotune_batch(jobs, opts=None):
    """rt_batch_autotune: the fastest (tile, chunk, blocks_per_cu) for these jobs and its time (ms)."""
    arr, keep = _make_jobs(jobs)
    o = opts if opts is not None else default_batch_opts()
    ms = ctypes.c_double()
    _check(lib().rt_batch_autotune(_context(), arr, len(jobs), ctypes.byref(o), ctypes.byref(ms)))
    return o, ms.value

