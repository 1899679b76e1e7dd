"""Python face of the shadow denoiser (include/rtrace_denoise.h) over ctypes.

Device tensors are torch CUDA tensors; the library only sees their pointers. The trained
parameters ship as `rtrace_amd/data/denoiser*.npy` (float32, the C ABI's parameter layout), one
set per light-sample count of the network's input render (WEIGHTS_BY_SAMPLES).
"""
import ctypes
import os

import numpy as np

from . import _check, _context, lib

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
WEIGHTS = os.path.join(_DATA, "denoiser.npy")
# Networks trained on input renders with this many light samples per hit (4: the configuration's
# own; 2 and 1: cheaper inputs). Each is trained by rtrace_amd.train_denoiser (its docstring
# gives the commands).
WEIGHTS_BY_SAMPLES = {4: WEIGHTS, 2: os.path.join(_DATA, "denoiser_s2.npy"), 1: os.path.join(_DATA, "denoiser_s1.npy")}
N_LAYERS = 15


def fallback_threshold(samples, at_four=0.06):
    """The tile-fallback threshold for `samples` light samples per hit: the binomial variance
    estimate m(1-m)/n is at most 0.25/n, so the threshold keeps the same fraction of its maximum
    as `at_four` does at 4 samples. With 1 sample every estimate is 0 and no tile is flagged."""
    return at_four * 4.0 / samples


class DenoiseStats(ctypes.Structure):
    _fields_ = [("ms_render", ctypes.c_double), ("ms_denoise", ctypes.c_double), ("ms_fallback", ctypes.c_double),
                ("ms_total", ctypes.c_double), ("tiles", ctypes.c_int32), ("tiles_fallback", ctypes.c_int32),
                ("rays", ctypes.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_bound = False


def _bind():
    global _bound
    if _bound:
        return lib()
    L = lib()
    P = ctypes.c_void_p
    L.rt_denoiser_create.argtypes = [P, ctypes.c_uint64, ctypes.POINTER(P)]
    L.rt_denoiser_destroy.argtypes = [P]
    L.rt_denoiser_param_count.restype = ctypes.c_int64
    L.rt_denoiser_layer_info.argtypes = [ctypes.c_int32, ctypes.POINTER(ctypes.c_int32),
                                         ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64),
                                         ctypes.POINTER(ctypes.c_int64)]
    L.rt_denoiser_get_params.argtypes = [P, P, ctypes.c_int64]
    L.rt_denoiser_set_params.argtypes = [P, P, ctypes.c_int64]
    L.rt_denoiser_train_step.argtypes = [P, P, P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                                         ctypes.POINTER(ctypes.c_float)]
    L.rt_denoiser_gradients.argtypes = [P, P, P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, P,
                                        ctypes.POINTER(ctypes.c_float)]
    L.rt_denoiser_activation.argtypes = [P, ctypes.c_int32, P, ctypes.c_int64]
    L.rt_denoise.argtypes = [P, P, ctypes.c_int32, ctypes.c_int32, P, P, ctypes.POINTER(ctypes.c_float)]
    L.rt_render_denoised.argtypes = [P, P, P, P, ctypes.c_int32, ctypes.c_float, ctypes.c_int32, P,
                                     ctypes.POINTER(DenoiseStats)]
    _bound = True
    return L


def param_count():
    return int(_bind().rt_denoiser_param_count())


def layer_info(layer):
    cin, cp = ctypes.c_int32(), ctypes.c_int32()
    wo, bo = ctypes.c_int64(), ctypes.c_int64()
    _check(_bind().rt_denoiser_layer_info(layer, ctypes.byref(cin), ctypes.byref(cp), ctypes.byref(wo),
                                          ctypes.byref(bo)))
    return cin.value, cp.value, wo.value, bo.value


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class Denoiser:
    """The U-Net: create(seed) for fresh He-normal weights, or load() the shipped ones."""

    def __init__(self, seed=0):
        L = _bind()
        self._h = ctypes.c_void_p()
        _check(L.rt_denoiser_create(_context(), seed, ctypes.byref(self._h)))
        self.n_params = param_count()

    @classmethod
    def load(cls, path=WEIGHTS):
        d = cls(0)
        d.set_params(np.load(path, allow_pickle=False))
        return d

    def save(self, path=WEIGHTS):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.save(path, self.get_params())

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is None or not h.value:
            return
        self._h = None
        try:
            lib().rt_denoiser_destroy(h)
        except Exception:  # interpreter shutdown: module globals may already be gone
            pass

    def get_params(self):
        p = np.zeros(self.n_params, np.float32)
        _check(_bind().rt_denoiser_get_params(self._h, p.ctypes.data, p.size))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(_bind().rt_denoiser_set_params(self._h, p.ctypes.data, p.size))

    def train_step(self, feat, target, lr):
        """One Adam step; feat [B, H, W, 13], target [B, H, W, 3] float32 device tensors."""
        B, H, W, _ = feat.shape
        loss = ctypes.c_float()
        _check(_bind().rt_denoiser_train_step(self._h, _ptr(feat), _ptr(target), B, H, W, lr, ctypes.byref(loss)))
        return loss.value

    def gradients(self, feat, target):
        B, H, W, _ = feat.shape
        g = np.zeros(self.n_params, np.float32)
        loss = ctypes.c_float()
        _check(_bind().rt_denoiser_gradients(self._h, _ptr(feat), _ptr(target), B, H, W, g.ctypes.data,
                                             ctypes.byref(loss)))
        return loss.value, g

    def activation(self, layer, shape):
        out = np.zeros(int(np.prod(shape)), np.float32)
        _check(_bind().rt_denoiser_activation(self._h, layer, out.ctypes.data, out.size))
        return out.reshape(shape)

    def denoise(self, feat, out_f32=None, out_u8=None):
        """feat [H, W, 13] device tensor -> fills out_f32 [H, W, 3] and/or out_u8; returns device ms."""
        H, W, _ = feat.shape
        ms = ctypes.c_float()
        _check(_bind().rt_denoise(self._h, _ptr(feat), H, W, _ptr(out_f32), _ptr(out_u8), ctypes.byref(ms)))
        return ms.value

    def render(self, renderer, scene, out_u8, fallback_samples=256, threshold=0.06, tile=16):
        """Denoised frame of `scene` with `renderer`'s settings (its light_samples is the low
        count) into out_u8 (device, H*W*3); returns DenoiseStats."""
        o = renderer._opts()
        st = DenoiseStats()
        _check(_bind().rt_render_denoised(_context(), scene._h, ctypes.byref(o), self._h, fallback_samples, threshold,
                                          tile, _ptr(out_u8), ctypes.byref(st)))
        return st


def evaluate(dn, seeds, width=1920, height=1080, low=4, high=256, threshold=0.06, tile=16, subdiv=6, repeats=3,
             work_dir=None, log=None):
    """Denoiser frames: per seed, the plain render with `high` light samples (the quality reference),
    the plain render with `low`, and the denoised render (low + network + tile fallback at
    `high`). Returns per-image PSNR / SSIM against that reference and the frame times (ms, median
    of `repeats` timed runs after one untimed)."""
    import tempfile
    import time

    import torch

    import rtrace_amd as R
    from . import scenes
    from .metrics import psnr, ssim

    work_dir = work_dir or tempfile.mkdtemp(prefix="rt_dn_eval_")
    r = R.Renderer(width, height)
    r.max_depth = 4
    out = torch.empty(height * width * 3, dtype=torch.uint8, device="cuda")
    rows = []

    def timed(fn):
        fn()
        ts = []
        for _ in range(repeats):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            res = fn()
            torch.cuda.synchronize()
            ts.append((time.perf_counter() - t0) * 1e3)
        return float(np.median(ts)), res

    for s in seeds:
        scene = R.Scene.from_json_file(scenes.denoise_scene(work_dir, s, subdiv=subdiv))
        r.light_samples = high
        ms_plain, st_plain = timed(lambda: r.render_device(scene, out))
        ref = out.cpu().numpy().reshape(height, width, 3).copy()
        r.light_samples = low
        ms_low, _ = timed(lambda: r.render_device(scene, out))
        noisy = out.cpu().numpy().reshape(height, width, 3).copy()
        ms_dn, st_dn = timed(lambda: dn.render(r, scene, out, fallback_samples=high, threshold=threshold, tile=tile))
        den = out.cpu().numpy().reshape(height, width, 3).copy()
        # stage split of the few-sample feature render (one profiled run after the timed ones)
        feat = torch.empty(height * width * 13, dtype=torch.float32, device="cuda")
        scratch = torch.empty_like(out)
        r.profile = True
        st_f = r.render_device_aux(scene, scratch, features=feat)
        r.profile = False
        # the network alone on those features (prep, convolutions, compose: device time)
        dn.denoise(feat.view(height, width, 13))
        ms_net = float(np.median([dn.denoise(feat.view(height, width, 13)) for _ in range(repeats)]))
        del feat, scratch
        render_stages = {k: round(v, 3) for k, v in st_f.as_dict()["ms_stage"].items()}
        row = {"seed": s, "ms_plain_high": round(ms_plain, 3), "ms_plain_low": round(ms_low, 3),
               "ms_denoised": round(ms_dn, 3), "speedup": round(ms_plain / ms_dn, 2),
               "psnr_low": round(psnr(noisy, ref), 2), "psnr_denoised": round(psnr(den, ref), 2),
               "ssim_low": round(ssim(noisy, ref), 5), "ssim_denoised": round(ssim(den, ref), 5),
               "tiles_fallback": st_dn.tiles_fallback, "tiles": st_dn.tiles,
               "stage_ms": {"render": round(st_dn.ms_render, 3), "denoise": round(st_dn.ms_denoise, 3),
                            "fallback": round(st_dn.ms_fallback, 3), "total": round(st_dn.ms_total, 3)},
               "ms_network": round(ms_net, 3),
               "render_stages_ms": render_stages,
               "rays_plain_high": int(st_plain.rays_total)}
        rows.append(row)
        if log:
            log(str(row))
    return rows
