"""CPU restatement of the shadow denoiser (the U-Net of include/rtrace_denoise.h).

TEST INFRASTRUCTURE ONLY: the parity oracle of the HIP kernels in
rtrace_amd/csrc/kernels/denoise.hip. Nothing in the product imports it.

The network has no counterpart in the reference (SeaRyanC/rtrace estimates area lights
with 16 shadow rays per hit, src/lighting.rs:145-217); this file states it in numpy with
an explicit backward pass. With `bf16=True` every tensor the device stores in bf16 is
rounded to bf16 (round to nearest even) at the same points as the device: the network
input, the weights, each post-ReLU activation, the pooled tensors and each dZ; sums run in
float64 (the device accumulates in f32). With `bf16=False` everything is float64, which
the CPU tests use to check the backward pass against finite differences and torch
autograd.
"""
import numpy as np

N_FEAT = 13
# (level, in0, up0, in1, cin, cout, relu); in0: -1 input, -2/-3/-4 pool of act 1/3/5
LAYERS = [
    (0, -1, 0, None, 32, 32, 1), (0, 0, 0, None, 32, 32, 1),
    (1, -2, 0, None, 32, 64, 1), (1, 2, 0, None, 64, 64, 1),
    (2, -3, 0, None, 64, 128, 1), (2, 4, 0, None, 128, 128, 1),
    (3, -4, 0, None, 128, 128, 1), (3, 6, 0, None, 128, 128, 1),
    (2, 7, 1, 5, 256, 128, 1), (2, 8, 0, None, 128, 128, 1),
    (1, 9, 1, 3, 192, 64, 1), (1, 10, 0, None, 64, 64, 1),
    (0, 11, 1, 1, 96, 32, 1), (0, 12, 0, None, 32, 32, 1),
    (0, 13, 0, None, 32, 3, 0),
]
POOL_SRC = {-2: 1, -3: 3, -4: 5}
OUT = len(LAYERS) - 1


def coutp(l):
    return (LAYERS[l][5] + 31) // 32 * 32


def offsets():
    o, res = 0, []
    for l, L in enumerate(LAYERS):
        w = o
        o += 9 * L[4] * coutp(l)
        b = o
        o += coutp(l)
        res.append((w, b))
    return res, o


OFFS, N_PARAMS = offsets()


def bf16(x):
    """Round to the nearest bfloat16 (ties to even); returns float32 values."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, ((u >> 16) | 0x40) << 16, r)
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


def split_params(params):
    ws, bs = [], []
    for l, L in enumerate(LAYERS):
        w, b = OFFS[l]
        ws.append(np.asarray(params[w:w + 9 * L[4] * coutp(l)], np.float64).reshape(9, L[4], coutp(l)))
        bs.append(np.asarray(params[b:b + coutp(l)], np.float64))
    return ws, bs


def prep(feat, Hp, Wp):
    """Network input (the device's k_prep): f32 normalisation, zero pad to Hp x Wp, 32 ch."""
    B, H, W, _ = feat.shape
    f = feat.astype(np.float32).copy()
    f[..., 0:3] = np.clip(f[..., 0:3], np.float32(0), np.float32(4))
    f[..., 9] = np.log1p(np.maximum(f[..., 9], np.float32(0))) * np.float32(0.25)
    f[..., 11] = np.sqrt(np.maximum(f[..., 11], np.float32(0))) * np.float32(4)
    x = np.zeros((B, Hp, Wp, 32), np.float32)
    x[:, :H, :W, :N_FEAT] = f
    return x


def conv(x, w, b=None):
    """3x3 'same' cross-correlation: out[p, co] = sum_tap sum_ci x[p + (dy-1, dx-1), ci] w[tap, ci, co],
    tap = 3 dy + dx; zero padding."""
    B, H, W, _ = x.shape
    xp = np.zeros((B, H + 2, W + 2, x.shape[3]), np.float64)
    xp[:, 1:-1, 1:-1] = x
    out = np.zeros((B, H, W, w.shape[2]), np.float64)
    for t in range(9):
        dy, dx = divmod(t, 3)
        out += np.tensordot(xp[:, dy:dy + H, dx:dx + W], w[t], axes=([3], [0]))
    if b is not None:
        out += b
    return out


def conv_wgrad(x, dz):
    """dW[tap, ci, co] = sum_p x[p + tap offset, ci] dz[p, co]."""
    B, H, W, _ = x.shape
    xp = np.zeros((B, H + 2, W + 2, x.shape[3]), np.float64)
    xp[:, 1:-1, 1:-1] = x
    g = np.zeros((9, x.shape[3], dz.shape[3]), np.float64)
    for t in range(9):
        dy, dx = divmod(t, 3)
        g[t] = np.tensordot(xp[:, dy:dy + H, dx:dx + W], dz, axes=([0, 1, 2], [0, 1, 2]))
    return g


def conv_bwd_data(dz, w):
    """Gradient of conv() with respect to its input: the correlation of dz with the weights
    flipped in both spatial axes and transposed (ci <-> co)."""
    wb = w[::-1].transpose(0, 2, 1)  # tap t of the flipped kernel is tap 8 - t
    return conv(dz, wb)


def pool(x):
    return ((x[:, 0::2, 0::2] + x[:, 0::2, 1::2]) + (x[:, 1::2, 0::2] + x[:, 1::2, 1::2])) * 0.25


def up(x):
    return x.repeat(2, axis=1).repeat(2, axis=2)


class UNet:
    def __init__(self, params, bf16_emulation=True):
        self.q = bf16 if bf16_emulation else (lambda a: np.asarray(a, np.float64))
        self.emu = bf16_emulation
        ws, bs = split_params(params)
        self.w = [np.asarray(self.q(w), np.float64) for w in ws]
        self.b = [np.asarray(b, np.float32 if bf16_emulation else np.float64).astype(np.float64) for b in bs]

    def _input(self, l, acts, pools):
        lv, in0, up0, in1, cin, cout, relu = LAYERS[l]
        src = self.x0 if in0 == -1 else pools[in0] if in0 < -1 else acts[in0]
        if up0:
            src = up(src)
        if in1 is not None:
            src = np.concatenate([src, acts[in1]], axis=3)
        return src

    def forward(self, feat):
        B, H, W, _ = feat.shape
        self.H, self.W = H, W
        Hp, Wp = (H + 7) // 8 * 8, (W + 7) // 8 * 8
        x0 = prep(feat, Hp, Wp)
        self.x0 = np.asarray(self.q(x0) if self.emu else x0, np.float64)
        acts, pools = {}, {}
        for l, L in enumerate(LAYERS):
            if l in (2, 4, 6):
                src = POOL_SRC[LAYERS[l][1]]
                pools[LAYERS[l][1]] = np.asarray(self.q(pool(acts[src])), np.float64)
            z = conv(self._input(l, acts, pools), self.w[l], self.b[l])
            if l == OUT:
                acts[l] = (z[..., :3].astype(np.float32) if self.emu else z[..., :3])
            else:
                a = np.maximum(z, 0.0) if L[6] else z
                acts[l] = np.asarray(self.q(a), np.float64)
        self.acts, self.pools = acts, pools
        return acts[OUT]

    def output(self, feat):
        y = self.forward(feat)
        return feat[..., 0:3] + y[:, :feat.shape[1], :feat.shape[2]]

    def loss_and_grads(self, feat, target):
        """Mean squared error of (noisy + y) against target and d(loss)/d(params) (flat, the
        device's parameter layout)."""
        y = self.forward(feat)
        B, H, W, _ = feat.shape
        Hp, Wp = y.shape[1], y.shape[2]
        diff = (feat[..., 0:3].astype(np.float64) + y[:, :H, :W]) - target
        loss = float(np.mean(diff * diff))
        dy = np.zeros((B, Hp, Wp, 32), np.float64)
        dy[:, :H, :W, :3] = 2.0 * diff / (B * H * W * 3)
        dz = np.asarray(self.q(dy), np.float64)
        grads = np.zeros(N_PARAMS, np.float64)
        g = {}
        acts, pools = self.acts, self.pools
        for l in range(OUT, -1, -1):
            lv, in0, up0, in1, cin, cout, relu = LAYERS[l]
            if l != OUT:
                gl = g[l]
                dz = np.where(acts[l] > 0, gl, 0.0)
                dz = np.asarray(self.q(dz), np.float64)
            x = self._input(l, acts, pools)
            wo, bo = OFFS[l]
            grads[wo:wo + 9 * cin * coutp(l)] = conv_wgrad(x, dz).ravel()
            grads[bo:bo + coutp(l)] = dz.sum(axis=(0, 1, 2))
            if l == 0:
                break
            dx = conv_bwd_data(dz, self.w[l])
            if in1 is not None:
                c0 = LAYERS[in0][5]
                d_up = dx[..., :c0]
                g[in0] = (d_up[:, 0::2, 0::2] + d_up[:, 0::2, 1::2]) + (d_up[:, 1::2, 0::2] + d_up[:, 1::2, 1::2])
                g[in1] = dx[..., c0:]
            elif in0 < -1:
                src = POOL_SRC[in0]
                g[src] = g[src] + 0.25 * up(dx)
            else:
                g[in0] = dx
        return loss, grads


def init_params(seed=0, scale_out=0.1):
    """He-normal weights (numpy generator; for CPU tests only)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, np.float32)
    for l, L in enumerate(LAYERS):
        cin, cout = L[4], L[5]
        fan = N_FEAT if l == 0 else cin
        w = rng.standard_normal((9, cin, coutp(l))) * np.sqrt(2.0 / (9 * fan)) * (scale_out if l == OUT else 1.0)
        w[:, fan:, :] = 0
        w[:, :, cout:] = 0
        wo, bo = OFFS[l]
        p[wo:wo + w.size] = w.ravel()
        p[bo:bo + cout] = rng.standard_normal(cout) * 0.01
    return p
