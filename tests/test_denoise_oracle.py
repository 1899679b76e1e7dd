"""Pins for the denoiser's CPU oracle (oracle/denoise_ref.py) before it checks the HIP
kernels: bf16 rounding against torch's bfloat16 conversion, the convolution against
torch.nn.functional.conv2d, the explicit backward pass against central finite
differences and against torch autograd of the same network, all in float64."""
import numpy as np
import pytest
import torch

from oracle import denoise_ref as D


def test_bf16_rounding_matches_torch_and_ties_to_even():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-30, 30, 20000),
                        np.float32([0.0, -0.0, 1.0, 1 + 2**-8, 1 + 3 * 2**-8, 65504.0, 3.4e38, 1e-40])])
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(D.bf16(x), want)
    assert D.bf16(np.float32(1 + 2**-8)) == 1.0 and D.bf16(np.float32(1 + 3 * 2**-8)) == np.float32(1 + 2**-6)


def test_conv_is_torch_cross_correlation():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 9, 7, 32))
    w = rng.standard_normal((9, 32, 64))
    b = rng.standard_normal(64)
    got = D.conv(x, w, b)
    tw = torch.from_numpy(w.reshape(3, 3, 32, 64).transpose(3, 2, 0, 1).copy())  # [co][ci][ky][kx]
    ref = torch.nn.functional.conv2d(torch.from_numpy(x.transpose(0, 3, 1, 2).copy()), tw, torch.from_numpy(b),
                                     padding=1).numpy().transpose(0, 2, 3, 1)
    assert np.abs(got - ref).max() < 1e-11


def test_conv_input_and_weight_gradients_are_the_adjoints():
    """<conv(x, w), y> = <x, conv_bwd_data(y, w)> = <w, conv_wgrad(x, y)>."""
    rng = np.random.default_rng(2)
    x = rng.standard_normal((1, 8, 16, 32))
    w = rng.standard_normal((9, 32, 32))
    y = rng.standard_normal((1, 8, 16, 32))
    lhs = np.sum(D.conv(x, w) * y)
    assert abs(lhs - np.sum(x * D.conv_bwd_data(y, w))) < 1e-9 * abs(lhs) + 1e-9
    assert abs(lhs - np.sum(w * D.conv_wgrad(x, y))) < 1e-9 * abs(lhs) + 1e-9


def _batch(seed, B=1, H=8, W=16):
    rng = np.random.default_rng(seed)
    feat = rng.random((B, H, W, D.N_FEAT)).astype(np.float32)
    feat[..., 6:9] = feat[..., 6:9] * 2 - 1
    feat[..., 9] *= 20
    feat[..., 11] *= 0.25
    target = np.clip(feat[..., 0:3] + 0.1 * rng.standard_normal((B, H, W, 3)), 0, 1).astype(np.float32)
    return feat, target


def test_backward_matches_finite_differences():
    params = D.init_params(3).astype(np.float64)
    params += np.random.default_rng(4).standard_normal(params.size) * 1e-3  # nonzero padded/biases
    feat, target = _batch(5, H=8, W=8)
    net = D.UNet(params, bf16_emulation=False)
    loss, g = net.loss_and_grads(feat, target)
    rng = np.random.default_rng(6)
    for l, L in enumerate(D.LAYERS):
        wo, bo = D.OFFS[l]
        cin, cout = L[4], L[5]
        for _ in range(3):
            tap, ci, co = rng.integers(9), rng.integers(min(cin, 13) if l == 0 else cin), rng.integers(cout)
            i = wo + (tap * cin + ci) * D.coutp(l) + co
            for idx in (i, bo + co):
                h = 1e-5
                p1, p2 = params.copy(), params.copy()
                p1[idx] += h
                p2[idx] -= h
                l1, _ = D.UNet(p1, False).loss_and_grads(feat, target)
                l2, _ = D.UNet(p2, False).loss_and_grads(feat, target)
                fd = (l1 - l2) / (2 * h)
                assert abs(fd - g[idx]) <= 1e-6 * max(abs(fd), abs(g[idx])) + 1e-11, (l, idx, fd, g[idx])


def _torch_unet(params, feat):
    """The same network written with torch modules (float64), for autograd."""
    ws, bs = D.split_params(params)
    tw = [torch.tensor(w.reshape(3, 3, w.shape[1], w.shape[2]).transpose(3, 2, 0, 1).copy(), requires_grad=True)
          for w in ws]
    tb = [torch.tensor(b, requires_grad=True) for b in bs]
    B, H, W, _ = feat.shape
    x = torch.from_numpy(D.prep(feat, (H + 7) // 8 * 8, (W + 7) // 8 * 8).astype(np.float64)).permute(0, 3, 1, 2)
    acts, pools = {}, {}
    F = torch.nn.functional
    for l, (lv, in0, up0, in1, cin, cout, relu) in enumerate(D.LAYERS):
        if l in (2, 4, 6):
            pools[in0] = F.avg_pool2d(acts[D.POOL_SRC[in0]], 2)
        src = x if in0 == -1 else pools[in0] if in0 < -1 else acts[in0]
        if up0:
            src = F.interpolate(src, scale_factor=2, mode="nearest")
        if in1 is not None:
            src = torch.cat([src, acts[in1]], 1)
        z = F.conv2d(src, tw[l], tb[l], padding=1)
        acts[l] = torch.relu(z) if relu else z
    return acts[D.OUT][:, :3, :H, :W].permute(0, 2, 3, 1), tw, tb


def test_backward_matches_torch_autograd():
    params = D.init_params(7).astype(np.float64)
    feat, target = _batch(8, B=2, H=16, W=24)
    loss, g = D.UNet(params, bf16_emulation=False).loss_and_grads(feat, target)
    y, tw, tb = _torch_unet(params, feat)
    out = torch.from_numpy(feat[..., :3].astype(np.float64)) + y
    tl = torch.mean((out - torch.from_numpy(target.astype(np.float64))) ** 2)
    tl.backward()
    assert abs(tl.item() - loss) < 1e-12
    for l, L in enumerate(D.LAYERS):
        wo, bo = D.OFFS[l]
        tg = tw[l].grad.numpy().transpose(2, 3, 1, 0).reshape(9, L[4], D.coutp(l))
        got = g[wo:wo + tg.size].reshape(tg.shape)
        assert np.abs(got - tg).max() <= 1e-9 * (np.abs(tg).max() + 1e-30), l
        assert np.abs(g[bo:bo + D.coutp(l)] - tb[l].grad.numpy()).max() <= 1e-9 * (np.abs(tb[l].grad.numpy()).max() + 1e-30)


def test_emulated_and_exact_gradients_agree_to_bf16_precision():
    """With the same (bf16) weights, rounding the stored activations and dZ to bf16 moves the
    gradients by a few percent at most: the emulation is a perturbation, not a different
    function."""
    params = D.init_params(9)
    pq = D.bf16(params)
    for l in range(len(D.LAYERS)):
        wo, bo = D.OFFS[l]
        pq[bo:bo + D.coutp(l)] = params[bo:bo + D.coutp(l)]
    feat, target = _batch(10, H=16, W=16)
    _, ge = D.UNet(params, True).loss_and_grads(feat, target)
    _, gx = D.UNet(pq, False).loss_and_grads(feat, target)
    for l in range(len(D.LAYERS)):
        wo, bo = D.OFFS[l]
        a, b = ge[wo:bo], gx[wo:bo]
        assert np.linalg.norm(a - b) <= 0.15 * np.linalg.norm(b), l
        assert a @ b >= 0.99 * np.linalg.norm(a) * np.linalg.norm(b), l
