"""Image quality against a reference frame: PSNR and SSIM on 8-bit RGB.

SSIM follows Wang, Bovik, Sheikh and Simoncelli (IEEE TIP 2004): an 11-tap Gaussian window
with sigma 1.5, K1 = 0.01, K2 = 0.03, L = 255, per channel, averaged over the image and
the three channels.
"""
import numpy as np
from scipy.ndimage import gaussian_filter


def psnr(a, b):
    """Peak signal-to-noise ratio in dB of uint8 images a against b (inf when equal)."""
    d = np.asarray(a, np.float64) - np.asarray(b, np.float64)
    mse = float(np.mean(d * d))
    return float("inf") if mse == 0 else 10.0 * np.log10(255.0 * 255.0 / mse)


def ssim(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    c1, c2 = (0.01 * 255) ** 2, (0.03 * 255) ** 2
    vals = []
    for ch in range(a.shape[2]):
        x, y = a[..., ch], b[..., ch]
        f = lambda z: gaussian_filter(z, 1.5, truncate=3.5, mode="reflect")  # noqa: E731 - 11 taps
        mx, my = f(x), f(y)
        sxx, syy, sxy = f(x * x) - mx * mx, f(y * y) - my * my, f(x * y) - mx * my
        m = ((2 * mx * my + c1) * (2 * sxy + c2)) / ((mx * mx + my * my + c1) * (sxx + syy + c2))
        vals.append(m.mean())
    return float(np.mean(vals))
