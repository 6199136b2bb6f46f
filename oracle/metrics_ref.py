"""Host FP64 image metrics. TEST INFRASTRUCTURE ONLY.

psnr: tools/diff.py:25-29 (20 log10(255 / sqrt(MSE)) over uint8 RGB, MSE in float64).
ssim: the scikit-image default that tools/diff.py:100-101 calls
(structural_similarity(channel_axis=2, data_range=255)): 7x7 uniform window, sample covariance
(N / (N - 1)), K1 = 0.01, K2 = 0.03, mean of the SSIM map over the interior where the window
fits (3-pixel border cropped), averaged over channels. scikit-image is not installed here; the
restatement is pinned by the reference's published numbers for its eval PNGs.
mad: mean absolute difference (tools/diff.py:102).
heat_lut / heatmap: the difference heatmap of tools/diff.py:32-45 (a 256-entry colour table
from three clipped quadratics; per pixel the channel-mean absolute difference times a gain,
clipped to [0, 255] and truncated, indexes it).
"""
import numpy as np


def psnr(a, b):
    mse = np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2)
    return float("inf") if mse == 0 else float(20.0 * np.log10(255.0 / np.sqrt(mse)))


def _box7(x):
    """Mean over every 7x7 window that fits, via a 2D cumulative sum (float64)."""
    c = np.cumsum(np.cumsum(np.pad(x, ((1, 0), (1, 0))), axis=0), axis=1)
    s = c[7:, 7:] - c[:-7, 7:] - c[7:, :-7] + c[:-7, :-7]
    return s / 49.0


def ssim(a, b, data_range=255.0):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    C1, C2 = (0.01 * data_range) ** 2, (0.03 * data_range) ** 2
    cov_norm = 49.0 / 48.0
    vals = []
    for ch in range(a.shape[2]):
        x, y = a[..., ch], b[..., ch]
        ux, uy = _box7(x), _box7(y)
        vx = cov_norm * (_box7(x * x) - ux * ux)
        vy = cov_norm * (_box7(y * y) - uy * uy)
        vxy = cov_norm * (_box7(x * y) - ux * uy)
        S = ((2 * ux * uy + C1) * (2 * vxy + C2)) / ((ux * ux + uy * uy + C1) * (vx + vy + C2))
        vals.append(S.mean())
    return float(np.mean(vals))


def mad(a, b):
    return float(np.mean(np.abs(a.astype(np.int16) - b.astype(np.int16))))


def heat_lut():
    """tools/diff.py:32-38: x = linspace(0, 1, 256); r, g, b quadratics clipped to [0, 1]; x 255;
    truncated to uint8."""
    x = np.linspace(0.0, 1.0, 256)
    r = np.clip(-0.2 + 2.4 * x - 1.3 * x ** 2, 0, 1)
    g = np.clip(0.1 + 0.8 * x + 0.1 * x ** 2, 0, 1)
    b = np.clip(0.4 + 0.8 * x - 1.1 * x ** 2, 0, 1)
    return (np.stack([r, g, b], axis=1) * 255).astype(np.uint8)


def heatmap(a, b, gain=4.0):
    """tools/diff.py:41-45: channel mean of |a - b| (float64), times gain, clip, truncate, colour."""
    d = np.mean(np.abs(a.astype(np.int16) - b.astype(np.int16)), axis=2)
    d = np.clip(d * gain, 0, 255).astype(np.uint8)
    return heat_lut()[d]
