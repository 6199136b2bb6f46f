"""Image metrics. The host restatement is pinned by the reference's held-out eval images and
their published numbers (docs/denoiser.md:30-32, README.md:103-105): noisy_8 / denoised /
clean_256 against gt_1024 give PSNR 34.52 / 42.04 / 49.57 dB and SSIM 0.815 / 0.964 / 0.991.
The GPU FP64 reductions are checked against the host restatement."""
import numpy as np
import pytest

from oracle import metrics_ref as M

PUBLISHED = {"noisy_8": (34.52, 0.815, 3.25), "denoised": (42.04, 0.964, 1.39), "clean_256": (49.57, 0.991, 0.51)}


def load(p):
    from PIL import Image
    return np.asarray(Image.open(p).convert("RGB"), dtype=np.uint8)


def test_metrics_reproduce_reference_published_numbers(reference_dir):
    gt = load(reference_dir / "ml" / "eval" / "gt_1024.png")
    for name, (p, s, m) in PUBLISHED.items():
        img = load(reference_dir / "ml" / "eval" / f"{name}.png")
        assert round(M.psnr(img, gt), 2) == p, name
        assert abs(M.ssim(img, gt) - s) < 1e-3, name   # published to 3 digits (0.8155 -> "0.815")
        assert round(M.mad(img, gt), 2) == m, name


def test_metric_closed_forms():
    rng = np.random.default_rng(0)
    a = rng.integers(0, 256, (40, 50, 3), dtype=np.uint8)
    assert M.psnr(a, a) == float("inf") and M.ssim(a, a) == pytest.approx(1.0) and M.mad(a, a) == 0.0
    b = a.astype(np.int16) + 3
    b = np.clip(b, 0, 255).astype(np.uint8)
    d = (b.astype(np.float64) - a).reshape(-1)
    assert M.psnr(a, b) == pytest.approx(20 * np.log10(255 / np.sqrt(np.mean(d ** 2))))
    # SSIM of a constant shift of a constant image: luminance term only
    c = np.full((20, 20, 3), 100, np.uint8)
    e = np.full((20, 20, 3), 110, np.uint8)
    C1 = (0.01 * 255) ** 2
    assert M.ssim(c, e) == pytest.approx((2 * 100 * 110 + C1) / (100 ** 2 + 110 ** 2 + C1))


@pytest.mark.gpu
def test_gpu_metrics_match_host(gpu):
    import torch
    from lgi import denoise
    rng = np.random.default_rng(1)
    a = rng.integers(0, 256, (96, 130, 3), dtype=np.uint8)
    b = np.clip(a.astype(np.int16) + rng.integers(-9, 10, a.shape), 0, 255).astype(np.uint8)
    mse, s, m = denoise.image_metrics(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda())
    assert 20 * np.log10(255 / np.sqrt(mse)) == pytest.approx(M.psnr(a, b), abs=1e-9)
    assert s == pytest.approx(M.ssim(a, b), abs=1e-12)
    assert m == pytest.approx(M.mad(a, b), abs=1e-12)
