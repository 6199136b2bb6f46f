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


def test_heat_lut_closed_forms():
    """The heatmap colour table against exact rational evaluation of its quadratics: equal
    wherever the exact value is not within 1e-9 of an integer (where float64 may land on either
    side), and hand-evaluated endpoints and midpoint."""
    from fractions import Fraction as F
    lut = M.heat_lut()
    coef = [(F(-2, 10), F(24, 10), F(-13, 10)), (F(1, 10), F(8, 10), F(1, 10)), (F(4, 10), F(8, 10), F(-11, 10))]
    for i in range(256):
        x = F(i, 255)
        for c, (a0, a1, a2) in enumerate(coef):
            v = min(max(a0 + a1 * x + a2 * x * x, F(0)), F(1)) * 255
            if abs(v - round(v)) > F(1, 10 ** 9):
                assert lut[i, c] == v.numerator // v.denominator, (i, c)
            else:
                assert lut[i, c] in (round(v) - 1, round(v)), (i, c)
    assert lut[0].tolist() == [0, 25, 102]      # (clip(-0.2), 0.1, 0.4) * 255
    assert lut[255].tolist() == [229, 255, 25]  # (0.9, 1.0, 0.1) * 255
    assert lut[128].tolist() == [172, 134, 133]


def test_library_heat_lut_equals_oracle():
    from lgi import metrics
    assert np.array_equal(metrics.heat_lut(), M.heat_lut())


def test_heatmap_closed_forms():
    """Index = min(255, floor(mean |a - b| * gain)): pixel pairs with known channel sums."""
    lut = M.heat_lut()
    a = np.zeros((1, 7, 3), np.uint8)
    b = np.array([[[0, 0, 0], [3, 0, 0], [1, 0, 0], [2, 0, 0], [64, 64, 64], [63, 64, 64], [0, 255, 255]]], np.uint8)
    want = [0, 4, 1, 2, 255, 254, 255]   # 0, 3/3*4, 1/3*4, 2/3*4, clip(256), 191/3*4 = 254.67, clip
    assert np.array_equal(M.heatmap(a, b), lut[want][None])
    assert np.array_equal(M.heatmap(b, a), lut[want][None])  # symmetric
    assert np.array_equal(M.heatmap(a, b, gain=0.0), lut[[0] * 7][None])
    assert np.array_equal(M.heatmap(a, b, gain=1.0), lut[[0, 1, 0, 0, 64, 63, 170]][None])


@pytest.mark.gpu
@pytest.mark.parametrize("h,w", [(1, 1), (37, 53), (864, 1728)])
def test_gpu_heatmap_matches_host(gpu, h, w):
    import torch
    from lgi import metrics
    rng = np.random.default_rng(h * 7 + w)
    a = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    b = np.clip(a.astype(np.int16) + rng.integers(-80, 81, a.shape), 0, 255).astype(np.uint8)
    ta, tb = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    for gain in (4.0, 1.5, 0.0, -2.0, 1e3):
        got = metrics.heatmap(ta, tb, gain).cpu().numpy()
        assert np.array_equal(got, M.heatmap(a, b, gain)), gain


GOLDEN_HEAT = __import__("pathlib").Path(__file__).resolve().parent / "golden" / "heatmap_golden.npz"


def test_heatmap_oracle_reproduces_reference_output():
    """Crops of the reference's own diff-tool output (docs/heatmap.png made by tools/diff.py from
    docs/engine.png and docs/reference.png; tests/golden/make_heatmap_golden.py)."""
    d = np.load(GOLDEN_HEAT)
    for k in range(3):
        e, r, h = d[f"c{k}_engine"], d[f"c{k}_reference"], d[f"c{k}_heatmap"]
        assert np.array_equal(M.heatmap(e, r), h), k


def test_heatmap_oracle_reproduces_reference_output_full(reference_dir):
    e, r, h = (load(reference_dir / "docs" / f"{n}.png") for n in ("engine", "reference", "heatmap"))
    assert np.array_equal(M.heatmap(e, r), h)


@pytest.mark.gpu
def test_gpu_heatmap_reproduces_reference_output(gpu):
    import torch
    from lgi import metrics
    d = np.load(GOLDEN_HEAT)
    for k in range(3):
        e, r = (torch.from_numpy(d[f"c{k}_{n}"]).cuda() for n in ("engine", "reference"))
        assert np.array_equal(metrics.heatmap(e, r).cpu().numpy(), d[f"c{k}_heatmap"]), k


def test_metrics_reproduce_reference_diff_numbers(reference_dir):
    """README.md:57-59 reports the diff of docs/engine.png against docs/reference.png as PSNR
    21.8 dB, SSIM 0.631, mean abs diff 13.4 / 255."""
    e, r = (load(reference_dir / "docs" / f"{n}.png") for n in ("engine", "reference"))
    assert round(M.psnr(e, r), 1) == 21.8
    assert round(M.ssim(e, r), 3) == 0.631
    assert round(M.mad(e, r), 1) == 13.4


@pytest.mark.gpu
def test_gpu_read_probe_times_a_launch(gpu):
    """lgi_probe_read_bw with ms: the launch is bracketed by events and waited for; a pass over
    a larger buffer takes longer than one over a smaller buffer."""
    import ctypes as C
    import torch
    from lgi import _abi
    L = _abi.lib()
    sink = torch.zeros(4, dtype=torch.int32, device="cuda")
    warm = torch.ones(1 << 18, dtype=torch.int32, device="cuda")  # the first launch also loads the kernel
    assert L.lgi_probe_read_bw(C.c_void_p(warm.data_ptr()), warm.numel() * 4, 1, 64, C.c_void_p(sink.data_ptr()),
                               None, None) == 0
    torch.cuda.synchronize()
    times = []
    for mib in (16, 1024):
        buf = torch.ones(mib << 18, dtype=torch.int32, device="cuda")
        ms = C.c_float(-1.0)
        st = L.lgi_probe_read_bw(C.c_void_p(buf.data_ptr()), buf.numel() * 4, 2, 2048, C.c_void_p(sink.data_ptr()),
                                 None, C.byref(ms))
        assert st == 0 and ms.value > 0.0
        times.append(ms.value)
    assert times[1] > times[0]
