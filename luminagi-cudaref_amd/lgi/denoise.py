"""Denoising pipeline on the GPU: render with per-pixel variance, UNet on log1p radiance,
variance-gated tile fallback to more samples, FP64 image metrics.

The plain tracer is both the fallback and the judge: an 8x8 tile whose denoised mean luminance
leaves the confidence band of the noisy tile mean (k standard errors, from the per-pixel
sample variance of the luminance samples) is rendered again with more samples, and its plain
estimate replaces the denoised one.
"""
import ctypes as C

import torch

from . import _abi
from ._abi import LgiError
from .metrics import image_metrics, psnr_from_mse  # noqa: F401  (re-exported)

LUM = (0.2126, 0.7152, 0.0722)


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def tonemap(ctx, hdr):
    """Reference sqrt tonemap (src/pathtracer.cu:413-429) of mean radiance HxWx3 -> uint8."""
    h, w, _ = hdr.shape
    out = torch.empty((h, w, 3), dtype=torch.uint8, device=hdr.device)
    ctx.tonemap(hdr.contiguous(), h * w, 1.0, out)
    return out


class Frame:
    """Device accumulators of one image: sums, Kahan terms, squared luminance."""

    def __init__(self, cam, device="cuda"):
        self.cam = cam
        n = cam.width * cam.height
        self.sum = torch.zeros(n * 3, device=device)
        self.comp = torch.zeros(n * 3, device=device)
        self.sumsq = torch.zeros(n, device=device)
        self.spp = torch.zeros((cam.height, cam.width), device=device)  # samples per pixel

    def render(self, ctx, spp, bounces=2, sample0=0, flags=0):
        ctx.accumulate(self.sum, self.comp, spp, bounces, flags=flags, sample0=sample0, d_sumsq=self.sumsq,
                       camera=self.cam)
        self.spp += spp

    def mean(self):
        h, w = self.cam.height, self.cam.width
        return self.sum.view(h, w, 3) / self.spp.unsqueeze(-1)

    def lum_se(self):
        """Standard error of the mean luminance per pixel (unbiased sample variance)."""
        h, w = self.cam.height, self.cam.width
        n = self.spp
        m = (self.mean() * torch.tensor(LUM, device=n.device)).sum(-1)
        var = (self.sumsq.view(h, w) / n - m * m) * n / (n - 1).clamp_min(1)
        return torch.sqrt(var.clamp_min(0) / n)


def denoise(model, hdr):
    """UNet on log1p radiance; hdr HxWx3 (H, W multiples of 16) -> denoised radiance."""
    with torch.no_grad():
        # log1p(max(hdr, 0)) in, max(expm1(y), 0) out, both inside the first and last kernels
        return model(hdr.unsqueeze(0).contiguous(), radiance=True)[0]


def radiance_input(png_u8, hdr_path=None):
    """HxWx3 float32 CUDA radiance for the denoiser from an 8-bit tonemapped image: the raw FP32
    HxWx3 sidecar that `cuda_ref --hdr` writes when one is given, otherwise the inverse of the
    renderer's sqrt tonemap (values the PNG clipped at 1.0 stay clipped)."""
    import numpy as np
    H, W = png_u8.shape[:2]
    if hdr_path:
        hdr = np.fromfile(hdr_path, dtype=np.float32)
        if hdr.size != H * W * 3:
            raise ValueError(f"{hdr_path}: {hdr.size} floats, expected {H}x{W}x3 = {H * W * 3}")
        return torch.from_numpy(hdr.reshape(H, W, 3)).cuda()
    return (torch.tensor(np.asarray(png_u8, dtype=np.uint8)).cuda().float() / 255.0) ** 2


def denoise_any_size(model, rad):
    """denoise() of an HxWx3 radiance image of any size: edge-replicated to a multiple of 16
    (four 2x pools), denoised and cropped back."""
    import torch.nn.functional as F
    H, W = rad.shape[:2]
    ph, pw = (16 - H % 16) % 16, (16 - W % 16) % 16
    x = F.pad(rad.permute(2, 0, 1).unsqueeze(0), (0, pw, 0, ph), mode="replicate")[0].permute(1, 2, 0)
    return denoise(model, x.contiguous())[:H, :W].contiguous()


class DenoiseGraph:
    """UNet inference at one fixed frame size, captured once in a HIP graph and replayed.

    Replays skip the per-layer launch overhead of the eager forward. The input is copied into
    a static buffer and the returned tensor is the graph's static output, overwritten by the
    next call (clone it to keep it). Convolution tile shapes are chosen during the warm-up,
    before capture."""

    def __init__(self, model, height, width):
        self.model = model
        self.x = torch.zeros((height, width, 3), device="cuda")
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(2):
                denoise(model, self.x)
        torch.cuda.current_stream().wait_stream(side)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.y = denoise(model, self.x)

    def __call__(self, hdr):
        self.x.copy_(hdr)
        self.graph.replay()
        return self.y


class FrameGraph:
    """One frame of the 8-spp + UNet path captured in a single HIP graph: render `spp` samples
    (from `sample0`) with the renderer `flags` names, its sums written over the accumulators
    (LGI_RENDER_SET_SUMS: no clearing pass, and no squared luminance, which only the variance
    gate reads), mean radiance, UNet, tonemap. A replay re-renders the same sample indices, so it
    reproduces the frame bit for bit; the outputs (`hdr`, `rgb`) are static tensors overwritten
    by each replay.

    The render needs its device buffers allocated and the UNet its convolution tiles chosen
    before capture, so the constructor runs the frame once eagerly first. The context is moved
    to the capture stream for the capture and back afterwards."""

    def __init__(self, ctx, model, cam, spp=8, bounces=2, sample0=0, flags=2):
        self.frame = Frame(cam)
        h, w = cam.height, cam.width
        args = (spp, bounces, sample0, flags)
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        prev = ctx.stream
        try:
            with torch.cuda.stream(side):
                ctx.set_stream(side.cuda_stream)
                for _ in range(2):
                    self._frame(ctx, model, h, w, *args)
            torch.cuda.current_stream().wait_stream(side)
            self.graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph):
                ctx.set_stream(torch.cuda.current_stream().cuda_stream)
                self.hdr, self.rgb = self._frame(ctx, model, h, w, *args)
        finally:
            ctx.set_stream(prev)

    def _frame(self, ctx, model, h, w, spp, bounces, sample0, flags):
        f = self.frame
        ctx.accumulate(f.sum, f.comp, spp, bounces, flags=flags | _abi.RENDER_SET_SUMS, sample0=sample0,
                       camera=f.cam)
        hdr = denoise(model, f.sum.view(h, w, 3) / float(spp))
        return hdr, tonemap(ctx, hdr)

    def replay(self):
        self.graph.replay()
        return self.hdr, self.rgb


def _tile_sum(x, th, tw):
    h, w = x.shape
    pad = torch.zeros((th * 8, tw * 8), dtype=x.dtype, device=x.device)
    pad[:h, :w] = x
    return pad.view(th, 8, tw, 8).sum(dim=(1, 3))


def gate_tiles(noisy_mean, se, denoised, k=4.0, rel_floor=0.02, abs_floor=1e-3):
    """Row-major indices of the 8x8 tiles whose denoised mean luminance leaves the confidence
    band of the noisy tile mean: |mean(den) - mean(noisy)| > k * SE_tile + floor, where SE_tile =
    sqrt(sum of per-pixel SE^2) / pixels and floor = rel_floor * mean(noisy) + abs_floor (the
    per-pixel variance from a few samples is itself noisy; the tile mean is not)."""
    lw = torch.tensor(LUM, device=noisy_mean.device)
    ln = (noisy_mean * lw).sum(-1)
    ld = (denoised * lw).sum(-1)
    h, w = ln.shape
    th, tw = (h + 7) // 8, (w + 7) // 8
    cnt = _tile_sum(torch.ones_like(ln), th, tw)
    mn = _tile_sum(ln, th, tw) / cnt
    md = _tile_sum(ld, th, tw) / cnt
    se_t = torch.sqrt(_tile_sum(se * se, th, tw)) / cnt
    bad = (md - mn).abs() > k * se_t + rel_floor * mn.abs() + abs_floor
    return torch.nonzero(bad.reshape(-1)).flatten().to(torch.int32)


def denoise_with_fallback(ctx, model, frame, extra_spp=248, bounces=2, sample0=1 << 22, k=4.0, rel_floor=0.02,
                          flags=2):
    """Denoise frame; re-render gated tiles with extra_spp more samples (sorted renderer on the
    listed tiles by default) and keep their plain mean. Returns (radiance HxWx3, number of
    fallback tiles)."""
    noisy = frame.mean()
    den = denoise(model, noisy)
    tiles = gate_tiles(noisy, frame.lum_se(), den, k=k, rel_floor=rel_floor)
    if tiles.numel() == 0:
        return den, 0
    cam = frame.cam
    ctx.render_tiles(tiles.cpu().numpy(), extra_spp, frame.sum, frame.comp, bounces, sample0=sample0,
                     d_sumsq=frame.sumsq, camera=cam, flags=flags)
    h, w = cam.height, cam.width
    th, tw = (h + 7) // 8, (w + 7) // 8
    flag = torch.zeros(th * tw, dtype=torch.bool, device=noisy.device)
    flag[tiles.long()] = True
    mask = flag.view(th, 1, tw, 1).expand(th, 8, tw, 8).reshape(th * 8, tw * 8)[:h, :w]
    frame.spp[mask] += extra_spp
    out = torch.where(mask.unsqueeze(-1), frame.mean(), den)
    return out, int(tiles.numel())
