"""Image comparison on the GPU: FP64 PSNR/SSIM/MAD, the difference heatmap and the HTML report
of the reference's engine-diff tool (tools/diff.py:25-123).

The reductions and the heatmap run in liblgi (lgi_image_metrics, lgi_diff_heatmap) on device
copies of the two images; the host only decodes and encodes files and fills in the report.
"""
import base64
import ctypes as C
import html
import io
import math
import pathlib

import numpy as np
import torch

from . import _abi
from ._abi import LgiError


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _device_u8(img):
    """HxWx3 uint8 (numpy or tensor) -> contiguous uint8 CUDA tensor."""
    t = torch.as_tensor(img)
    assert t.dtype == torch.uint8 and t.dim() == 3 and t.shape[2] == 3, "expected HxWx3 uint8"
    return t.cuda().contiguous()


def image_metrics(a_u8, b_u8):
    """(MSE, SSIM, MAD) of two HxWx3 uint8 CUDA tensors, reduced in FP64 on the GPU."""
    assert a_u8.dtype == torch.uint8 and a_u8.shape == b_u8.shape and a_u8.is_cuda
    h, w, _ = a_u8.shape
    scratch = torch.empty(3 * 1024 + 8, dtype=torch.float64, device=a_u8.device)
    out = torch.empty(3, dtype=torch.float64, device=a_u8.device)
    st = _abi.lib().lgi_image_metrics(C.c_void_p(a_u8.contiguous().data_ptr()), C.c_void_p(b_u8.contiguous().data_ptr()),
                                      h, w, C.c_void_p(scratch.data_ptr()), C.c_void_p(out.data_ptr()), _stream())
    if st != 0:
        raise LgiError(st, "lgi_image_metrics")
    mse, ssim, mad = out.cpu().tolist()
    return mse, ssim, mad


def psnr_from_mse(mse):
    """20 log10(255 / sqrt(MSE)); infinite for identical images (tools/diff.py:25-29)."""
    return float("inf") if mse == 0 else 20.0 * math.log10(255.0 / math.sqrt(mse))


def heat_lut():
    """The heatmap's 256x3 uint8 colour table (lgi_heat_lut)."""
    out = np.zeros((256, 3), np.uint8)
    st = _abi.lib().lgi_heat_lut(out.ctypes.data_as(C.c_void_p))
    if st != 0:
        raise LgiError(st, "lgi_heat_lut")
    return out


def heatmap(a_u8, b_u8, gain=4.0):
    """HxWx3 uint8 CUDA tensor: the colour of min(255, mean |a - b| * gain) per pixel."""
    assert a_u8.dtype == torch.uint8 and a_u8.shape == b_u8.shape and a_u8.is_cuda and a_u8.shape[2] == 3
    h, w, _ = a_u8.shape
    a, b = a_u8.contiguous(), b_u8.contiguous()
    out = torch.empty_like(a)
    st = _abi.lib().lgi_diff_heatmap(C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()), h, w, float(gain),
                                     C.c_void_p(out.data_ptr()), _stream())
    if st != 0:
        raise LgiError(st, "lgi_diff_heatmap")
    return out


def compare(engine, ref, gain=4.0):
    """PSNR, SSIM, MAD and heatmap of two HxWx3 uint8 images of the same size."""
    a, b = _device_u8(engine), _device_u8(ref)
    if a.shape != b.shape:
        raise ValueError(f"image sizes differ: {tuple(a.shape)} vs {tuple(b.shape)}")
    mse, ssim, mad = image_metrics(a, b)
    heat = heatmap(a, b, gain).cpu().numpy()
    return {"psnr": psnr_from_mse(mse), "ssim": ssim, "mad": mad, "mse": mse, "heat": heat}


def png_bytes(img):
    from PIL import Image
    buf = io.BytesIO()
    Image.fromarray(np.ascontiguousarray(img)).save(buf, format="PNG")
    return buf.getvalue()


_REPORT = """<!doctype html>
<html><head><meta charset="utf-8"><title>{title}</title>
<style>
body {{ font-family: system-ui, sans-serif; margin: 24px; background: #1d1f21; color: #e8e8e8; }}
table.m td {{ padding: 3px 14px 3px 0; }}
table.m td:first-child {{ color: #9a9a9a; }}
.row {{ display: flex; gap: 12px; margin-top: 16px; }}
.row figure {{ flex: 1; margin: 0; }}
.row img {{ width: 100%; image-rendering: pixelated; border: 1px solid #444; }}
figcaption {{ color: #9a9a9a; font-size: 90%; margin-top: 4px; }}
</style></head><body>
<h1>{title}</h1>
<table class="m">
<tr><td>Resolution</td><td>{w}x{h}</td></tr>
<tr><td>PSNR</td><td>{psnr:.2f} dB</td></tr>
<tr><td>SSIM</td><td>{ssim:.4f}</td></tr>
<tr><td>Mean abs diff (per channel)</td><td>{mad:.2f} / 255</td></tr>
<tr><td>Engine</td><td>{engine}</td></tr>
<tr><td>Reference</td><td>{ref}</td></tr>
</table>
<div class="row">
<figure><img src="data:image/png;base64,{img_engine}"><figcaption>Engine</figcaption></figure>
<figure><img src="data:image/png;base64,{img_ref}"><figcaption>Reference (path tracer)</figcaption></figure>
<figure><img src="data:image/png;base64,{img_heat}"><figcaption>Abs diff &times; {gain:g}</figcaption></figure>
</div>
</body></html>
"""


def html_report(engine, ref, result, engine_path, ref_path, gain=4.0, title="Engine vs path-traced reference"):
    """Self-contained HTML page: the metrics table and the three images inlined as PNG."""
    h, w = ref.shape[:2]
    b64 = lambda img: base64.b64encode(png_bytes(img)).decode("ascii")  # noqa: E731
    return _REPORT.format(title=html.escape(title), w=w, h=h, psnr=result["psnr"], ssim=result["ssim"],
                          mad=result["mad"], gain=gain,
                          engine=html.escape(pathlib.Path(engine_path).resolve().as_posix()),
                          ref=html.escape(pathlib.Path(ref_path).resolve().as_posix()),
                          img_engine=b64(engine), img_ref=b64(ref), img_heat=b64(result["heat"]))


def read_rgb(path, size=None):
    """HxWx3 uint8 of an image file; with size=(w, h) and another size on disk, resampled to it
    with PIL's bilinear filter (the reference tool resizes the engine image this way)."""
    from PIL import Image
    with Image.open(path) as im:
        im = im.convert("RGB")
        if size is not None and im.size != tuple(size):
            im = im.resize(tuple(size), Image.BILINEAR)
        return np.array(im, dtype=np.uint8)


def report_line(r):
    """The diff tool's one-line summary (tools/diff.py's report format)."""
    return f"PSNR {r['psnr']:.2f} dB | SSIM {r['ssim']:.4f} | MAD {r['mad']:.2f}"


def diff_files(engine_path, ref_path, out_html, gain=4.0, warn=None):
    """Engine screenshot vs reference image on the GPU: metrics and heatmap of the two files (the
    engine image resampled to the reference's size when they differ), written as out_html and
    its <stem>.heatmap.png sidecar. Returns compare()'s dict plus the two images."""
    from .tracer import write_png
    ref = read_rgb(ref_path)
    engine = read_rgb(engine_path)
    if engine.shape != ref.shape:
        if warn:
            warn(f"resizing engine {engine.shape[1::-1]} -> {ref.shape[1::-1]}")
        engine = read_rgb(engine_path, size=ref.shape[1::-1])
    r = compare(engine, ref, gain=gain)
    out = pathlib.Path(out_html)
    out.parent.mkdir(parents=True, exist_ok=True)
    write_png(out.with_suffix(".heatmap.png"), r["heat"])
    out.write_text(html_report(engine, ref, r, engine_path, ref_path, gain=gain), encoding="utf-8")
    r["engine"], r["ref"] = engine, ref
    return r
