"""Denoiser UNet on the project's HIP kernels (BF16 MFMA convolutions, FP32 accumulation).

Architecture (reference ml/train.py:37-75, in the configuration of this project): encoder
widths 32/64/128/256 with two 3x3 conv + ReLU per level and 2x2 max-pool between levels, a
512-wide bottleneck, decoder levels of bilinear 2x upsample + skip concat + two 3x3 conv +
ReLU, and a 1x1 conv to 3 channels added to the input (residual). Inputs and outputs are
log1p radiance, NHWC FP32 [N, H, W, 3]; H and W must be multiples of 16.

Every layer is a torch.autograd.Function whose forward and backward call liblgi through its C
ABI on the current PyTorch stream; activations are BF16 NHWC, master weights FP32.
"""
import ctypes as C

import torch

from . import _abi
from ._abi import LgiError

ENC = (32, 64, 128, 256)
BOTTLENECK = 512

WGRAD_SCRATCH_MAX = 64 << 20  # floats (256 MB) for weight-gradient partial sums


def _call(name, *args):
    st = getattr(_abi.lib(), name)(*args)
    if st != 0:
        raise LgiError(st, name)


def _p(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _check(t, dtype, channels=None):
    assert t.is_cuda and t.dtype == dtype and t.is_contiguous(), (t.device, t.dtype, t.is_contiguous())
    if channels is not None:
        assert t.shape[-1] == channels, (t.shape, channels)


# ------------------------------------------------------------------ tile-shape selection
CONV_VARIANTS = {0: "8x16 px, 32/64 ch", 1: "16x16 px, 64 ch, rb", 2: "32x16 px, 32 ch, rb",
                 3: "16x16 px, 32 ch, rb", 4: "32x16 px, 64 ch, rb", 5: "8x32 px, 64 ch, 32x32 MFMA",
                 6: "8x32 px, 32 ch, 32x32 MFMA", 7: "16x32 px, 32 ch, 32x32 MFMA",
                 8: "16x32 px, 64 ch, direct LDS loads", 9: "8x32 px, 64 ch, direct LDS loads",
                 10: "16x32 px, 32 ch, direct LDS loads", 11: "8x32 px, 64 ch, FP8 operands",
                 12: "4x32 px, 64 ch, FP8 operands", 13: "8x32 px, 32 ch, FP8 operands"}


def conv_tuning(enabled=True):
    """Time the tiling variants on the first call of each conv shape (default) or use a fixed
    choice."""
    _call("lgi_unet_conv_tuning", int(bool(enabled)))


def force_conv_variant(variant):
    """Every conv call uses `variant` (None restores per-shape selection)."""
    _call("lgi_unet_conv_force_variant", -1 if variant is None else int(variant))


def conv_tune_table():
    """[(shape dict, chosen variant, {variant: us per launch})] of the shapes timed so far."""
    import numpy as np
    n = C.c_int32()
    _call("lgi_unet_conv_tune_table", None, None, 0, C.byref(n))
    cap = n.value
    rows = np.zeros((max(cap, 1), 10), np.int32)
    us = np.zeros((max(cap, 1), len(CONV_VARIANTS)), np.float32)
    _call("lgi_unet_conv_tune_table", C.c_void_p(rows.ctypes.data), C.c_void_p(us.ctypes.data), cap, C.byref(n))
    keys = ("n", "h", "w", "c0", "c1", "cout", "cout_split", "kind", "relu")
    return [(dict(zip(keys, map(int, r[:9]))), int(r[9]), {v: float(t) for v, t in enumerate(u) if t > 0})
            for r, u in zip(rows[:min(cap, n.value)], us)]


# ------------------------------------------------------------------ ops
def conv3x3_fwd(src0, src1, w_bf16, bias, relu):
    n, h, w, c0 = src0.shape
    c1 = src1.shape[-1] if src1 is not None else 0
    cout = w_bf16.shape[0]
    _check(src0, torch.bfloat16)
    if src1 is not None:
        _check(src1, torch.bfloat16)
    out = torch.empty((n, h, w, cout), dtype=torch.bfloat16, device=src0.device)
    _call("lgi_unet_conv3x3_fwd", _p(src0), c0, _p(src1), c1, _p(w_bf16), _p(bias), cout, n, h, w, int(relu),
          _p(out), _stream())
    return out


def conv3x3_pool_fwd(src0, src1, w_bf16, bias, relu):
    """conv3x3_fwd and its 2x2 max-pool from one kernel: returns (y, maxpool2(y))."""
    n, h, w, c0 = src0.shape
    c1 = src1.shape[-1] if src1 is not None else 0
    cout = w_bf16.shape[0]
    _check(src0, torch.bfloat16)
    if src1 is not None:
        _check(src1, torch.bfloat16)
    out = torch.empty((n, h, w, cout), dtype=torch.bfloat16, device=src0.device)
    pooled = torch.empty((n, h // 2, w // 2, cout), dtype=torch.bfloat16, device=src0.device)
    _call("lgi_unet_conv3x3_pool_fwd", _p(src0), c0, _p(src1), c1, _p(w_bf16), _p(bias), cout, n, h, w, int(relu),
          _p(out), _p(pooled), _stream())
    return out, pooled


def up_phase_weights(w, cu):
    """Phase weights [4, cout, 9, cu] BF16 of conv3x3_up_fwd from FP32 weights w [cout, 9, cin]:
    the 3x3 convolution of a bilinear 2x upsample, rewritten per output phase as a 3x3
    convolution of the upsample's input (first cu input channels of w)."""
    w = w.detach().float().contiguous()
    cout, _, cin = w.shape
    out = torch.empty((4, cout, 9, cu), dtype=torch.bfloat16, device=w.device)
    _call("lgi_unet_up_phase_weights", _p(w), cout, cin, cu, _p(out), _stream())
    return out


def conv3x3_up_fwd(lo, skip, w_phase, w_bf16, bias, relu):
    """relu?(conv3x3(cat(upsample2x(lo), skip)) + bias) without the upsampled tensor: lo
    [N,h,w,cu], skip [N,2h,2w,cs] or None, w_bf16 [cout,9,cu+cs], w_phase from up_phase_weights."""
    n, h, w, cu = lo.shape
    _check(lo, torch.bfloat16)
    cs = 0
    if skip is not None:
        _check(skip, torch.bfloat16)
        assert skip.shape[:3] == (n, 2 * h, 2 * w), (skip.shape, lo.shape)
        cs = skip.shape[-1]
    cout = w_bf16.shape[0]
    assert w_bf16.shape == (cout, 9, cu + cs) and w_bf16.dtype == torch.bfloat16 and w_bf16.is_contiguous()
    assert w_phase.shape == (4, cout, 9, cu) and w_phase.dtype == torch.bfloat16 and w_phase.is_contiguous()
    out = torch.empty((n, 2 * h, 2 * w, cout), dtype=torch.bfloat16, device=lo.device)
    _call("lgi_unet_conv3x3_up_fwd", _p(lo), cu, _p(skip), cs, _p(w_phase), _p(w_bf16), _p(bias), cout, n, h, w,
          int(relu), _p(out), _stream())
    return out


E4M3_MAX = 448.0


def quantize_weight_e4m3(w):
    """FP32 weights [cout, 9, cin] -> (E4M3 codes as uint8 [cout, 9, cin], FP32 scale [cout]) with
    w ~= code * scale[co]: each output channel's largest magnitude maps to 448, rounding to nearest
    even (torch's float8_e4m3fn cast)."""
    w = w.detach().float()
    scale = (w.abs().amax(dim=(1, 2)) / E4M3_MAX).clamp_min(1e-30)
    q = (w / scale[:, None, None]).clamp(-E4M3_MAX, E4M3_MAX).to(torch.float8_e4m3fn)
    return q.view(torch.uint8).contiguous(), scale.contiguous()


def quantize_act_e4m3(x, act_scale):
    """The kernels' on-chip activation quantization, restated in torch: e4m3(clamp(x / s, +-448))."""
    return (x.float() / act_scale).clamp(-E4M3_MAX, E4M3_MAX).to(torch.float8_e4m3fn)


def conv3x3_fp8_fwd(src0, src1, wq, w_scale, act_scale, bias, relu, pool=False):
    """conv3x3_fwd (or conv3x3_pool_fwd with pool=True) on FP8 operands: inputs quantized with the
    per-tensor act_scale inside the kernel, wq/w_scale from quantize_weight_e4m3."""
    n, h, w, c0 = src0.shape
    c1 = src1.shape[-1] if src1 is not None else 0
    cout = wq.shape[0]
    _check(src0, torch.bfloat16)
    if src1 is not None:
        _check(src1, torch.bfloat16)
    assert wq.dtype == torch.uint8 and wq.is_contiguous() and w_scale.dtype == torch.float32
    out = torch.empty((n, h, w, cout), dtype=torch.bfloat16, device=src0.device)
    pooled = torch.empty((n, h // 2, w // 2, cout), dtype=torch.bfloat16, device=src0.device) if pool else None
    _call("lgi_unet_conv3x3_fp8_fwd", _p(src0), c0, _p(src1), c1, _p(wq), _p(w_scale), C.c_float(float(act_scale)),
          _p(bias), cout, n, h, w, int(relu), _p(out), _p(pooled), _stream())
    return (out, pooled) if pool else out


def conv3x3_first_fwd(x, w_bf16, bias, relu=True, radiance=False):
    """First layer straight from the FP32 3-channel input x [N,H,W,3] -> BF16 [N,H,W,32];
    radiance=True: x is radiance and the layer reads log1p(max(x, 0))."""
    n, h, w, c = x.shape
    _check(x, torch.float32, 3)
    assert w_bf16.shape[0] == 32 and w_bf16.dtype == torch.bfloat16 and w_bf16.is_contiguous()
    out = torch.empty((n, h, w, 32), dtype=torch.bfloat16, device=x.device)
    name = "lgi_unet_conv3x3_first_radiance_fwd" if radiance else "lgi_unet_conv3x3_first_fwd"
    _call(name, _p(x), _p(w_bf16), w_bf16.shape[2], _p(bias), n, h, w, int(relu), _p(out), _stream())
    return out


def conv3x3_tail_radiance_fwd(src0, w_bf16, bias, tail_w, tail_b, hdr, relu=True):
    """Last decoder convolution (32 channels) with the radiance-domain tail in its epilogue:
    max(expm1(log1p(max(hdr, 0)) + conv1x1(relu(conv3x3(src0)))), 0) without writing the features."""
    n, h, w, c0 = src0.shape
    _check(src0, torch.bfloat16)
    _check(hdr, torch.float32, 3)
    assert w_bf16.shape[0] == 32
    out = torch.empty((n, h, w, 3), dtype=torch.float32, device=src0.device)
    _call("lgi_unet_conv3x3_tail_radiance_fwd", _p(src0), c0, _p(w_bf16), _p(bias), n, h, w, int(relu),
          _p(tail_w.contiguous()), _p(tail_b.contiguous()), _p(hdr), _p(out), _stream())
    return out


def tail_radiance_fwd(feat, hdr, w, b):
    """Inference tail in the radiance domain: max(expm1(log1p(max(hdr, 0)) + conv1x1(feat)), 0)."""
    n, h, wd, c = feat.shape
    _check(feat, torch.bfloat16, 32)
    _check(hdr, torch.float32, 3)
    out = torch.empty((n, h, wd, 3), dtype=torch.float32, device=feat.device)
    _call("lgi_unet_tail_radiance_fwd", _p(feat), _p(hdr), _p(w.contiguous()), _p(b.contiguous()), n * h * wd,
          _p(out), _stream())
    return out


def conv3x3_bwd(dy, y_relu, w_bf16, src0, src1, need_dx=True):
    """Returns (dx0, dx1, dw fp32 [cout,9,cin], db fp32)."""
    n, h, w, cout = dy.shape
    c0 = src0.shape[-1]
    c1 = src1.shape[-1] if src1 is not None else 0
    cin = c0 + c1
    dy = dy.contiguous()
    dx0 = dx1 = None
    if need_dx:
        dx0 = torch.empty((n, h, w, c0), dtype=torch.bfloat16, device=dy.device)
        dx1 = torch.empty((n, h, w, c1), dtype=torch.bfloat16, device=dy.device) if c1 else None
        wt = torch.empty(cin * 9 * cout, dtype=torch.bfloat16, device=dy.device)
        _call("lgi_unet_conv3x3_bwd_data", _p(dy), _p(y_relu), _p(w_bf16), cout, cin, c0, n, h, w, _p(dx0), _p(dx1),
              _p(wt), _stream())
    dw = torch.empty((cout, 9, cin), dtype=torch.float32, device=dy.device)
    db = torch.empty(cout, dtype=torch.float32, device=dy.device)
    # the library's preferred workspace (its pixel splits fill the GPU), capped at 256 MB
    want = _abi.lib().lgi_unet_wgrad_scratch_floats(cout, cin, n, h, w)
    scratch = torch.empty(max(min(want, WGRAD_SCRATCH_MAX), cout * 9 * cin), dtype=torch.float32, device=dy.device)
    _call("lgi_unet_conv3x3_bwd_weight", _p(dy), _p(y_relu), cout, _p(src0), c0, _p(src1), c1, n, h, w, _p(dw),
          _p(db), _p(scratch), scratch.numel(), _stream())
    return dx0, dx1, dw, db


class Conv3x3Fn(torch.autograd.Function):
    """relu?(conv3x3(cat(src0, src1)) + b); weights FP32 [cout, 9, cin] cast to BF16 per call."""

    @staticmethod
    def forward(ctx, src0, src1, weight, bias, relu):
        wb = weight.to(torch.bfloat16).contiguous()
        y = conv3x3_fwd(src0, src1, wb, bias, relu)
        ctx.save_for_backward(src0, src1, wb, y)
        ctx.relu = relu
        ctx.has_src1 = src1 is not None
        return y

    @staticmethod
    def backward(ctx, dy):
        src0, src1, wb, y = ctx.saved_tensors
        need_dx = ctx.needs_input_grad[0] or ctx.needs_input_grad[1]
        dx0, dx1, dw, db = conv3x3_bwd(dy.to(torch.bfloat16), y if ctx.relu else None, wb, src0,
                                       src1 if ctx.has_src1 else None, need_dx)
        return dx0, dx1 if ctx.has_src1 else None, dw, db, None


class MaxPoolFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        n, h, w, c = x.shape
        _check(x, torch.bfloat16)
        out = torch.empty((n, h // 2, w // 2, c), dtype=torch.bfloat16, device=x.device)
        _call("lgi_unet_maxpool_fwd", _p(x), n, h, w, c, _p(out), _stream())
        ctx.save_for_backward(x)
        return out

    @staticmethod
    def backward(ctx, dout):
        (x,) = ctx.saved_tensors
        n, h, w, c = x.shape
        dout = dout.to(torch.bfloat16).contiguous()
        din = torch.empty_like(x)
        _call("lgi_unet_maxpool_bwd", _p(x), _p(dout), n, h, w, c, _p(din), _stream())
        return din


class UpsampleFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        n, h, w, c = x.shape
        _check(x, torch.bfloat16)
        out = torch.empty((n, 2 * h, 2 * w, c), dtype=torch.bfloat16, device=x.device)
        _call("lgi_unet_upsample_fwd", _p(x), n, h, w, c, _p(out), _stream())
        ctx.shape = (n, h, w, c)
        return out

    @staticmethod
    def backward(ctx, dout):
        n, h, w, c = ctx.shape
        dout = dout.to(torch.bfloat16).contiguous()
        din = torch.empty((n, h, w, c), dtype=torch.bfloat16, device=dout.device)
        _call("lgi_unet_upsample_bwd", _p(dout), n, h, w, c, _p(din), _stream())
        return din


class TailFn(torch.autograd.Function):
    """out = x + conv1x1(feat) (32 -> 3), FP32 output; x is the (non-differentiated) input."""

    @staticmethod
    def forward(ctx, feat, x, w, b):
        n, h, wd, c = feat.shape
        _check(feat, torch.bfloat16, 32)
        _check(x, torch.float32, 3)
        out = torch.empty((n, h, wd, 3), dtype=torch.float32, device=feat.device)
        _call("lgi_unet_tail_fwd", _p(feat), _p(x), _p(w.contiguous()), _p(b.contiguous()), n * h * wd, _p(out),
              _stream())
        ctx.save_for_backward(feat, w)
        return out

    @staticmethod
    def backward(ctx, dout):
        feat, w = ctx.saved_tensors
        n, h, wd, c = feat.shape
        npx = n * h * wd
        dout = dout.float().contiguous()
        dfeat = torch.empty_like(feat)
        dw = torch.empty((3, 32), dtype=torch.float32, device=feat.device)
        db = torch.empty(3, dtype=torch.float32, device=feat.device)
        blocks = min(1024, (npx + 255) // 256)  # the kernel's grid: per-block partial sums
        scratch = torch.empty(blocks * 99, dtype=torch.float32, device=feat.device)
        _call("lgi_unet_tail_bwd", _p(feat), _p(dout), _p(w.contiguous()), npx, _p(dfeat), _p(dw), _p(db), _p(scratch),
              scratch.numel(), _stream())
        return dfeat, dout, dw, db


def pack_input(x):
    n, h, w, _ = x.shape
    _check(x, torch.float32, 3)
    out = torch.empty((n, h, w, 32), dtype=torch.bfloat16, device=x.device)
    _call("lgi_unet_pack_input", _p(x), n * h * w, _p(out), _stream())
    return out


# ------------------------------------------------------------------ model
def conv_shapes():
    """(name, cin, cout) of every 3x3 convolution in forward order (cin of the first = 32 padded)."""
    shapes = []
    cin = 32
    for i, c in enumerate(ENC):
        shapes += [(f"enc{i}a", cin, c), (f"enc{i}b", c, c)]
        cin = c
    shapes += [("mid_a", ENC[-1], BOTTLENECK), ("mid_b", BOTTLENECK, BOTTLENECK)]
    up = BOTTLENECK
    for i in reversed(range(len(ENC))):
        c = ENC[i]
        shapes += [(f"dec{i}a", up + c, c), (f"dec{i}b", c, c)]
        up = c
    return shapes


class UNet(torch.nn.Module):
    """Denoiser on log1p radiance: forward(x [N,H,W,3] fp32) -> [N,H,W,3] fp32."""

    def __init__(self, seed=0, recompute=False):
        super().__init__()
        g = torch.Generator().manual_seed(seed)
        self.shapes = conv_shapes()
        self.weights = torch.nn.ParameterDict()
        self.biases = torch.nn.ParameterDict()
        for name, cin, cout in self.shapes:
            fan_in = 9 * (3 if name == "enc0a" else cin)
            w = torch.randn(cout, 9, cin, generator=g) * (2.0 / fan_in) ** 0.5  # He init
            if name == "enc0a":
                w[:, :, 3:] = 0.0  # padded input channels
            self.weights[name] = torch.nn.Parameter(w)
            self.biases[name] = torch.nn.Parameter(torch.zeros(cout))
        self.tail_w = torch.nn.Parameter(torch.randn(3, 32, generator=g) * (1.0 / 32) ** 0.5 * 0.1)
        self.tail_b = torch.nn.Parameter(torch.zeros(3))
        self.recompute = recompute
        self._wcache = {}  # name -> ((version, data_ptr), BF16 copy) for inference
        self._pcache = {}  # name -> ((version, data_ptr), phase weights of conv3x3_up_fwd)
        # decoder levels whose first convolution reads the upsample's input directly in inference
        # (conv3x3_up_fwd); on a 1728x864 frame it wins at level 0 only (143 us against 61 + 125 us
        # for the upsample and the convolution; the FP8 path is faster at the deeper levels)
        self.fused_up_levels = (0,)
        self._fp8 = {}     # name -> (E4M3 weights, weight scales, activation scale) when FP8 is on
        self._amax = None  # name -> largest input magnitude, while calibrating

    def _wb(self, name):
        """BF16 copy of a 3x3 weight, rebuilt only when the FP32 weight has changed."""
        w = self.weights[name]
        key = (w._version, w.data_ptr())
        hit = self._wcache.get(name)
        if hit is None or hit[0] != key:
            hit = (key, w.detach().to(torch.bfloat16).contiguous())
            self._wcache[name] = hit
        return hit[1]

    def _wp(self, name, cu):
        """Phase weights of conv3x3_up_fwd for a decoder level's first convolution, rebuilt only
        when the FP32 weight has changed."""
        w = self.weights[name]
        key = (w._version, w.data_ptr())
        hit = self._pcache.get(name)
        if hit is None or hit[0] != key:
            hit = (key, up_phase_weights(w, cu))
            self._pcache[name] = hit
        return hit[1]

    def conv_up(self, level, lo, skip):
        """relu(conv3x3(cat(upsample2x(lo), skip))) of decoder level `level` (inference): fused
        when the level is in fused_up_levels, else the upsample kernel and self.conv."""
        name = f"dec{level}a"
        if not torch.is_grad_enabled() and level in self.fused_up_levels and self._amax is None:
            return conv3x3_up_fwd(lo, skip, self._wp(name, lo.shape[-1]), self._wb(name), self.biases[name], True)
        return self.conv(name, UpsampleFn.apply(lo), skip)

    def conv(self, name, a, b=None, relu=True, pool=False):
        """relu(conv3x3(cat(a, b))), and with pool=True also its 2x2 max-pool (inference only)."""
        if not torch.is_grad_enabled():
            if self._amax is not None:
                m = a.abs().amax() if b is None else torch.maximum(a.abs().amax(), b.abs().amax())
                self._amax[name] = max(self._amax.get(name, 0.0), float(m))
            if name in self._fp8:
                wq, ws, sa = self._fp8[name]
                return conv3x3_fp8_fwd(a, b, wq, ws, sa, self.biases[name], relu, pool=pool)
            if pool:
                return conv3x3_pool_fwd(a, b, self._wb(name), self.biases[name], relu)
            return conv3x3_fwd(a, b, self._wb(name), self.biases[name], relu)
        assert not pool
        return Conv3x3Fn.apply(a, b, self.weights[name], self.biases[name], relu)

    def fp8_layers(self):
        """The convolutions that can run on FP8 operands: 64-channel input chunks of one source."""
        out = []
        for name, cin, cout in self.shapes:
            if name in ("enc0a", "enc1a"):
                continue  # 32-channel inputs
            srcs = (cin - ENC[int(name[3])], ENC[int(name[3])]) if name.endswith("a") and name.startswith("dec") \
                else (cin,)
            if all(c % 64 == 0 for c in srcs):
                out.append(name)
        return out

    @torch.no_grad()
    def enable_fp8(self, calib_x, margin=4.0, layers=None, radiance=False):
        """Switches the FP8-capable convolutions (fp8_layers(), or `layers`) to E4M3 operands: per
        output channel weight scales, and per-tensor activation scales of margin x the largest input
        magnitude a BF16 forward of calib_x produces (larger inputs saturate at 448 x scale)."""
        self._fp8 = {}
        self._amax = {}
        try:
            self.forward(calib_x, radiance=radiance)
        finally:
            amax, self._amax = self._amax, None
        for name in (layers if layers is not None else self.fp8_layers()):
            wq, ws = quantize_weight_e4m3(self.weights[name])
            sa = max(amax[name], 1e-6) * margin / E4M3_MAX
            self._fp8[name] = (wq.to(calib_x.device), ws.to(calib_x.device), sa)
        return dict(self._fp8)

    def disable_fp8(self):
        self._fp8 = {}

    def block(self, prefix, a, b=None):
        y = self.conv(prefix + "a", a, b)
        return self.conv(prefix + "b", y)

    def _block(self, prefix, a, b=None):
        if self.recompute and torch.is_grad_enabled():
            from torch.utils.checkpoint import checkpoint
            return checkpoint(self.block, prefix, a, b, use_reentrant=False)
        return self.block(prefix, a, b)

    def forward(self, x, taps=None, radiance=False):
        """taps (optional dict) receives every level's activation for per-layer error reports.
        radiance=True (inference only): x is radiance, and the log1p(max(x, 0)) input transform
        and the max(expm1(y), 0) output transform run inside the first layer and the tail."""
        x = x.contiguous()
        if radiance and torch.is_grad_enabled():
            raise ValueError("radiance=True is an inference path; run it under torch.no_grad()")
        skips = []
        for i in range(len(ENC)):
            if not torch.is_grad_enabled():
                # inference: the first conv reads the 3-channel input directly, and the second
                # conv of each level writes its max-pooled copy from the same kernel
                if i == 0:
                    y = conv3x3_first_fwd(x, self._wb("enc0a"), self.biases["enc0a"], radiance=radiance)
                else:
                    y = self.conv(f"enc{i}a", h)
                h, pooled = self.conv(f"enc{i}b", y, pool=True)
            else:
                h = self._block(f"enc{i}", pack_input(x) if i == 0 else h)
                pooled = None
            skips.append(h)
            if taps is not None:
                taps[f"enc{i}"] = h
            h = MaxPoolFn.apply(h) if pooled is None else pooled
        h = self._block("mid_", h)
        if taps is not None:
            taps["mid"] = h
        for i in reversed(range(len(ENC))):
            if radiance and i == 0 and taps is None:
                # the last convolution with the tail in its epilogue: dec0's output never leaves the chip
                y = self.conv_up(0, h, skips[0])
                return conv3x3_tail_radiance_fwd(y, self._wb("dec0b"), self.biases["dec0b"], self.tail_w, self.tail_b, x)
            if not torch.is_grad_enabled():
                h = self.conv(f"dec{i}b", self.conv_up(i, h, skips[i]))
            else:
                h = self._block(f"dec{i}", UpsampleFn.apply(h), skips[i])
            if taps is not None:
                taps[f"dec{i}"] = h
        if radiance:
            return tail_radiance_fwd(h, x, self.tail_w, self.tail_b)
        return TailFn.apply(h, x, self.tail_w, self.tail_b)

    def state_for_checkpoint(self):
        return {k: v.detach().cpu() for k, v in self.state_dict().items()}
