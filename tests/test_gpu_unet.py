"""Denoiser kernels against float64 CPU references (torch.nn.functional) and finite
differences. BF16 operands, FP32 accumulation: forward and data-gradient outputs are BF16
(relative rounding 2^-9), so the bar is stated as a relative error on the tensor's scale."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import unet_ref

pytestmark = pytest.mark.gpu

BF16_TOL = 5e-3   # max |err| / max |ref| for BF16 outputs of FP32-accumulated sums: 2.5x the 2^-9 output rounding


def rel(a, b):
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    return float((a - b).abs().max() / b.abs().max().clamp_min(1e-30))


def bf16_rand(*shape, scale=1.0, seed=0):
    g = torch.Generator().manual_seed(seed)
    return (torch.randn(*shape, generator=g) * scale).to(torch.bfloat16)


def ref_conv_nhwc(x0, x1, w, b, relu):
    x = x0.double() if x1 is None else torch.cat([x0.double(), x1.double()], -1)
    y = F.conv2d(x.permute(0, 3, 1, 2), unet_ref.torch_weight(w.double()), None if b is None else b.double(), padding=1)
    if relu:
        y = F.relu(y)
    return y.permute(0, 2, 3, 1)


CONV_CASES = [  # n, h, w, c0, c1, cout, relu
    (2, 16, 32, 64, 0, 64, True),
    (1, 12, 20, 32, 32, 32, True),      # ragged tiles + fused concat
    (1, 8, 16, 128, 64, 64, False),
    (2, 4, 6, 256, 0, 96, True),        # image smaller than one tile, Cout not a multiple of 64
    (1, 10, 45, 64, 64, 128, True),     # ragged 32-column tiles, concat, two 64-channel blocks
]


VARIANTS = [None] + list(range(11))     # lgi.unet.CONV_VARIANTS; None = per-shape selection
TN64 = (1, 4, 5, 8, 9)                  # tilings of 64 output channels
DIRECT = (8, 9, 10)                     # direct-to-LDS staging: no ReLU-masked input


@pytest.fixture(params=VARIANTS, ids=lambda v: "tuned" if v is None else f"variant{v}")
def conv_variant(request, gpu):
    """Every tiling variant forced in turn (None = per-shape selection by timing)."""
    from lgi import unet
    v = request.param
    cout = request.node.callspec.params["case"][5]
    if v in TN64 and cout % 64:
        pytest.skip("64-channel tiling needs cout % 64 == 0")
    if v in DIRECT and request.node.name.startswith("test_conv3x3_backward") and request.node.callspec.params["case"][6]:
        pytest.skip("the data gradient of a ReLU layer reads a masked input, which direct staging does not apply")
    unet.force_conv_variant(v)
    yield v
    unet.force_conv_variant(None)


@pytest.mark.parametrize("case", CONV_CASES, ids=lambda c: "x".join(map(str, c)))
def test_conv3x3_forward(conv_variant, case):
    from lgi import unet
    n, h, w, c0, c1, cout, relu = case
    x0 = bf16_rand(n, h, w, c0, seed=1)
    x1 = bf16_rand(n, h, w, c1, seed=2) if c1 else None
    wt = bf16_rand(cout, 9, c0 + c1, scale=(2.0 / (9 * (c0 + c1))) ** 0.5, seed=3)
    b = torch.randn(cout, generator=torch.Generator().manual_seed(4)) * 0.1
    y = unet.conv3x3_fwd(x0.cuda(), None if x1 is None else x1.cuda(), wt.cuda(), b.cuda(), relu)
    ref = ref_conv_nhwc(x0, x1, wt, b, relu)
    e = rel(y, ref)
    print(f"conv fwd {case} variant {conv_variant}: rel err {e:.2e}", flush=True)
    assert y.shape == ref.shape and e < BF16_TOL


@pytest.mark.parametrize("case", CONV_CASES, ids=lambda c: "x".join(map(str, c)))
def test_conv3x3_backward(conv_variant, case):
    from lgi import unet
    n, h, w, c0, c1, cout, relu = case
    x0 = bf16_rand(n, h, w, c0, seed=5)
    x1 = bf16_rand(n, h, w, c1, seed=6) if c1 else None
    wt = bf16_rand(cout, 9, c0 + c1, scale=(2.0 / (9 * (c0 + c1))) ** 0.5, seed=7)
    b = torch.randn(cout, generator=torch.Generator().manual_seed(8)) * 0.1
    dy = bf16_rand(n, h, w, cout, seed=9)
    y = unet.conv3x3_fwd(x0.cuda(), None if x1 is None else x1.cuda(), wt.cuda(), b.cuda(), relu)
    dx0, dx1, dw, db = unet.conv3x3_bwd(dy.cuda(), y if relu else None, wt.cuda(), x0.cuda(),
                                        None if x1 is None else x1.cuda())
    # float64 reference; the ReLU mask comes from the kernel's own BF16 output so that ties at
    # exactly zero are decided the same way
    xs = [x0.double().requires_grad_()] + ([x1.double().requires_grad_()] if c1 else [])
    wd = wt.double().requires_grad_()
    bd = b.double().requires_grad_()
    z = ref_conv_nhwc(xs[0], xs[1] if c1 else None, wd, bd, False)
    mask = (y.cpu().double() > 0).double() if relu else torch.ones_like(z)
    (z * mask * dy.double()).sum().backward()
    errs = {"dx0": rel(dx0, xs[0].grad), "dw": rel(dw, wd.grad), "db": rel(db, bd.grad)}
    if c1:
        errs["dx1"] = rel(dx1, xs[1].grad)
    print(f"conv bwd {case} variant {conv_variant}: {errs}", flush=True)
    assert errs["dx0"] < BF16_TOL and errs.get("dx1", 0) < BF16_TOL
    assert errs["dw"] < 1e-2 and errs["db"] < 1e-2   # FP32 outputs of BF16 products


def test_maxpool_forward_backward(gpu):
    from lgi import unet
    x = bf16_rand(2, 8, 12, 64, seed=10)
    x[0, 0, 0, :] = x[0, 0, 1, :]  # exact ties: gradient goes to the first maximum
    y = unet.MaxPoolFn.apply(x.cuda())
    ref = F.max_pool2d(x.double().permute(0, 3, 1, 2), 2).permute(0, 2, 3, 1)
    assert torch.equal(y.cpu().double(), ref)
    dy = bf16_rand(2, 4, 6, 64, seed=11)
    xd = x.double().requires_grad_()
    (F.max_pool2d(xd.permute(0, 3, 1, 2), 2).permute(0, 2, 3, 1) * dy.double()).sum().backward()
    xc = x.cuda().requires_grad_()
    unet.MaxPoolFn.apply(xc).backward(dy.cuda())
    assert torch.equal(xc.grad.cpu().double(), xd.grad)


def test_upsample_forward_backward(gpu):
    from lgi import unet
    x = bf16_rand(2, 5, 7, 32, seed=12)
    y = unet.UpsampleFn.apply(x.cuda())
    ref = unet_ref.up(x.double().permute(0, 3, 1, 2)).permute(0, 2, 3, 1)
    assert rel(y, ref) < 1e-2
    dy = bf16_rand(2, 10, 14, 32, seed=13)
    xd = x.double().requires_grad_()
    (unet_ref.up(xd.permute(0, 3, 1, 2)).permute(0, 2, 3, 1) * dy.double()).sum().backward()
    xc = x.cuda().requires_grad_()
    unet.UpsampleFn.apply(xc).backward(dy.cuda())
    assert rel(xc.grad, xd.grad) < 1e-2


def _up_exact(x):
    """The upsample's arithmetic in FP32, one operation at a time (the kernel is built without
    contraction): lin_src's source rows/columns and weights, then the blend, rounded to BF16."""
    n, h, w, c = x.shape
    xf = x.float().numpy()

    def src(d, size):
        s = (np.arange(d, dtype=np.float32) + np.float32(0.5)) * np.float32(0.5) - np.float32(0.5)
        s = np.maximum(s, np.float32(0.0))
        i0 = s.astype(np.int64)
        return i0, np.minimum(i0 + 1, size - 1), (s - i0.astype(np.float32)).astype(np.float32)

    y0, y1, ly = src(2 * h, h)
    x0, x1, lx = src(2 * w, w)
    ly = ly[None, :, None, None]
    lx = lx[None, None, :, None]
    one = np.float32(1.0)
    a, b = xf[:, y0][:, :, x0], xf[:, y0][:, :, x1]
    cc, d = xf[:, y1][:, :, x0], xf[:, y1][:, :, x1]
    v = (one - ly) * ((one - lx) * a + lx * b) + ly * ((one - lx) * cc + lx * d)
    return torch.from_numpy(v.astype(np.float32)).to(torch.bfloat16)


@pytest.mark.parametrize("shape", [(2, 9, 13, 32), (1, 1, 1, 8), (1, 2, 3, 16), (1, 54, 108, 64)],
                         ids=lambda s: "x".join(map(str, s)))
def test_upsample_forward_bit_exact(gpu, shape):
    """The upsample's output, bit for bit, against its arithmetic restated in FP32 on the CPU:
    interior threads (which share one 3 x 3 neighbourhood among their 4 outputs) and the border."""
    from lgi import unet
    x = bf16_rand(*shape, seed=14)
    y = unet.UpsampleFn.apply(x.cuda()).cpu()
    assert torch.equal(y.view(torch.int16), _up_exact(x).view(torch.int16))


UP_CASES = [  # n, h, w (upsample input), cu, cs, cout, relu
    (1, 8, 32, 64, 32, 32, True),      # one tile
    (2, 11, 45, 32, 32, 64, True),     # ragged tiles, two images, two output-channel blocks
    (1, 3, 5, 64, 64, 32, False),      # image smaller than one tile
    (1, 1, 1, 32, 32, 32, True),       # 2x2 output: every pixel on the border ring
    (1, 20, 40, 128, 0, 32, True),     # no skip tensor
]


@pytest.mark.parametrize("case", UP_CASES, ids=lambda c: "x".join(map(str, c)))
def test_conv3x3_up_matches_upsample_then_conv(gpu, case):
    """The fused decoder convolution (phase weights inside the image, explicitly upsampled
    neighbourhoods on the border ring) against the FP64 upsample + convolution, and against the
    two-kernel BF16 path; the ring pixels separately, since they take their own code path."""
    from lgi import unet
    n, h, w, cu, cs, cout, relu = case
    lo = bf16_rand(n, h, w, cu, seed=21)
    skip = bf16_rand(n, 2 * h, 2 * w, cs, seed=22) if cs else None
    wt = bf16_rand(cout, 9, cu + cs, scale=(2.0 / (9 * (cu + cs))) ** 0.5, seed=23)
    b = torch.randn(cout, generator=torch.Generator().manual_seed(24)) * 0.1
    wp = unet.up_phase_weights(wt.float().cuda(), cu)
    y = unet.conv3x3_up_fwd(lo.cuda(), None if skip is None else skip.cuda(), wp, wt.cuda(), b.cuda(), relu)
    up = unet_ref.up(lo.double().permute(0, 3, 1, 2)).permute(0, 2, 3, 1)
    ref = ref_conv_nhwc(up, skip, wt, b, relu)
    two = unet.conv3x3_fwd(unet.UpsampleFn.apply(lo.cuda()), None if skip is None else skip.cuda(), wt.cuda(),
                           b.cuda(), relu)
    ring = torch.zeros(2 * h, 2 * w, dtype=torch.bool)
    ring[0, :] = ring[-1, :] = ring[:, 0] = ring[:, -1] = True
    yc, tc = y.cpu(), two.cpu()
    errs = {"fp64": rel(yc, ref), "two_kernel": rel(yc, tc), "ring_fp64": rel(yc[:, ring], ref[:, ring]),
            "ring_two_kernel": rel(yc[:, ring], tc[:, ring])}
    print(f"conv up {case}: {errs}", flush=True)
    assert y.shape == ref.shape
    # each path is within BF16_TOL of the FP64 result, so they differ by up to twice that
    assert errs["fp64"] < BF16_TOL and errs["ring_fp64"] < BF16_TOL, errs
    assert errs["two_kernel"] < 2 * BF16_TOL and errs["ring_two_kernel"] < 2 * BF16_TOL, errs


def test_conv3x3_up_many_tiles_per_block(gpu):
    """More tiles than the persistent grid has blocks (one per CU), two output-channel groups:
    each block streams several tiles' chunks through its three buffers, across tile boundaries."""
    from lgi import unet
    n, h, w, cu, cs, cout = 2, 72, 272, 64, 32, 64   # 2 x 9 x 9 tiles x 2 channel groups = 324 items
    g = torch.Generator(device="cuda").manual_seed(31)
    lo = torch.randn(n, h, w, cu, device="cuda", generator=g).to(torch.bfloat16)
    skip = torch.randn(n, 2 * h, 2 * w, cs, device="cuda", generator=g).to(torch.bfloat16)
    wt = (torch.randn(cout, 9, cu + cs, device="cuda", generator=g) * (2.0 / (9 * (cu + cs))) ** 0.5).to(torch.bfloat16)
    b = torch.randn(cout, device="cuda", generator=g) * 0.1
    y = unet.conv3x3_up_fwd(lo, skip, unet.up_phase_weights(wt.float(), cu), wt, b, True)
    two = unet.conv3x3_fwd(unet.UpsampleFn.apply(lo), skip, wt, b, True)
    e = rel(y, two)
    print(f"conv up, many tiles per block: rel diff to the two-kernel path {e:.2e}", flush=True)
    assert e < 2 * BF16_TOL


def test_conv3x3_up_rejects_bad_shapes(gpu):
    from lgi import unet
    from lgi._abi import LgiError
    lo = torch.zeros((1, 4, 4, 48), dtype=torch.bfloat16, device="cuda")   # 48 channels: not a multiple of 32
    wt = torch.zeros((32, 9, 48), dtype=torch.bfloat16, device="cuda")
    wp = torch.zeros((4, 32, 9, 48), dtype=torch.bfloat16, device="cuda")
    with pytest.raises(LgiError):
        unet.conv3x3_up_fwd(lo, None, wp, wt, None, True)


def test_tail_forward_backward(gpu):
    from lgi import unet
    feat = bf16_rand(1, 8, 16, 32, seed=14)
    x = torch.rand(1, 8, 16, 3)
    w = torch.randn(3, 32) * 0.2
    b = torch.randn(3) * 0.1
    fc, wc, bc = feat.cuda().requires_grad_(), w.cuda().requires_grad_(), b.cuda().requires_grad_()
    out = unet.TailFn.apply(fc, x.cuda(), wc, bc)
    fd, wd, bd = feat.double().requires_grad_(), w.double().requires_grad_(), b.double().requires_grad_()
    ref = x.double() + torch.einsum("nhwc,oc->nhwo", fd, wd) + bd
    assert rel(out, ref) < 1e-5
    g = torch.randn(1, 8, 16, 3)
    out.backward(g.cuda())
    (ref * g.double()).sum().backward()
    assert rel(fc.grad, fd.grad) < BF16_TOL and rel(wc.grad, wd.grad) < 1e-5 and rel(bc.grad, bd.grad) < 1e-5


def make_model_and_input(seed=0, n=1, h=32, w=32):
    from lgi import unet
    model = unet.UNet(seed=seed).cuda()
    g = torch.Generator().manual_seed(seed + 100)
    x = torch.log1p(torch.rand(n, h, w, 3, generator=g) * 2.0)
    return model, x


def test_unet_forward_per_layer_error_vs_fp64(gpu):
    """Per-level error of the BF16 MFMA forward against the FP64 CPU forward of the same
    (BF16-representable) weights."""
    model, x = make_model_and_input(seed=1, n=2, h=32, w=64)
    with torch.no_grad():
        for p in model.parameters():
            p.copy_(p.to(torch.bfloat16).float())
    taps, rtaps = {}, {}
    with torch.no_grad():
        y = model(x.cuda(), taps=taps)
    ref = unet_ref.forward(model.state_dict(), x, taps=rtaps)
    report = {k: rel(taps[k], rtaps[k]) for k in rtaps}
    report["output"] = rel(y, ref)
    report["output_minus_input"] = rel(y - x.cuda(), ref - x.double())
    print("per-layer max-rel error vs fp64:", {k: f"{v:.2e}" for k, v in report.items()}, flush=True)
    assert all(v < 5e-2 for k, v in report.items() if k != "output_minus_input")
    assert report["output_minus_input"] < 5e-2


def test_unet_backward_vs_fp64_and_finite_differences(gpu):
    from lgi import unet
    model, x = make_model_and_input(seed=2, n=1, h=32, w=32)
    with torch.no_grad():
        for p in model.parameters():
            p.copy_(p.to(torch.bfloat16).float())
    g = torch.Generator().manual_seed(7)
    target = torch.randn(1, 32, 32, 3, generator=g)
    xc, tc = x.cuda(), target.cuda()
    out = model(xc)
    loss = (out * tc).sum()
    loss.backward()
    # Float64 autograd of the reference network, twice: (a) plain FP64 activations, (b) FP64
    # math with every stored activation rounded to BF16 (straight-through), as the GPU stores
    # them. Against (a) the BF16 network differs mostly where a BF16 activation lands on the
    # other side of a ReLU (few pixels at the 2x2 / 4x4 levels of a 32x32 crop), so (a) gets a
    # direction bar; (b) follows the same masks and is held to an L2 bar of a few BF16 ulps per
    # layer (the two computations round independently, so their activations drift apart by
    # about one BF16 ulp per layer over the 18 convolutions).
    named = dict(model.named_parameters())
    for bf16_acts, max_l2, min_cos in ((False, 0.25, 0.97), (True, 0.06, 0.998)):
        params = {k: v.detach().cpu().double().requires_grad_() for k, v in model.state_dict().items()}
        rloss = (unet_ref.forward(params, x, bf16_activations=bf16_acts) * target.double()).sum()
        rloss.backward()
        worst, l2 = {}, {}
        for k, p in params.items():
            gd = named[k].grad.detach().cpu().double()
            if k == "weights.enc0a":
                gd, pg = gd[:, :, :3], p.grad[:, :, :3]
            else:
                pg = p.grad
            worst[k] = rel(gd, pg)
            l2[k] = float((gd - pg).norm() / pg.norm().clamp_min(1e-30))
            cos = float((gd * pg).sum() / (gd.norm() * pg.norm()).clamp_min(1e-30))
            assert cos > min_cos, (bf16_acts, k, cos)
        tag = "bf16-activation fp64" if bf16_acts else "fp64"
        print(f"grad max-rel error vs {tag}:", {k: f"{v:.1e}" for k, v in sorted(worst.items())}, flush=True)
        print(f"grad rel-L2 error vs {tag}:", {k: f"{v:.1e}" for k, v in sorted(l2.items())}, flush=True)
        assert max(l2.values()) < max_l2, tag


def test_conv_gradients_by_finite_differences(gpu):
    """Central differences of the GPU forward kernel against the GPU gradient kernels, on a
    32x32 crop. Without ReLU, L(W) = sum(conv(X, W) * G) is linear in W and in X, so the
    central difference along a BF16 direction V with a power-of-two step is exact up to the
    BF16 rounding of the outputs; steps are taken from W = 0 (resp. X = 0) so every perturbed
    operand is exactly representable."""
    from lgi import unet
    n, h, w, c0, c1, cout = 1, 32, 32, 64, 32, 64
    X0, X1 = bf16_rand(n, h, w, c0, seed=20).cuda(), bf16_rand(n, h, w, c1, seed=21).cuda()
    Wt = bf16_rand(cout, 9, c0 + c1, scale=0.05, seed=22).cuda()
    G = bf16_rand(n, h, w, cout, seed=23).cuda()
    V = bf16_rand(cout, 9, c0 + c1, scale=0.05, seed=24).cuda()
    VX0, VX1 = bf16_rand(n, h, w, c0, seed=25).cuda(), bf16_rand(n, h, w, c1, seed=26).cuda()

    def L(x0, x1, wt):
        return float((unet.conv3x3_fwd(x0, x1, wt, None, False).double() * G.double()).sum())

    dx0, dx1, dw, _ = unet.conv3x3_bwd(G, None, Wt, X0, X1)
    eps = 2.0 ** -3
    zero_w = torch.zeros_like(Wt)
    fd_w = (L(X0, X1, (zero_w + eps * V.float()).to(torch.bfloat16)) -
            L(X0, X1, (zero_w - eps * V.float()).to(torch.bfloat16))) / (2 * eps)
    an_w = float((dw.double() * V.double()).sum())
    zx0, zx1 = torch.zeros_like(X0), torch.zeros_like(X1)
    fd_x = (L((eps * VX0.float()).to(torch.bfloat16), (eps * VX1.float()).to(torch.bfloat16), Wt) -
            L((-eps * VX0.float()).to(torch.bfloat16), (-eps * VX1.float()).to(torch.bfloat16), Wt)) / (2 * eps)
    an_x = float((dx0.double() * VX0.double()).sum() + (dx1.double() * VX1.double()).sum())
    print(f"finite differences: dW {fd_w:.6g} vs {an_w:.6g}; dX {fd_x:.6g} vs {an_x:.6g}", flush=True)
    # the forward's outputs are rounded to BF16 (2^-9 relative); summed against random G
    # that leaves ~0.2-0.4% on the difference quotient
    assert abs(fd_w - an_w) <= 1e-2 * abs(an_w)
    assert abs(fd_x - an_x) <= 1e-2 * abs(an_x)
    # pooling is piecewise linear: the four values of every window are distinct on a 2/16 grid,
    # so steps of 1/64 cannot change which element is the maximum
    g = torch.Generator().manual_seed(27)
    base = torch.randint(-20, 21, (1, 16, 16, 32, 1), generator=g).float()
    offs = torch.argsort(torch.rand(1, 16, 16, 32, 4, generator=g), dim=-1).float() * 2
    win = (base + offs) / 16.0                                  # [1, 16, 16, 32, 4] (dy, dx) row-major
    Xp = win.reshape(1, 16, 16, 32, 2, 2).permute(0, 1, 4, 2, 5, 3).reshape(1, 32, 32, 32)
    Xp = Xp.to(torch.bfloat16).cuda()
    Vp = torch.randint(-1, 2, (1, 32, 32, 32), generator=g).to(torch.bfloat16).cuda()
    Gp = bf16_rand(1, 16, 16, 32, seed=28).cuda()
    Xr = Xp.clone().requires_grad_()
    unet.MaxPoolFn.apply(Xr).backward(Gp)

    def Lp(x):
        return float((unet.MaxPoolFn.apply(x).double() * Gp.double()).sum())
    fd_p = (Lp((Xp.float() + Vp.float() / 64).to(torch.bfloat16)) -
            Lp((Xp.float() - Vp.float() / 64).to(torch.bfloat16))) * 32
    an_p = float((Xr.grad.double() * Vp.double()).sum())
    print(f"finite differences: maxpool {fd_p:.6g} vs {an_p:.6g}", flush=True)
    assert abs(fd_p - an_p) <= 1e-2 * max(1.0, abs(an_p))


def test_conv_variant_unsupported_shape_is_an_error(gpu):
    from lgi import unet
    from lgi._abi import LgiError
    x = torch.zeros(1, 8, 16, 32, dtype=torch.bfloat16, device="cuda")
    w = torch.zeros(96, 9, 32, dtype=torch.bfloat16, device="cuda")
    unet.force_conv_variant(1)
    try:
        with pytest.raises(LgiError, match="UNSUPPORTED"):
            unet.conv3x3_fwd(x, None, w, torch.zeros(96, device="cuda"), True)
    finally:
        unet.force_conv_variant(None)


def test_conv_tuning_table_records_each_shape_once(gpu):
    from lgi import unet
    unet.conv_tuning(True)
    x = bf16_rand(1, 40, 48, 64, seed=11).cuda()
    w = bf16_rand(64, 9, 64, scale=0.05, seed=12).cuda()
    b = torch.zeros(64, device="cuda")
    y1 = unet.conv3x3_fwd(x, None, w, b, True)
    y2 = unet.conv3x3_fwd(x, None, w, b, True)
    rows = [r for r in unet.conv_tune_table() if (r[0]["h"], r[0]["w"], r[0]["c0"], r[0]["cout"]) == (40, 48, 64, 64)]
    assert len(rows) == 1
    shape, chosen, times = rows[0]
    assert set(times) == set(range(11)) and chosen == min(times, key=times.get)
    assert torch.equal(y1, y2)


def test_first_layer_kernel_matches_packed_path_and_fp64(gpu):
    """The inference first layer (FP32 3-channel input read directly, one K=27 MFMA row per
    pixel) against the packed-input 3x3 conv and the FP64 reference, ragged tile edges."""
    from lgi import unet
    g = torch.Generator().manual_seed(21)
    x = torch.log1p(torch.rand(2, 20, 37, 3, generator=g) * 3.0)
    w = torch.zeros(32, 9, 32)
    w[:, :, :3] = torch.randn(32, 9, 3, generator=g) * (2.0 / 27) ** 0.5
    wb = w.to(torch.bfloat16)
    b = torch.randn(32, generator=g) * 0.1
    y1 = unet.conv3x3_first_fwd(x.cuda(), wb.cuda(), b.cuda(), True)
    y0 = unet.conv3x3_fwd(unet.pack_input(x.cuda()), None, wb.cuda(), b.cuda(), True)
    xb = torch.zeros(2, 20, 37, 32, dtype=torch.float64)
    xb[..., :3] = x.to(torch.bfloat16).double()
    ref = ref_conv_nhwc(xb, None, wb, b, True)
    e1, e0 = rel(y1, ref), rel(y0, ref)
    d = (y1.float() - y0.float()).abs().max().item() / y0.float().abs().max().item()
    print(f"first layer: rel err {e1:.2e} (packed path {e0:.2e}), max diff between paths {d:.2e}", flush=True)
    assert e1 < BF16_TOL and d < 1e-2


def test_first_layer_kernel_many_tiles_per_block(gpu):
    """A frame with more 8x16 tiles than the persistent first layer has blocks (each block walks
    several tiles, loading the next one's halo ahead), ragged in both directions, against an
    FP32 convolution of the same BF16-rounded operands."""
    from lgi import unet
    g = torch.Generator().manual_seed(22)
    n, h, w = 2, 517, 1030  # 2 x 65 x 65 = 8450 tiles
    x = torch.log1p(torch.rand(n, h, w, 3, generator=g) * 3.0)
    wt = torch.zeros(32, 9, 32)
    wt[:, :, :3] = torch.randn(32, 9, 3, generator=g) * (2.0 / 27) ** 0.5
    wb = wt.to(torch.bfloat16)
    b = torch.randn(32, generator=g) * 0.1
    y = unet.conv3x3_first_fwd(x.cuda(), wb.cuda(), b.cuda(), True).float().cpu()
    xin = x.to(torch.bfloat16).float().permute(0, 3, 1, 2)
    wk = wb.float()[:, :, :3].reshape(32, 3, 3, 3).permute(0, 3, 1, 2)  # [co][c][ky][kx]
    ref = torch.relu(torch.nn.functional.conv2d(xin, wk, b, padding=1)).permute(0, 2, 3, 1)
    err = ((y - ref).abs().max() / ref.abs().max()).item()
    print(f"first layer, {n}x{h}x{w}: rel err {err:.2e}", flush=True)
    assert err < BF16_TOL


POOL_CASES = [  # n, h, w, c0, c1, cout
    (2, 16, 32, 32, 0, 32),
    (1, 14, 46, 64, 0, 64),     # ragged tiles in both directions
    (1, 6, 70, 128, 0, 128),    # ragged 32-column tiles
    (1, 10, 18, 32, 32, 64),    # concatenated input
]


@pytest.mark.parametrize("variant", VARIANTS, ids=lambda v: "tuned" if v is None else f"variant{v}")
@pytest.mark.parametrize("case", POOL_CASES, ids=lambda c: "x".join(map(str, c)))
def test_conv3x3_pool_equals_conv_then_maxpool(gpu, variant, case):
    """The conv with its pooled copy (written from the output tile by the 32x32-MFMA tilings,
    by a separate pool launch after the others): y bit-identical to the plain conv with the
    same tiling, the pooled copy bit-identical to MaxPool2d(2) of y."""
    from lgi import unet
    n, h, w, c0, c1, cout = case
    if variant in TN64 and cout % 64:
        pytest.skip("64-channel tiling needs cout % 64 == 0")
    x0 = bf16_rand(n, h, w, c0, seed=31).cuda()
    x1 = bf16_rand(n, h, w, c1, seed=32).cuda() if c1 else None
    wt = bf16_rand(cout, 9, c0 + c1, scale=(2.0 / (9 * (c0 + c1))) ** 0.5, seed=33).cuda()
    b = (torch.randn(cout, generator=torch.Generator().manual_seed(34)) * 0.1).cuda()
    unet.conv_tuning(True)
    unet.force_conv_variant(variant)
    try:
        y, p = unet.conv3x3_pool_fwd(x0, x1, wt, b, True)
        if variant is None:
            chosen = [r for r in unet.conv_tune_table()
                      if (r[0]["h"], r[0]["w"], r[0]["c0"], r[0]["c1"], r[0]["cout"], r[0]["kind"]) ==
                      (h, w, c0, c1, cout, 2)]
            assert len(chosen) == 1 and chosen[0][1] == min(chosen[0][2], key=chosen[0][2].get)
            unet.force_conv_variant(chosen[0][1])
        y0 = unet.conv3x3_fwd(x0, x1, wt, b, True)
    finally:
        unet.force_conv_variant(None)
    p0 = unet.MaxPoolFn.apply(y0)
    assert p.shape == (n, h // 2, w // 2, cout)
    assert torch.equal(y, y0) and torch.equal(p, p0)
    assert rel(y, ref_conv_nhwc(x0.cpu(), None if x1 is None else x1.cpu(), wt.cpu(), b.cpu(), True)) < BF16_TOL


def test_conv3x3_pool_rejects_odd_sizes(gpu):
    from lgi import unet
    from lgi._abi import LgiError
    w = torch.zeros(32, 9, 32, dtype=torch.bfloat16, device="cuda")
    for h, wd in ((7, 16), (8, 15)):
        with pytest.raises(LgiError, match="INVALID_ARG"):
            unet.conv3x3_pool_fwd(torch.zeros(1, h, wd, 32, dtype=torch.bfloat16, device="cuda"), None, w,
                                  torch.zeros(32, device="cuda"), True)


def test_unet_inference_forward_matches_training_forward(gpu):
    """Inference runs the first-layer kernel and the pooled convs; training runs the plain
    layers. Both are BF16 paths of the same network, so they agree to BF16 rounding."""
    model, x = make_model_and_input(seed=5, n=1, h=48, w=80)
    x = x.cuda()
    with torch.no_grad():
        yi = model(x)
    with torch.enable_grad():
        yt = model(x).detach()
    e = rel(yi, yt)
    print(f"inference vs training forward: rel diff {e:.2e}", flush=True)
    assert e < 2e-2


def test_radiance_entry_points_match_torch_transforms(gpu):
    """The first layer reading log1p(max(hdr, 0)) and the tail writing max(expm1(.), 0) against
    torch's log1p/clamp_min/expm1 around the log-domain network, with negative radiance in the
    input (clamped) and a radiance range of 0..60."""
    from lgi import unet
    from lgi import denoise as D
    model, _ = make_model_and_input(seed=7)
    g = torch.Generator().manual_seed(8)
    hdr = (torch.rand(48, 80, 3, generator=g) ** 4 * 60.0)
    hdr[::7, ::5, 1] = -0.5
    hdr = hdr.cuda()
    with torch.no_grad():
        y_new = D.denoise(model, hdr)
        y_old = torch.expm1(model(torch.log1p(hdr.clamp_min(0)).unsqueeze(0).contiguous())[0]).clamp_min(0)
        f0 = unet.conv3x3_first_fwd(torch.log1p(hdr.clamp_min(0)).unsqueeze(0).contiguous(), model._wb("enc0a"),
                                    model.biases["enc0a"])
        f1 = unet.conv3x3_first_fwd(hdr.unsqueeze(0).contiguous(), model._wb("enc0a"), model.biases["enc0a"],
                                    radiance=True)
    e = rel(y_new, y_old)
    print(f"radiance path vs torch transforms: rel diff {e:.2e}, first layer equal {torch.equal(f0, f1)}", flush=True)
    assert bool((y_new >= 0).all()) and e < 1e-3
    assert torch.equal(f0, f1) or rel(f1, f0) < 1e-2
    with pytest.raises(ValueError):
        model(hdr.unsqueeze(0), radiance=True)


@pytest.mark.parametrize("variant", [None, 6, 7, 10], ids=lambda v: "tuned" if v is None else f"variant{v}")
@pytest.mark.parametrize("shape", [(1, 16, 64, 64), (2, 13, 45, 32)], ids=lambda s: "x".join(map(str, s)))
def test_conv3x3_tail_equals_conv_then_tail(gpu, variant, shape):
    """The last convolution with the radiance-domain tail in its epilogue: bit-identical to the
    same tiling's convolution followed by the tail kernel (ragged tiles, negative radiance)."""
    from lgi import unet
    n, h, w, c0 = shape
    x0 = bf16_rand(n, h, w, c0, seed=51).cuda()
    wt = bf16_rand(32, 9, c0, scale=(2.0 / (9 * c0)) ** 0.5, seed=52).cuda()
    b = (torch.randn(32, generator=torch.Generator().manual_seed(53)) * 0.1).cuda()
    g = torch.Generator().manual_seed(54)
    tw, tb = (torch.randn(3, 32, generator=g) * 0.1).cuda(), (torch.randn(3, generator=g) * 0.1).cuda()
    hdr = (torch.rand(n, h, w, 3, generator=g) * 3.0 - 0.2).cuda()
    unet.conv_tuning(True)
    unet.force_conv_variant(variant)
    try:
        out = unet.conv3x3_tail_radiance_fwd(x0, wt, b, tw, tb, hdr)
        if variant is None:
            chosen = [r for r in unet.conv_tune_table()
                      if (r[0]["h"], r[0]["w"], r[0]["c0"], r[0]["cout"], r[0]["kind"]) == (h, w, c0, 32, 3)]
            assert len(chosen) == 1 and chosen[0][1] in (6, 7, 10)
            unet.force_conv_variant(chosen[0][1])
        y = unet.conv3x3_fwd(x0, None, wt, b, True)
    finally:
        unet.force_conv_variant(None)
    ref = unet.tail_radiance_fwd(y, hdr, tw, tb)
    assert out.shape == (n, h, w, 3) and torch.equal(out, ref)


FP8_CASES = [  # n, h, w, c0, c1, cout, pool
    (1, 16, 64, 64, 0, 64, False),
    (2, 12, 40, 128, 64, 64, True),     # concat of two sources, ragged tiles, pooled copy
    (1, 9, 33, 64, 0, 96, False),       # Cout not a multiple of 64
    (1, 8, 32, 256, 256, 128, True),
]


@pytest.mark.parametrize("variant", [None, 11, 12, 13], ids=lambda v: "tuned" if v is None else f"variant{v}")
@pytest.mark.parametrize("case", FP8_CASES, ids=lambda c: "x".join(map(str, c)))
def test_conv3x3_fp8(gpu, variant, case):
    """FP8 (E4M3) convolution. Against the FP64 convolution of the same quantized operands (torch's
    float8_e4m3fn rounding of x / act_scale and of the per-channel scaled weights): only the FP32
    sums and the BF16 output round, so the BF16 bar holds. Against the unquantized FP64 reference:
    the 3-bit E4M3 mantissa bounds the error. The pooled copy is MaxPool2d(2) of the output."""
    from lgi import unet
    n, h, w, c0, c1, cout, pool = case
    if variant in (11, 12) and cout % 64:
        pytest.skip("64-channel tiling needs cout % 64 == 0")
    g = torch.Generator().manual_seed(61)
    x0 = (torch.rand(n, h, w, c0, generator=g) * 3.0).to(torch.bfloat16)
    x1 = (torch.rand(n, h, w, c1, generator=g) * 2.0).to(torch.bfloat16) if c1 else None
    wt = torch.randn(cout, 9, c0 + c1, generator=g) * (2.0 / (9 * (c0 + c1))) ** 0.5
    b = torch.randn(cout, generator=g) * 0.1
    sa = 3.0 * 2 / unet.E4M3_MAX
    wq, ws = unet.quantize_weight_e4m3(wt)
    unet.conv_tuning(True)
    unet.force_conv_variant(variant)
    try:
        r = unet.conv3x3_fp8_fwd(x0.cuda(), None if x1 is None else x1.cuda(), wq.cuda(), ws.cuda(), sa, b.cuda(), True,
                                 pool=pool)
    finally:
        unet.force_conv_variant(None)
    y, p = r if pool else (r, None)
    # FP64 of the quantized operands
    xq0 = unet.quantize_act_e4m3(x0, sa).double() * sa
    xq1 = unet.quantize_act_e4m3(x1, sa).double() * sa if c1 else None
    wdq = wq.view(torch.float8_e4m3fn).double() * ws.double()[:, None, None]
    ref_q = ref_conv_nhwc(xq0, xq1, wdq, b, True)
    ref = ref_conv_nhwc(x0, x1, wt, b, True)
    eq, e = rel(y, ref_q), rel(y, ref)
    print(f"fp8 conv {case} variant {variant}: vs quantized operands {eq:.2e}, vs FP64 {e:.2e}", flush=True)
    assert y.shape == ref.shape and eq < BF16_TOL and e < 8e-2
    if pool:
        assert torch.equal(p, unet.MaxPoolFn.apply(y))
