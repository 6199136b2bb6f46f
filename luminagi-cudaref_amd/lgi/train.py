"""Denoiser training on pairs rendered by the project's own tracer, and the time-to-quality
evaluation of 8 spp + UNet against the plain tracer.

Training recipe (reference ml/train.py:128-225, in this project's configuration): random
256x256 crops of (8-spp noisy, 256-spp clean) log1p radiance pairs, batch 16, L1 loss, Adam with
FP32 master weights, BF16 activations and gradients inside the kernels, dynamic loss scaling,
activations recomputed per UNet level; the best validation checkpoint is kept. Checkpoints
keep the reference keys {"model", "base", "in_ch"} and add {"arch", "optimizer", "loss_scale"}.
"""
import argparse
import json
import math
import pathlib
import time

import numpy as np
import torch

from . import denoise as D
from .poses import camera_for, orbit_poses
from .tracer import Context
from .unet import UNet

ARCH = "unet-relu-bilinear-32-64-128-256-512-log1p"


class LossScaler:
    """Dynamic loss scaling: halve on inf/nan gradients (step skipped), double every `interval`
    clean steps."""

    def __init__(self, scale=2.0 ** 12, interval=200, max_scale=2.0 ** 24):
        self.scale, self.interval, self.clean, self.max_scale = scale, interval, 0, max_scale

    def step(self, params, opt):
        grads = [p.grad for p in params if p.grad is not None]
        # one fused pass over all gradients: unscale in place (exact, the scale is a power of
        # two) and flag any inf/nan; one host synchronization per step
        found = torch.zeros(1, device=grads[0].device)
        inv = torch.full((1,), 1.0 / self.scale, device=grads[0].device)
        torch._amp_foreach_non_finite_check_and_unscale_(grads, found, inv)
        if found.item():
            self.scale /= 2.0
            self.clean = 0
            opt.zero_grad(set_to_none=True)
            return False
        opt.step()
        opt.zero_grad(set_to_none=True)
        self.clean += 1
        if self.clean % self.interval == 0:
            self.scale = min(self.scale * 2.0, self.max_scale)
        return True


def make_scene(kind=1, seed=1, width=1728, height=864):
    ctx = Context()
    ctx.generate(kind, seed, width, height)
    ctx.build_bvh()
    return ctx


def render_mean(ctx, cam, spp, sample0, bounces=2, batch=None):
    """Mean radiance HxWx3 (CUDA) of spp samples starting at sample0."""
    f = D.Frame(cam)
    step = batch or spp
    done = 0
    while done < spp:
        s = min(step, spp - done)
        f.render(ctx, s, bounces=bounces, sample0=sample0 + done)
        done += s
    return f.mean().contiguous()


def make_dataset(ctx, n_poses=50, noisy_spp=8, clean_spp=256, seed=0, bounces=2):
    """(noisy, clean) log1p radiance pairs [n, H, W, 3] on the GPU from orbit poses."""
    base = ctx.camera()
    noisy, clean = [], []
    for pos, tgt, _ in orbit_poses(n_poses, seed=seed):
        cam = camera_for(base, pos, tgt)
        noisy.append(torch.log1p(render_mean(ctx, cam, noisy_spp, 0, bounces)))
        clean.append(torch.log1p(render_mean(ctx, cam, clean_spp, 1 << 16, bounces, batch=64)))
    return torch.stack(noisy), torch.stack(clean)


def random_crops(noisy, clean, batch, crop, gen):
    n, h, w, _ = noisy.shape
    idx = torch.randint(0, n, (batch,), generator=gen)
    ys = torch.randint(0, h - crop + 1, (batch,), generator=gen)
    xs = torch.randint(0, w - crop + 1, (batch,), generator=gen)
    flips = torch.randint(0, 2, (batch,), generator=gen)
    xb, yb = [], []
    for i, y, x, f in zip(idx.tolist(), ys.tolist(), xs.tolist(), flips.tolist()):
        a = noisy[i, y:y + crop, x:x + crop]
        b = clean[i, y:y + crop, x:x + crop]
        if f:
            a, b = a.flip(1), b.flip(1)
        xb.append(a)
        yb.append(b)
    return torch.stack(xb).contiguous(), torch.stack(yb).contiguous()


def train(noisy, clean, steps=2000, batch=16, crop=256, lr=1e-4, seed=42, val_frac=0.2, time_budget=None,
          recompute=True, cosine=False, log=print):
    """Adam on L1(log1p) with dynamic loss scaling; with cosine=True the learning rate follows a
    cosine from lr down to lr/20 over `steps`. Returns (model at best validation, checkpoint, s)."""
    n = noisy.shape[0]
    perm = np.random.RandomState(seed).permutation(n)  # seeded split (ml/train.py:152-158)
    n_val = max(1, int(n * val_frac))
    val_idx, tr_idx = perm[:n_val].tolist(), perm[n_val:].tolist()
    tr_n, tr_c = noisy[tr_idx], clean[tr_idx]
    torch.manual_seed(seed)
    model = UNet(seed=seed, recompute=recompute).cuda()
    params = list(model.parameters())
    opt = torch.optim.Adam(params, lr=lr)
    scaler = LossScaler()
    gen = torch.Generator().manual_seed(seed)
    best, best_state, hist = float("inf"), None, []
    t0 = time.perf_counter()
    for it in range(steps):
        xb, yb = random_crops(tr_n, tr_c, batch, crop, gen)
        loss = (model(xb) - yb).abs().mean()
        (loss * scaler.scale).backward()
        if cosine:
            for g in opt.param_groups:
                g["lr"] = lr * (0.05 + 0.475 * (1.0 + math.cos(math.pi * it / steps)))
        scaler.step(params, opt)
        if (it + 1) % 100 == 0 or it + 1 == steps:
            with torch.no_grad():
                vl = float(np.mean([(model(noisy[i:i + 1]) - clean[i:i + 1]).abs().mean().item() for i in val_idx]))
            hist.append((it + 1, float(loss), vl))
            tag = ""
            if vl < best:
                best, tag = vl, "  *"
                best_state = {k: v.detach().clone() for k, v in model.state_dict().items()}
            log(f"step {it + 1:5d}/{steps}  train={float(loss):.4f}  val={vl:.4f}  scale={scaler.scale:g}{tag}")
        if time_budget and time.perf_counter() - t0 > time_budget:
            log(f"time budget reached after {it + 1} steps")
            break
    model.load_state_dict(best_state)
    ckpt = {"model": {k: v.cpu() for k, v in best_state.items()}, "base": 32, "in_ch": 3, "arch": ARCH,
            "optimizer": opt.state_dict(), "loss_scale": scaler.scale, "best_val": best, "history": hist}
    return model, ckpt, time.perf_counter() - t0


def export_checkpoint(ckpt):
    """Inference checkpoint: reference keys {"model", "base", "in_ch"} plus "arch"; 3x3 conv
    weights stored as BF16 (the convolution kernels round them to BF16 anyway, so the network's
    output is unchanged), biases and the 1x1 tail in FP32; no optimizer state."""
    model = {k: (v.to(torch.bfloat16) if k.startswith("weights.") else v) for k, v in ckpt["model"].items()}
    return {"model": model, "base": ckpt["base"], "in_ch": ckpt["in_ch"], "arch": ckpt["arch"],
            "best_val": ckpt.get("best_val")}


def load_model(path):
    ckpt = torch.load(path, map_location="cpu", weights_only=True)
    if ckpt.get("arch") != ARCH:
        raise ValueError(f"checkpoint arch {ckpt.get('arch')!r} is not {ARCH!r}")
    model = UNet().cuda()
    model.load_state_dict(ckpt["model"])
    return model.eval()


def timed(fn):
    torch.cuda.synchronize()
    t = time.perf_counter()
    out = fn()
    torch.cuda.synchronize()
    return out, (time.perf_counter() - t) * 1e3


def timed_median(fn, reps=5):
    """timed() of a repeatable call (same inputs, same outputs: a graph replay, an eager
    forward): the median of reps calls, so that one call's launch jitter does not decide a
    few-millisecond figure. Returns the last call's output."""
    out, times = None, []
    for _ in range(reps):
        out, ms = timed(fn)
        times.append(ms)
    return out, sorted(times)[len(times) // 2]


def plain_spp_for(pts, psnr):
    """spp at which the plain curve reaches `psnr`, from MSE ~ 1/spp between the two bracketing
    measurements (None when the curve does not bracket it)."""
    if pts and pts[0]["psnr"] >= psnr:
        return pts[0]["spp"] * 10 ** ((psnr - pts[0]["psnr"]) / 10.0)
    for a, b in zip(pts, pts[1:]):
        if a["psnr"] < psnr <= b["psnr"]:
            return a["spp"] * 10 ** ((psnr - a["psnr"]) / 10.0)
    return None


RENDERERS = {0: "inline", 2: "sorted", 2 | 8: "wavefront without ray sort"}


def pick_renderer(ctx, cam, bounces=2, spp=8):
    """The fastest renderer for this scene and camera: (flags, {flags: ms}). All three give the
    same bits: the inline megakernel, the wavefront renderer with its ray sort between bounces,
    and the wavefront renderer without it (LGI_RENDER_NO_RAY_SORT). Which one wins depends on
    the scene and the view. Times spp samples with each after a warm-up."""
    times = {}
    for fl in RENDERERS:
        D.Frame(cam).render(ctx, spp, bounces, sample0=1 << 25, flags=fl)
        f = D.Frame(cam)
        _, times[fl] = timed(lambda: f.render(ctx, spp, bounces, sample0=1 << 25, flags=fl))
    return min(times, key=times.get), times


def fp8_calibration_frame(ctx, spp=8, bounces=2):
    """Radiance (1, H, W, 3) of a training pose at 8 spp: the frame the FP8 activation scales are
    calibrated on (the first orbit pose of the training set, not the evaluated view)."""
    (pos, tgt, _), = orbit_poses(1, seed=0)
    return render_mean(ctx, camera_for(ctx.camera(), pos, tgt), spp, 1 << 26, bounces).unsqueeze(0).contiguous()


def evaluate(ctx, model, cam, ref_spp=4096, bounces=2, fallback=True, flags=None, fp8=True, log=print):
    """PSNR/SSIM (8-bit sqrt-tonemapped domain, as tools/diff.py) of plain renders at growing spp
    and of 8 spp + UNet against a ref_spp plain render; times to 40 dB. flags selects the
    renderer used for the timed renders (None: the faster one on this scene, pick_renderer;
    both give the same bits). fp8=True also measures the UNet on FP8 operands (activation scales
    calibrated on a training pose, fp8_calibration_frame). Every method that reaches 40 dB is
    timed; "unet_ms" is the fastest of them. The plain renders and the fallback add samples to
    their frames, so each is timed once; the repeatable graph replays and forwards report the
    median of 5 calls."""
    gt = D.tonemap(ctx, render_mean(ctx, cam, ref_spp, 0, bounces, batch=256))
    test0 = 1 << 24  # sample streams disjoint from the reference's
    res = {"ref_spp": ref_spp, "plain": []}
    if flags is None:
        flags, rt = pick_renderer(ctx, cam, bounces)
        res["renderer_ms_8spp"] = {RENDERERS[k]: v for k, v in rt.items()}
    res["renderer"] = RENDERERS.get(flags, str(flags))
    # plain curve: progressive accumulation, timed per added chunk; a throw-away render at the
    # largest chunk first, so that no timed chunk pays for growing the renderer's buffers
    D.Frame(cam).render(ctx, 512, bounces, sample0=test0 + (1 << 25), flags=flags)
    f = D.Frame(cam)
    total_ms, spp = 0.0, 0
    for target in (8, 16, 32, 64, 128, 256, 512, 1024):
        _, ms = timed(lambda: f.render(ctx, target - spp, bounces, sample0=test0 + spp, flags=flags))
        total_ms += ms
        spp = target
        mse, ssim, mad = D.image_metrics(D.tonemap(ctx, f.mean()), gt)
        res["plain"].append({"spp": spp, "ms": total_ms, "psnr": D.psnr_from_mse(mse), "ssim": ssim, "mad": mad})
        log(f"plain {spp:5d} spp: {D.psnr_from_mse(mse):6.2f} dB  ssim {ssim:.4f}  {total_ms:8.1f} ms")
    ms_per_spp = res["plain"][-1]["ms"] / 1024
    # 8 spp + UNet (+ gated fallback)
    fr = D.Frame(cam)
    _, t_render = timed(lambda: fr.render(ctx, 8, bounces, sample0=test0 + (1 << 20), flags=flags))
    D.denoise(model, fr.mean())  # warm-up: first-call allocations are not part of the steady state
    den, t_den = timed_median(lambda: D.denoise(model, fr.mean()))
    graph = D.DenoiseGraph(model, cam.height, cam.width)
    graph(fr.mean())
    den_g, t_den_g = timed_median(lambda: graph(fr.mean()))
    assert torch.equal(den_g, den), "graph replay differs from the eager forward"
    mse, ssim, mad = D.image_metrics(D.tonemap(ctx, den), gt)
    res["unet"] = {"spp": 8, "render_ms": t_render, "denoise_ms": t_den_g, "denoise_eager_ms": t_den,
                   "psnr": D.psnr_from_mse(mse), "ssim": ssim}
    # the whole 8-spp + UNet frame (clear, render, sort, accumulate, UNet, tonemap) as one graph
    fg = D.FrameGraph(ctx, model, cam, spp=8, bounces=bounces, sample0=test0 + (1 << 20), flags=flags)
    fg.replay()
    (hdr_g, _), t_frame = timed_median(fg.replay)
    assert torch.equal(hdr_g, den), "frame graph differs from the eager render + denoise"
    res["unet"]["frame_graph_ms"] = t_frame
    mse8, ssim8, _ = D.image_metrics(D.tonemap(ctx, fr.mean()), gt)
    res["noisy8"] = {"psnr": D.psnr_from_mse(mse8), "ssim": ssim8}
    log(f"8 spp raw: {D.psnr_from_mse(mse8):.2f} dB; 8 spp + UNet: {res['unet']['psnr']:.2f} dB ssim {ssim:.4f} "
        f"({t_render:.1f} + {t_den_g:.1f} ms; eager denoise {t_den:.1f} ms; one-graph frame {t_frame:.1f} ms)")
    if fp8:
        # the same 8-spp frame through the UNet on FP8 operands (before the fallback below adds
        # samples to some of its tiles)
        model.enable_fp8(fp8_calibration_frame(ctx, bounces=bounces), radiance=True)
        try:
            D.denoise(model, fr.mean())
            den8, t8 = timed_median(lambda: D.denoise(model, fr.mean()))
            g8 = D.DenoiseGraph(model, cam.height, cam.width)
            g8(fr.mean())
            den8_g, t8_g = timed_median(lambda: g8(fr.mean()))
            fg8 = D.FrameGraph(ctx, model, cam, spp=8, bounces=bounces, sample0=test0 + (1 << 20), flags=flags)
            fg8.replay()
            (hdr8, _), t_frame8 = timed_median(fg8.replay)
            assert torch.equal(den8_g, den8) and torch.equal(hdr8, den8), "FP8 graphs differ from the eager forward"
        finally:
            model.disable_fp8()
        mse, ssim, _ = D.image_metrics(D.tonemap(ctx, den8), gt)
        res["unet_fp8"] = {"spp": 8, "render_ms": t_render, "denoise_ms": t8_g, "denoise_eager_ms": t8,
                           "frame_graph_ms": t_frame8, "psnr": D.psnr_from_mse(mse), "ssim": ssim,
                           "ms": min(t_render + t8_g, t_frame8)}
        log(f"8 spp + UNet (FP8): {res['unet_fp8']['psnr']:.2f} dB ssim {ssim:.4f} "
            f"({t_render:.1f} + {t8_g:.2f} ms; one-graph frame {t_frame8:.2f} ms)")
    if fallback:
        warm = D.Frame(cam)  # warm-up of the tile path on a throw-away frame
        warm.render(ctx, 8, bounces, sample0=test0 + (1 << 22), flags=flags)
        D.denoise_with_fallback(ctx, model, warm, bounces=bounces, sample0=test0 + (1 << 23))
        (out, nt), t_fb = timed(lambda: D.denoise_with_fallback(ctx, model, fr, bounces=bounces,
                                                                sample0=test0 + (1 << 21)))
        mse, ssim, _ = D.image_metrics(D.tonemap(ctx, out), gt)
        res["unet_fallback"] = {"tiles": nt, "ms": t_render + t_fb, "psnr": D.psnr_from_mse(mse), "ssim": ssim}
        log(f"8 spp + UNet + gated fallback ({nt} tiles): {D.psnr_from_mse(mse):.2f} dB ({t_render + t_fb:.1f} ms)")
    spp40 = plain_spp_for(res["plain"], 40.0)
    res["plain_spp_40db"] = spp40
    res["plain_ms_40db"] = None if spp40 is None else spp40 * ms_per_spp
    res["unet"]["ms"] = min(res["unet"]["render_ms"] + res["unet"]["denoise_ms"], res["unet"]["frame_graph_ms"])
    # every method: its time, the plain tracer's time to its PSNR, and both speedups
    methods = {}
    for k in ("unet", "unet_fp8", "unet_fallback"):
        if k not in res:
            continue
        m = {"ms": res[k]["ms"], "psnr": res[k]["psnr"], "reaches_40db": res[k]["psnr"] >= 40.0}
        if m["reaches_40db"] and res["plain_ms_40db"]:
            m["speedup_40db"] = res["plain_ms_40db"] / m["ms"]
        spp_eq = plain_spp_for(res["plain"], m["psnr"])
        if spp_eq is not None:
            m["plain_spp_same_psnr"] = spp_eq
            m["plain_ms_same_psnr"] = spp_eq * ms_per_spp
            m["speedup_same_psnr"] = m["plain_ms_same_psnr"] / m["ms"]
        methods[k] = m
    res["methods"] = methods
    cands = [(m["ms"], k) for k, m in methods.items() if m["reaches_40db"]]
    res["unet_reaches_40db"] = bool(cands)
    if cands:
        res["unet_ms"], res["unet_variant_40db"] = min(cands)
        if res["plain_ms_40db"]:
            res["speedup_40db"] = res["plain_ms_40db"] / res["unet_ms"]
    # plain time to the PSNR that the fastest method (the one unet_ms times) reaches
    best = methods[res.get("unet_variant_40db", "unet")]
    if "plain_spp_same_psnr" in best:
        for k in ("plain_spp_same_psnr", "plain_ms_same_psnr", "speedup_same_psnr"):
            res[k] = best[k]
    return res


def off_orbit_camera(ctx):
    """A view the training set has no neighbour of: 6 m from the scene centre at 2.5 m height,
    looking across it (the training poses stand on a 15 m orbit, 3-12 m high)."""
    return camera_for(ctx.camera(), (-4.2, 4.3, 2.5), (3.0, -2.0, 3.0))


def held_out_camera(ctx):
    """The evaluation pose: an orbit pose drawn from a seed the training poses do not use."""
    (pos, tgt, _), = orbit_poses(1, seed=12345)
    return camera_for(ctx.camera(), pos, tgt)


def main(argv=None):
    ap = argparse.ArgumentParser(description="train the denoiser on tracer-generated pairs and evaluate it")
    ap.add_argument("--poses", type=int, default=50)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--time-budget", type=float, default=None, help="seconds of training at most")
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--crop", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-4)
    ap.add_argument("--width", type=int, default=1728)
    ap.add_argument("--height", type=int, default=864)
    ap.add_argument("--scene-seed", type=int, default=1)
    ap.add_argument("--ref-spp", type=int, default=4096)
    ap.add_argument("--out", default="denoiser.pt")
    ap.add_argument("--report", default=None, help="write the evaluation as JSON here")
    ap.add_argument("--cosine", action="store_true", help="cosine learning-rate decay over --steps")
    ap.add_argument("--export", default=None, help="also write the compact inference checkpoint here")
    ap.add_argument("--eval-only", default=None, metavar="CKPT", help="skip training; evaluate this checkpoint")
    args = ap.parse_args(argv)
    ctx = make_scene(1, args.scene_seed, args.width, args.height)
    extra = {}
    if args.eval_only:
        model = load_model(args.eval_only)
    else:
        (data, t_data) = timed(lambda: make_dataset(ctx, args.poses))
        print(f"dataset: {args.poses} poses x (8, 256) spp at {args.width}x{args.height} in {t_data / 1e3:.1f} s")
        model, ckpt, t_train = train(*data, steps=args.steps, batch=args.batch, crop=args.crop, lr=args.lr,
                                     time_budget=args.time_budget, cosine=args.cosine)
        del data
        torch.save(ckpt, args.out)
        if args.export:
            torch.save(export_checkpoint(ckpt), args.export)
        print(f"training: {t_train:.1f} s, best val L1 {ckpt['best_val']:.4f}, saved {args.out}")
        extra = {"train_s": t_train, "dataset_s": t_data / 1e3,
                 "steps": len(ckpt["history"]) and ckpt["history"][-1][0]}
    res = evaluate(ctx, model.eval(), held_out_camera(ctx), ref_spp=args.ref_spp)
    res.update(extra)
    print(json.dumps({k: v for k, v in res.items() if k != "plain"}))
    if args.report:
        pathlib.Path(args.report).write_text(json.dumps(res, indent=1))


if __name__ == "__main__":
    main()
