"""eval_checkpoints.py <noisy.png> <gt.png> <ckpt1.pt> [ckpt2.pt ...] [--save-dir DIR] [--hdr noisy.f32]

Drop-in for the reference's tools/eval_checkpoints.py: a markdown table of inference time
(best of three after a warm-up), PSNR and SSIM against the ground truth, for the raw input and
for each denoiser checkpoint. The input goes back to radiance as in denoise.py (the FP32
sidecar when given, else the inverse sqrt tonemap), the output is tonemapped like the renderer,
and PSNR/SSIM are the FP64 GPU reductions of liblgi. --save-dir writes <ckpt-stem>.denoised.png.
"""
import argparse
import pathlib
import sys
import time

import numpy as np

sys.path.insert(0, str(pathlib.Path(__file__).resolve().parent.parent))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("noisy_png")
    ap.add_argument("gt_png")
    ap.add_argument("checkpoints", nargs="+")
    ap.add_argument("--save-dir", help="if set, write <ckpt-stem>.denoised.png here")
    ap.add_argument("--hdr", default=None, help="raw FP32 HxWx3 radiance of the noisy image (cuda_ref --hdr)")
    args = ap.parse_args(argv)

    import torch
    from PIL import Image

    import lgi
    from lgi import denoise as D
    from lgi import metrics as M
    from lgi.train import load_model

    if not torch.cuda.is_available():
        sys.exit("eval_checkpoints.py needs a GPU: the denoiser and the metrics run only on the HIP kernels")
    save_dir = pathlib.Path(args.save_dir) if args.save_dir else None
    if save_dir:
        save_dir.mkdir(parents=True, exist_ok=True)
    noisy = np.asarray(Image.open(args.noisy_png).convert("RGB"), dtype=np.uint8)
    gt = torch.from_numpy(np.asarray(Image.open(args.gt_png).convert("RGB"), dtype=np.uint8)).cuda()
    if tuple(gt.shape) != noisy.shape:
        sys.exit(f"size mismatch: noisy {noisy.shape[1]}x{noisy.shape[0]}, gt {gt.shape[1]}x{gt.shape[0]}")
    try:
        rad = D.radiance_input(noisy, args.hdr)
    except ValueError as e:
        sys.exit(str(e))
    ctx = lgi.Context()

    def scores(img_u8):
        mse, ssim, _ = M.image_metrics(img_u8, gt)
        return M.psnr_from_mse(mse), ssim

    rows = []
    print("| checkpoint | infer ms | PSNR (dB) | SSIM |")
    print("|---|--:|--:|--:|")
    p, s = scores(torch.from_numpy(noisy).cuda())
    print(f"| (raw input) | — | {p:.2f} | {s:.4f} |")
    rows.append(("(raw input)", None, p, s))
    for path in args.checkpoints:
        model = load_model(path)
        D.denoise_any_size(model, rad)
        best = float("inf")
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            y = D.denoise_any_size(model, rad)
            torch.cuda.synchronize()
            best = min(best, (time.perf_counter() - t0) * 1000.0)
        out = D.tonemap(ctx, y)
        p, s = scores(out)
        name = pathlib.Path(path).name
        print(f"| {name} | {best:.1f} | {p:.2f} | {s:.4f} |")
        rows.append((name, best, p, s))
        if save_dir:
            lgi.write_png(save_dir / f"{pathlib.Path(path).stem}.denoised.png", out.cpu().numpy())
    return rows


if __name__ == "__main__":
    main()
