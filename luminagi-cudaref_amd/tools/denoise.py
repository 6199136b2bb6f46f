"""denoise.py <checkpoint.pt> <noisy.png> <out.png> [--hdr noisy.f32]

Drop-in for the reference's tools/denoise.py (same positional arguments, same one-line report).
The network runs on the project's BF16 MFMA kernels and works on log1p radiance, so the input
is taken back to radiance first: from the raw FP32 HxWx3 sidecar that `cuda_ref --hdr` writes
when one is given, otherwise by inverting the reference's sqrt tonemap of the PNG (values that
the PNG clipped at 1.0 stay clipped). The output PNG uses the same sqrt tonemap as the renderer
(src/pathtracer.cu:413-429). The image is padded to a multiple of 16 by edge replication (four
2x pools) and cropped back, as the reference pads to a multiple of 8 for its three pools.
"""
import argparse
import pathlib
import sys
import time

import numpy as np

sys.path.insert(0, str(pathlib.Path(__file__).resolve().parent.parent))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("checkpoint")
    ap.add_argument("noisy_png")
    ap.add_argument("out_png")
    ap.add_argument("--hdr", default=None, help="raw FP32 HxWx3 radiance written by cuda_ref --hdr")
    args = ap.parse_args(argv)

    import torch
    from PIL import Image

    import lgi
    from lgi import denoise as D
    from lgi.train import load_model

    if not torch.cuda.is_available():
        sys.exit("denoise.py needs a GPU: the denoiser runs only on the HIP kernels")
    png = np.asarray(Image.open(args.noisy_png).convert("RGB"), dtype=np.uint8)
    H, W = png.shape[:2]
    try:
        rad = D.radiance_input(png, args.hdr)
    except ValueError as e:
        sys.exit(str(e))
    model = load_model(args.checkpoint)
    D.denoise_any_size(model, rad)  # first call allocates; time the second
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    y = D.denoise_any_size(model, rad)
    torch.cuda.synchronize()
    dt_ms = (time.perf_counter() - t0) * 1000.0
    out = D.tonemap(lgi.Context(), y).cpu().numpy()
    Image.fromarray(out).save(args.out_png)
    print(f"inference: {dt_ms:.1f} ms ({W}x{H}, base=32)  wrote {args.out_png}")


if __name__ == "__main__":
    main()
