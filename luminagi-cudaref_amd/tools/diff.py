"""diff.py <engine.png> <ref.png> [--out report.html] [--gain G]

Drop-in for the reference's tools/diff.py (same arguments, outputs and report lines): PSNR,
SSIM and mean absolute difference of the engine screenshot against the path-traced reference,
a difference heatmap written next to the report (<out>.heatmap.png), and a self-contained HTML
page with the three images. The metrics (FP64) and the heatmap are computed on the GPU by
liblgi; an engine image of another size is first resized to the reference's with PIL's
bilinear filter, as the reference does.
"""
import argparse
import pathlib
import sys

import numpy as np

sys.path.insert(0, str(pathlib.Path(__file__).resolve().parent.parent))


def load_rgb(path):
    from PIL import Image
    img = Image.open(path).convert("RGB")
    return np.asarray(img, dtype=np.uint8), img.size


def match_size(a, target_wh):
    from PIL import Image
    if (a.shape[1], a.shape[0]) == tuple(target_wh):
        return a
    return np.asarray(Image.fromarray(a).resize(tuple(target_wh), Image.BILINEAR), dtype=np.uint8)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("engine", help="engine screenshot")
    ap.add_argument("ref", help="path-traced reference image")
    ap.add_argument("--out", default="diff_report.html")
    ap.add_argument("--gain", type=float, default=4.0, help="heatmap gain on the mean absolute difference")
    args = ap.parse_args(argv)

    import torch
    from lgi import metrics

    if not torch.cuda.is_available():
        sys.exit("diff.py needs a GPU: the metrics run only on the HIP kernels")
    engine, eng_size = load_rgb(args.engine)
    ref, ref_size = load_rgb(args.ref)
    if eng_size != ref_size:
        print(f"resizing engine {eng_size} -> {ref_size}", file=sys.stderr)
        engine = match_size(engine, ref_size)
    r = metrics.compare(engine, ref, gain=args.gain)

    import lgi
    out = pathlib.Path(args.out)
    if out.parent != pathlib.Path(""):
        out.parent.mkdir(parents=True, exist_ok=True)
    lgi.write_png(out.with_suffix(".heatmap.png"), r["heat"])
    out.write_text(metrics.html_report(engine, ref, r, args.engine, args.ref, gain=args.gain), encoding="utf-8")
    print(f"PSNR {r['psnr']:.2f} dB | SSIM {r['ssim']:.4f} | MAD {r['mad']:.2f}")
    print(f"wrote {args.out}")
    return r


if __name__ == "__main__":
    main()
