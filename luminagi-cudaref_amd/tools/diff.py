"""diff.py <engine.png> <ref.png> [--out report.html] [--gain G]

Engine screenshot against the path-traced reference, with the reference tools/diff.py's
arguments, report line and outputs: PSNR, SSIM and mean absolute difference, a difference
heatmap next to the report (<out>.heatmap.png) and a self-contained HTML page with the three
images. Everything past decoding the files runs on the GPU through liblgi (FP64 reductions,
heatmap kernel); lgi.metrics.diff_files does the work.
"""
import argparse
import pathlib
import sys

sys.path.insert(0, str(pathlib.Path(__file__).resolve().parent.parent))


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
    r = metrics.diff_files(args.engine, args.ref, args.out, gain=args.gain,
                           warn=lambda msg: print(msg, file=sys.stderr))
    print(metrics.report_line(r))
    print(f"wrote {args.out}")
    return r


if __name__ == "__main__":
    main()
