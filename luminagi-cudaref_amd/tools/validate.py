"""validate.py <scene.json> [--open] [--out-dir DIR]

Drop-in for the reference's tools/validate.py: renders the scene with bin/cuda_ref into
<out-dir>/ref_<stem>.png, then runs diff.py on the engine screenshot saved next to the scene
(<scene>.png) against it, writing <out-dir>/report_<stem>.html. The output directory defaults
to output/ under the package root. --open shows the report in the default browser.
"""
import argparse
import pathlib
import subprocess
import sys
import time

PKG = pathlib.Path(__file__).resolve().parent.parent
CUDA_REF = PKG / "bin" / "cuda_ref"
DIFF_PY = PKG / "tools" / "diff.py"


def run(cmd, label, cwd):
    print(f"--- {label} ---")
    print("  " + " ".join(f'"{c}"' if " " in str(c) else str(c) for c in cmd))
    t0 = time.perf_counter()
    r = subprocess.run([str(c) for c in cmd], cwd=cwd, capture_output=True, text=True)
    dt = time.perf_counter() - t0
    if r.stdout:
        print(r.stdout.rstrip())
    if r.stderr:
        print(r.stderr.rstrip(), file=sys.stderr)
    if r.returncode != 0:
        sys.exit(f"{label} failed (exit {r.returncode})")
    print(f"  ({dt:.1f}s)")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("scene_json")
    ap.add_argument("--open", action="store_true", help="open the HTML report after running")
    ap.add_argument("--out-dir", default=str(PKG / "output"))
    args = ap.parse_args(argv)

    scene = pathlib.Path(args.scene_json).resolve()
    if not scene.is_file():
        sys.exit(f"no such file: {scene}")
    engine_png = scene.with_suffix(".png")
    if not engine_png.is_file():
        sys.exit(f"engine screenshot missing: {engine_png}")
    if not CUDA_REF.is_file():
        sys.exit(f"{CUDA_REF} is not built (make -C {PKG})")

    out_dir = pathlib.Path(args.out_dir).resolve()
    out_dir.mkdir(parents=True, exist_ok=True)
    stem = scene.stem
    ref_png = out_dir / f"ref_{stem}.png"
    report = out_dir / f"report_{stem}.html"

    run([CUDA_REF, scene, ref_png], "path-traced reference render", PKG)
    run([sys.executable, DIFF_PY, engine_png, ref_png, "--out", report], "image diff", PKG)

    print()
    print(f"report: {report}")
    if args.open:
        import webbrowser
        webbrowser.open(report.as_uri())
    return report


if __name__ == "__main__":
    main()
