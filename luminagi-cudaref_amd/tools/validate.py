"""validate.py <scene.json> [--open] [--out-dir DIR]

Checks an engine capture against the path tracer in one process. A capture is a scene file plus
the engine's screenshot saved beside it under the same stem (<scene>.png). The scene is loaded
and rendered through liblgi with the cuda_ref defaults (64 spp, 2 bounces), the render is written
to <out-dir>/ref_<stem>.png, and the screenshot is compared with it on the GPU, giving
<out-dir>/report_<stem>.html and its heatmap. The arguments and the output file names are those
of the reference's tools/validate.py; --out-dir defaults to output/ under the package root.
"""
import argparse
import contextlib
import pathlib
import sys
import time

PKG = pathlib.Path(__file__).resolve().parent.parent
sys.path.insert(0, str(PKG))

REF_SPP, REF_BOUNCES = 64, 2  # cuda_ref's defaults


@contextlib.contextmanager
def stage(name):
    t0 = time.perf_counter()
    yield
    print(f"[validate] {name}: {time.perf_counter() - t0:.2f} s")


def capture_paths(scene_json, out_dir):
    """(scene, screenshot, reference PNG, report) paths of one capture; raises FileNotFoundError
    naming every input that is not there."""
    scene = pathlib.Path(scene_json).resolve()
    shot = scene.with_suffix(".png")
    absent = [str(p) for p in (scene, shot) if not p.is_file()]
    if absent:
        raise FileNotFoundError("capture incomplete, not found: " + ", ".join(absent))
    out = pathlib.Path(out_dir).resolve()
    return scene, shot, out / f"ref_{scene.stem}.png", out / f"report_{scene.stem}.html"


def render_reference(scene, out_png, spp=REF_SPP, bounces=REF_BOUNCES):
    import lgi
    with lgi.Context() as ctx:
        ctx.load_json(scene)
        ctx.build_bvh()
        img = ctx.render(spp, bounces)["rgb"]
    out_png.parent.mkdir(parents=True, exist_ok=True)
    lgi.write_png(out_png, img)
    return img


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("scene_json")
    ap.add_argument("--open", action="store_true", help="open the HTML report after running")
    ap.add_argument("--out-dir", default=str(PKG / "output"))
    args = ap.parse_args(argv)
    try:
        scene, shot, ref_png, report = capture_paths(args.scene_json, args.out_dir)
    except FileNotFoundError as e:
        sys.exit(f"validate: {e}")

    from lgi import metrics

    with stage(f"reference render ({REF_SPP} spp, {REF_BOUNCES} bounces) -> {ref_png.name}"):
        render_reference(scene, ref_png)
    with stage(f"comparison -> {report.name}"):
        r = metrics.diff_files(shot, ref_png, report, warn=lambda msg: print(msg, file=sys.stderr))
    print(metrics.report_line(r))
    print(f"report: {report}")
    if args.open:
        import webbrowser
        webbrowser.open(report.as_uri())
    return report


if __name__ == "__main__":
    main()
