"""The drop-in entry points on the GPU: the cuda_ref CLI (single render, G-buffer, batch) and
tools/denoise.py, checked against the library calls they wrap."""
import importlib.util
import pathlib
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PKG = pathlib.Path(__file__).resolve().parent.parent / "luminagi-cudaref_amd"
CLI = PKG / "bin" / "cuda_ref"


def read_png(path):
    from PIL import Image
    return np.asarray(Image.open(path).convert("RGB"), dtype=np.uint8)


def small_scene(tmp_path, w=96, h=64):
    import lgi
    ctx = lgi.Context()
    ctx.generate(1, 7, w, h)
    path = tmp_path / "scene.json"
    ctx.save_json(path)
    return ctx, path


def test_cli_render_gbuffer_and_hdr_match_library(gpu, tmp_path):
    import lgi
    ctx, scene = small_scene(tmp_path)
    out = tmp_path / "img.png"
    hdr = tmp_path / "img.f32"
    r = subprocess.run([str(CLI), str(scene), "-o", str(out), "--spp", "4", "--bounces", "2", "--gbuffer",
                        "--hdr", str(hdr)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    assert "render:" in r.stdout and f"wrote {out}" in r.stdout
    ref = lgi.Context()
    ref.load_json(scene)
    ref.build_bvh()
    want = ref.render(4, 2, hdr=True, gbuffer=True)
    assert np.array_equal(read_png(out), want["rgb"])
    got_hdr = np.fromfile(hdr, dtype=np.float32).reshape(want["hdr"].shape)
    assert np.array_equal(got_hdr, want["hdr"])
    for name, key in (("albedo", "albedo"), ("normal", "normal"), ("worldpos", "worldpos")):
        assert np.array_equal(read_png(tmp_path / f"img.{name}.png"), want[key]), name


def test_cli_batch_mode_renders_every_job(gpu, tmp_path):
    import lgi
    ctx, scene = small_scene(tmp_path)
    jobs = tmp_path / "batch.txt"
    a, b = tmp_path / "out" / "a.png", tmp_path / "out" / "b.png"
    jobs.write_text(f"# two poses\n{a} 2 1 0 -20 8 0 0 4\n{b} 3 2 5 -18 10 0 0 5\n")
    r = subprocess.run([str(CLI), str(scene), "--batch", str(jobs)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and "batch: 2 jobs" in r.stdout
    ref = lgi.Context()
    ref.load_json(scene)
    ref.build_bvh()
    for path, spp, bounces, pos, tgt in ((a, 2, 1, (0, -20, 8), (0, 0, 4)), (b, 3, 2, (5, -18, 10), (0, 0, 5))):
        cam = lgi.camera_look_at(ref.camera(), pos, tgt)
        assert np.array_equal(read_png(path), ref.render(spp, bounces, camera=cam)["rgb"]), path


def test_tools_denoise_drop_in(gpu, tmp_path, capsys):
    """tools/denoise.py <ckpt> <noisy.png> <out.png> [--hdr]: PNG in, PNG out, reference report
    line; the HDR sidecar path equals the library's denoise of the same radiance."""
    import torch
    import lgi
    from lgi import denoise as D
    from lgi.unet import UNet
    spec = importlib.util.spec_from_file_location("lgi_denoise_tool", PKG / "tools" / "denoise.py")
    tool = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(tool)
    model = UNet(seed=4)
    ckpt = tmp_path / "d.pt"
    from lgi.train import ARCH
    torch.save({"model": model.state_dict(), "base": 32, "in_ch": 3, "arch": ARCH}, ckpt)
    ctx, scene = small_scene(tmp_path, w=72, h=40)  # not a multiple of 16: exercises the padding
    ctx.build_bvh()
    img = ctx.render(8, 2, hdr=True)
    noisy, hdr = tmp_path / "noisy.png", tmp_path / "noisy.f32"
    lgi.write_png(noisy, img["rgb"])
    img["hdr"].astype(np.float32).tofile(hdr)
    out = tmp_path / "den.png"
    tool.main([str(ckpt), str(noisy), str(out), "--hdr", str(hdr)])
    assert "inference:" in capsys.readouterr().out
    got = read_png(out)
    assert got.shape == img["rgb"].shape
    x = torch.from_numpy(img["hdr"]).cuda()
    xp = torch.nn.functional.pad(x.permute(2, 0, 1).unsqueeze(0), (0, 8, 0, 8), mode="replicate")[0].permute(1, 2, 0)
    model = model.cuda().eval()
    want = D.tonemap(ctx, D.denoise(model, xp.contiguous())[:40, :72].contiguous()).cpu().numpy()
    assert np.array_equal(got, want)
    out2 = tmp_path / "den2.png"
    tool.main([str(ckpt), str(noisy), str(out2)])  # PNG-only input (inverted tonemap)
    assert read_png(out2).shape == img["rgb"].shape


def test_cli_reads_binary_scene(gpu, tmp_path):
    import lgi
    ctx, scene = small_scene(tmp_path)
    ctx.save_bin(tmp_path / "scene.lgis")
    a, b = tmp_path / "a.png", tmp_path / "b.png"
    for src, out in ((scene, a), (tmp_path / "scene.lgis", b)):
        r = subprocess.run([str(CLI), str(src), "-o", str(out), "--spp", "2"], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
    assert np.array_equal(read_png(a), read_png(b))


def load_tool(name):
    spec = importlib.util.spec_from_file_location(f"lgi_tool_{name}", PKG / "tools" / f"{name}.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def fmt_line(a, b):
    from oracle import metrics_ref as M
    return f"PSNR {M.psnr(a, b):.2f} dB | SSIM {M.ssim(a, b):.4f} | MAD {M.mad(a, b):.2f}"


def test_tools_diff_report(gpu, tmp_path, capsys):
    """diff.py <engine> <ref> --out: the report line, the heatmap sidecar and the HTML page, each
    against the oracle's restatement of the reference's metrics and heatmap."""
    import lgi
    from oracle import metrics_ref as M
    ctx, _ = small_scene(tmp_path, w=120, h=72)
    ctx.build_bvh()
    eng, ref = ctx.render(2, 2)["rgb"], ctx.render(64, 2)["rgb"]
    lgi.write_png(tmp_path / "eng.png", eng)
    lgi.write_png(tmp_path / "ref.png", ref)
    out = tmp_path / "rep" / "report.html"
    load_tool("diff").main([str(tmp_path / "eng.png"), str(tmp_path / "ref.png"), "--out", str(out)])
    lines = capsys.readouterr().out.splitlines()
    assert lines[0] == fmt_line(eng, ref) and lines[1] == f"wrote {out}"
    assert np.array_equal(read_png(tmp_path / "rep" / "report.heatmap.png"), M.heatmap(eng, ref))
    page = out.read_text()
    assert f"{M.psnr(eng, ref):.2f} dB" in page and f"{M.ssim(eng, ref):.4f}" in page
    assert page.count("data:image/png;base64,") == 3 and "120x72" in page


def test_tools_diff_resizes_engine(gpu, tmp_path, capsys):
    from PIL import Image
    from oracle import metrics_ref as M
    rng = np.random.default_rng(9)
    ref = rng.integers(0, 256, (48, 64, 3), dtype=np.uint8)
    eng = rng.integers(0, 256, (30, 40, 3), dtype=np.uint8)
    Image.fromarray(ref).save(tmp_path / "ref.png")
    Image.fromarray(eng).save(tmp_path / "eng.png")
    load_tool("diff").main([str(tmp_path / "eng.png"), str(tmp_path / "ref.png"), "--out", str(tmp_path / "r.html")])
    cap = capsys.readouterr()
    assert "resizing engine (40, 30) -> (64, 48)" in cap.err
    resized = np.asarray(Image.fromarray(eng).resize((64, 48), Image.BILINEAR), dtype=np.uint8)
    assert cap.out.splitlines()[0] == fmt_line(resized, ref)


def test_tools_validate_flow(gpu, tmp_path, capsys):
    """validate.py <scene.json>: the reference render (cuda_ref defaults) is written next to the
    report and compared with the engine screenshot (<scene>.png); the report's numbers are the
    oracle's on those files."""
    import lgi
    ctx, scene = small_scene(tmp_path)
    ctx.build_bvh()
    lgi.write_png(scene.with_suffix(".png"), ctx.render(1, 1)["rgb"])
    report = load_tool("validate").main([str(scene), "--out-dir", str(tmp_path / "out")])
    out = capsys.readouterr().out
    ref_png = tmp_path / "out" / "ref_scene.png"
    assert report == tmp_path / "out" / "report_scene.html" and report.is_file()
    assert "[validate] reference render (64 spp, 2 bounces)" in out and "[validate] comparison" in out
    assert (tmp_path / "out" / "report_scene.heatmap.png").is_file()
    assert fmt_line(read_png(scene.with_suffix(".png")), read_png(ref_png)) in out
    ref = lgi.Context()
    ref.load_json(scene)
    ref.build_bvh()
    assert np.array_equal(read_png(ref_png), ref.render(64, 2)["rgb"])  # cuda_ref defaults: 64 spp, 2 bounces


def test_tools_eval_checkpoints_table(gpu, tmp_path, capsys):
    import torch
    import lgi
    from oracle import metrics_ref as M
    from lgi.train import ARCH
    from lgi.unet import UNet
    ckpt = tmp_path / "a.pt"
    torch.save({"model": UNet(seed=2).state_dict(), "base": 32, "in_ch": 3, "arch": ARCH}, ckpt)
    ctx, _ = small_scene(tmp_path, w=80, h=48)
    ctx.build_bvh()
    noisy = ctx.render(4, 2, hdr=True)
    gt = ctx.render(256, 2)["rgb"]
    lgi.write_png(tmp_path / "noisy.png", noisy["rgb"])
    noisy["hdr"].astype(np.float32).tofile(tmp_path / "noisy.f32")
    lgi.write_png(tmp_path / "gt.png", gt)
    rows = load_tool("eval_checkpoints").main([str(tmp_path / "noisy.png"), str(tmp_path / "gt.png"), str(ckpt),
                                              "--hdr", str(tmp_path / "noisy.f32"), "--save-dir", str(tmp_path / "s")])
    out = capsys.readouterr().out.splitlines()
    assert out[0] == "| checkpoint | infer ms | PSNR (dB) | SSIM |"
    raw = noisy["rgb"]
    assert out[2] == f"| (raw input) | — | {M.psnr(raw, gt):.2f} | {M.ssim(raw, gt):.4f} |"
    den = read_png(tmp_path / "s" / "a.denoised.png")
    assert rows[1][0] == "a.pt" and rows[1][1] > 0
    assert rows[1][2] == pytest.approx(M.psnr(den, gt), abs=1e-9)
    assert rows[1][3] == pytest.approx(M.ssim(den, gt), abs=1e-12)


def test_cli_renderers_write_the_same_png(gpu, tmp_path):
    """cuda_ref with and without --sort, and with --sort --no-ray-sort: one image."""
    _, scene = small_scene(tmp_path)
    outs = []
    for k, extra in enumerate(([], ["--sort"], ["--sort", "--no-ray-sort"])):
        out = tmp_path / f"r{k}.png"
        r = subprocess.run([str(CLI), str(scene), "-o", str(out), "--spp", "4"] + extra, capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        outs.append(read_png(out))
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
