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
