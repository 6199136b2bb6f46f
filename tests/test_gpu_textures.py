"""Textured materials on the GPU against the oracle: the inline and the sorted renderer, the
G-buffer albedo, the asset-root rule and the reference's texture log line."""
import json

import numpy as np
import pytest

from oracle import oracle as orc
from parity import assert_close_images

pytestmark = pytest.mark.gpu


def textured_scene(tmp_path, w=128, h=64):
    """The 16-mesh scene with textures on materials 0-2 and a missing file on material 3, laid
    out as the reference expects: <root>/scenes/scene.json, textures under <root>/tex/."""
    import lgi
    gen = lgi.Context()
    gen.generate(1, 5, w, h)
    (tmp_path / "scenes").mkdir()
    (tmp_path / "tex").mkdir()
    path = tmp_path / "scenes" / "scene.json"
    gen.save_json(path)
    doc = json.loads(path.read_text())
    rng = np.random.default_rng(8)
    textures = {}
    for m, (tw, th) in enumerate([(16, 8), (5, 7), (64, 64)]):
        img = rng.integers(0, 256, (th, tw, 3), dtype=np.uint8)
        lgi.write_png(tmp_path / "tex" / f"m{m}.png", img)
        doc["materials"][m]["diffuse_tex"] = f"tex/m{m}.png"
        textures[m] = np.concatenate([img, np.full((th, tw, 1), 255, np.uint8)], -1)
    doc["materials"][3]["diffuse_tex"] = "tex/missing.png"
    path.write_text(json.dumps(doc))
    ctx = lgi.Context()
    ctx.load_json(path)
    s = ctx.scene()
    osc = orc.Scene(s["tris"], s["materials"][:, 0, :], s["points"], s["sun"], s["camera"], textures=textures)
    ctx.build_bvh()
    return ctx, osc, s


def compare(gpu_sum, cpu_sum, label):
    assert_close_images(gpu_sum, cpu_sum, label)


@pytest.mark.parametrize("spp,bounces,flags", [(1, 0, 0), (16, 2, 0), (16, 2, 2)],
                         ids=["direct", "2-bounce", "2-bounce-sorted"])
def test_textured_render_matches_oracle(gpu, tmp_path, spp, bounces, flags):
    import torch
    ctx, osc, s = textured_scene(tmp_path)
    cam = s["camera"]
    n = cam.width * cam.height
    acc = [torch.zeros(n * 3, device="cuda"), torch.zeros(n * 3, device="cuda")]
    ctx.accumulate(acc[0], acc[1], spp, bounces, flags=flags & 2, camera=cam)
    torch.cuda.synchronize()
    o = osc.render(spp, bounces)
    compare(acc[0].cpu().numpy(), o["sum"], f"textured spp={spp} bounces={bounces} flags={flags}")
    # the textures change the image: the same scene with constant albedos differs
    plain = orc.Scene(s["tris"], s["materials"][:, 0, :], s["points"], s["sun"], cam)
    assert not np.array_equal(plain.render(spp, bounces)["sum"], o["sum"])


def test_textured_gbuffer_and_log_line(gpu, tmp_path, capfd):
    ctx, osc, s = textured_scene(tmp_path)
    out = ctx.render(1, 0, gbuffer=True)
    log = capfd.readouterr()
    assert "textures: 3 loaded, 1 missing (of" in log.out
    assert "texture not found:" in log.err and "missing.png" in log.err
    a, nrm, wp = osc.gbuffer()
    assert np.array_equal(out["albedo"], a)
    assert np.array_equal(out["normal"], nrm)


def test_asset_root_override(gpu, tmp_path, capfd):
    ctx, _, _ = textured_scene(tmp_path)
    ctx.set_asset_root(tmp_path / "elsewhere")  # nothing there: every texture goes missing
    ctx.build_bvh()
    ctx.render(1, 0)
    assert "textures: 0 loaded, 4 missing" in capfd.readouterr().out
    ctx.set_asset_root(None)
    ctx.build_bvh()
    ctx.render(1, 0)
    assert "textures: 3 loaded, 1 missing" in capfd.readouterr().out
