"""Host-side checks of the training loop's pieces (no GPU): the dynamic loss scaler's overflow and
skip path and its growth schedule, and the seeded train/validation split (ml/train.py:152-158)."""
import math

import numpy as np
import torch

from lgi.train import LossScaler, plain_spp_for


def _params(grads):
    ps = []
    for g in grads:
        p = torch.nn.Parameter(torch.ones(3))
        p.grad = torch.tensor(g, dtype=torch.float32)
        ps.append(p)
    return ps


def test_loss_scaler_skips_overflowing_steps_and_halves():
    ps = _params([[1.0, 2.0, 3.0], [math.inf, 0.0, 0.0]])
    opt = torch.optim.SGD(ps, lr=0.5)
    sc = LossScaler(scale=1024.0, interval=3)
    before = [p.detach().clone() for p in ps]
    assert sc.step(ps, opt) is False                 # inf gradient: no update
    assert sc.scale == 512.0 and sc.clean == 0
    assert all(torch.equal(p, b) for p, b in zip(ps, before))
    assert all(p.grad is None for p in ps)           # gradients cleared for the next step
    # NaN is an overflow too
    ps[0].grad = torch.tensor([math.nan, 0.0, 0.0])
    ps[1].grad = torch.zeros(3)
    assert sc.step(ps, opt) is False and sc.scale == 256.0


def test_loss_scaler_unscales_and_grows_after_clean_steps():
    ps = _params([[256.0, 512.0, 0.0]])
    opt = torch.optim.SGD(ps, lr=1.0)
    sc = LossScaler(scale=256.0, interval=2, max_scale=512.0)
    assert sc.step(ps, opt) is True
    # the update used the unscaled gradient (1, 2, 0)
    assert torch.allclose(ps[0].detach(), torch.tensor([0.0, -1.0, 1.0]))
    assert sc.scale == 256.0 and sc.clean == 1
    ps[0].grad = torch.zeros(3)
    assert sc.step(ps, opt) is True and sc.scale == 512.0      # doubled after `interval` clean steps
    for _ in range(4):
        ps[0].grad = torch.zeros(3)
        sc.step(ps, opt)
    assert sc.scale == 512.0                                    # capped at max_scale


def test_plain_spp_interpolation_follows_mse_over_spp():
    # MSE ~ 1/spp: +10*log10(2) dB per doubling; 40 dB between 64 and 128 spp here
    pts = [{"spp": s, "psnr": 30.0 + 10 * math.log10(s / 8)} for s in (8, 16, 32, 64, 128)]
    spp = plain_spp_for(pts, 40.0)
    assert abs(spp - 80.0) < 1e-9
    assert plain_spp_for(pts, 50.0) is None
    assert abs(plain_spp_for(pts, 29.0) - 8 * 10 ** -0.1) < 1e-12


def test_seeded_split_is_reproducible():
    a = np.random.RandomState(42).permutation(50)
    b = np.random.RandomState(42).permutation(50)
    assert np.array_equal(a, b) and sorted(a.tolist()) == list(range(50))
