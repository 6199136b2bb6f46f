"""GPU tracer parity against the CPU oracle (same PCG32 / xorshift32 streams).

Bars: primary-hit triangle ids bit-exact except exact distance ties;
direct-only 1-spp and 2-bounce images within 1e-5 relative per pixel; G-buffer and tonemap
bytes exact. The kernels evaluate the reference expressions in the same order with no FP
contraction, so the sums agree bit for bit on these scenes: no value may differ by more than
1e-5, and tonemapped and G-buffer bytes must be equal.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as orc
from parity import REL_TOL, assert_close_images, gpu_accumulate, rel_err

pytestmark = pytest.mark.gpu


def make_pair(kind=1, seed=1, w=512, h=256):
    import lgi
    ctx = lgi.Context()
    ctx.generate(kind, seed, w, h)
    s = ctx.scene()
    oscene = orc.Scene(s["tris"], s["materials"][:, 0, :], s["points"], s["sun"], s["camera"])
    ctx.build_bvh()
    return ctx, oscene, s


def small_camera(cam, w, h):
    import lgi
    c = lgi.Camera.from_buffer_copy(cam)
    c.width, c.height = w, h
    return c


def test_primary_hits_bit_exact_config1(gpu):
    import torch
    ctx, osc, s = make_pair()
    cam = s["camera"]
    H, W = cam.height, cam.width
    tri = torch.empty(H * W, dtype=torch.int32, device="cuda")
    t = torch.empty(H * W, dtype=torch.float32, device="cuda")
    ctx.primary_hits(tri, t, camera=cam)
    torch.cuda.synchronize()
    gtri, gt = tri.cpu().numpy().reshape(H, W), t.cpu().numpy().reshape(H, W)
    otri, ot = osc.primary_hits()
    assert (otri >= 0).mean() > 0.3, "camera must see the scene"
    mism = gtri != otri
    print(f"primary: {mism.sum()} id mismatches of {mism.size}; t equal {np.mean(gt == ot):.6f}", flush=True)
    assert np.array_equal(gt, ot)            # hit distances bit-identical everywhere
    if mism.any():                           # a different id is only allowed at an exact tie
        s_tris = s["tris"]
        assert np.all(gt[mism] == ot[mism])
        assert mism.mean() < 1e-3


def test_direct_1spp_config1(gpu):
    ctx, osc, s = make_pair()
    cam = s["camera"]
    gs = gpu_accumulate(ctx, cam, 1, 0)
    o = osc.render(1, 0)
    assert_close_images(gs[0].cpu().numpy(), o["sum"], "direct 1spp")
    assert rel_err(gs[2].cpu().numpy().reshape(o["sumsq"].shape), o["sumsq"]).max() <= REL_TOL


@pytest.mark.parametrize("flags", [0, 1, 4], ids=["pcg32", "xorshift_ref", "no_rr"])
def test_two_bounce_64spp_small(gpu, flags):
    ctx, osc, s = make_pair()
    cam = small_camera(s["camera"], 128, 64)
    gs = gpu_accumulate(ctx, cam, 64, 2, flags=flags)
    o = osc.render(64, 2, flags=flags, camera=cam)
    assert_close_images(gs[0].cpu().numpy(), o["sum"], f"2-bounce 64spp flags={flags}")
    # Kahan FP32 sum agrees with the FP64 sum of the same samples
    g = gs[0].cpu().numpy().reshape(o["sum64"].shape)
    assert np.all(np.abs(g - o["sum64"]) <= 1e-6 * np.abs(o["sum64"]) + 1e-6)


@pytest.mark.parametrize("flags", [2, 2 | 4, 2 | 8], ids=["sorted", "sorted_no_rr", "no_ray_sort"])
def test_sorted_wavefront_matches_inline_and_oracle(gpu, flags):
    """The wavefront renderer, with or without its ray sort, produces the inline kernel's sums
    bit for bit."""
    ctx, osc, s = make_pair(seed=6)
    cam = small_camera(s["camera"], 200, 100)
    inline = gpu_accumulate(ctx, cam, 16, 2, flags=flags & 4)
    wave = gpu_accumulate(ctx, cam, 16, 2, flags=flags)
    for a, b in zip(inline, wave):
        a, b = a.cpu().numpy(), b.cpu().numpy()
        print(f"sorted vs inline: {np.mean(a == b):.6f} equal", flush=True)
        assert np.array_equal(a, b)
    o = osc.render(16, 2, flags=flags & 4, camera=cam)
    assert_close_images(wave[0].cpu().numpy(), o["sum"], f"sorted flags={flags}")


@pytest.mark.parametrize("bounces", [0, 1, 3])
@pytest.mark.parametrize("flags", [2, 2 | 8], ids=["sorted", "no_ray_sort"])
def test_wavefront_bounce_counts_match_inline(gpu, flags, bounces):
    """Other bounce counts: the last bounce's shadow rays are traced in the launch that finds its
    closest hits once there is a bounce (with 3, after a launch that does not), or in a launch of
    their own without one; the sums equal the inline kernel's bit for bit either way."""
    ctx, _, s = make_pair(seed=6)
    cam = small_camera(s["camera"], 200, 100)
    inline = gpu_accumulate(ctx, cam, 8, bounces)
    wave = gpu_accumulate(ctx, cam, 8, bounces, flags=flags)
    for a, b in zip(inline, wave):
        assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())


def test_converged_means_agree_across_rng_policies(gpu):
    """4096-spp means of the PCG32 sorted renderer and of the reference's xorshift32 seeding
    (src/pathtracer.cu:345-364) agree within 3 standard errors of their difference: the RNG
    choice changes the noise, not the estimator. Independent sample sets, per-pixel
    luminance, 64x32 pixels."""
    import lgi
    ctx, _, s = make_pair(seed=3)
    cam = small_camera(s["camera"], 64, 32)
    lw = np.array([0.2126, 0.7152, 0.0722])
    stats = []
    for flags in (lgi.RENDER_SORTED, lgi.RNG_XORSHIFT_REF):
        g = gpu_accumulate(ctx, cam, 4096, 2, flags=flags)
        n = 4096.0
        m = (g[0].cpu().numpy().reshape(-1, 3) / n) @ lw
        var = np.maximum(g[2].cpu().numpy() / n - m * m, 0.0) * n / (n - 1)
        stats.append((m, var / n))
    (m0, v0), (m1, v1) = stats
    lit = (v0 + v1) > 0
    z = np.abs(m0 - m1)[lit] / np.sqrt(v0 + v1)[lit]
    print(f"converged 4096 spp: {lit.sum()} lit pixels, max |z| {z.max():.2f}, "
          f"frac |z|>3 {np.mean(z > 3):.4f}, mean z^2 {np.mean(z * z):.3f}", flush=True)
    assert lit.sum() > 0.3 * lit.size
    assert np.mean(z > 3) < 0.01             # ~0.27% expected beyond 3 sigma
    assert 0.7 < np.mean(z * z) < 1.3        # chi-square per pixel ~ 1 when unbiased
    # image-wide means agree within 3 standard errors of the mean over pixels
    d = (m0 - m1)[lit]
    assert abs(d.mean()) < 3 * np.sqrt((v0 + v1)[lit].sum()) / lit.sum()


def test_sorted_rejects_xorshift(gpu):
    import lgi
    ctx, _, s = make_pair(seed=6, w=32, h=16)
    with pytest.raises(lgi.LgiError) as e:
        gpu_accumulate(ctx, s["camera"], 1, 0, flags=lgi.RENDER_SORTED | lgi.RNG_XORSHIFT_REF)
    assert e.value.status == 8


def test_chunked_accumulation_equals_single_call(gpu):
    ctx, osc, s = make_pair(seed=2)
    cam = small_camera(s["camera"], 96, 48)
    one = gpu_accumulate(ctx, cam, 24, 2)
    two = gpu_accumulate(ctx, cam, 10, 2)
    two = gpu_accumulate(ctx, cam, 14, 2, sample0=10, sums=two)
    for a, b in zip(one, two):
        assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())


def test_render_deterministic_and_tonemap(gpu):
    ctx, osc, s = make_pair(seed=3)
    cam = small_camera(s["camera"], 160, 80)
    a = ctx.render(8, 2, camera=cam, hdr=True)
    b = ctx.render(8, 2, camera=cam, hdr=True)
    assert np.array_equal(a["rgb"], b["rgb"]) and np.array_equal(a["hdr"], b["hdr"])
    o = osc.render(8, 2, camera=cam)
    ref_rgb = orc.tonemap(o["sum"], np.float32(1.0) / np.float32(8))
    diff = a["rgb"] != ref_rgb
    assert not diff.any(), f"{diff.sum()} tonemapped bytes differ"


def test_tonemap_kernel_exact(gpu):
    import torch
    ctx, _, _ = make_pair(seed=4, w=32, h=16)
    rng = np.random.default_rng(5)
    acc = rng.uniform(-1, 40, (257, 3)).astype(np.float32)
    acc[:5] = [[0, 0, 0], [16, 16, 16], [15.99, 0.001, 1e-9], [np.inf, -np.inf, 8], [1e30, 3, 2]]
    d = torch.from_numpy(acc.reshape(-1)).cuda()
    out = torch.empty(acc.size, dtype=torch.uint8, device="cuda")
    inv = np.float32(1.0) / np.float32(16)
    ctx.tonemap(d, 257, float(inv), out)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy().reshape(-1, 3), orc.tonemap(acc, inv))


def test_gbuffer_matches_oracle(gpu):
    import torch
    ctx, osc, s = make_pair(seed=1)
    cam = s["camera"]
    H, W = cam.height, cam.width
    bufs = [torch.empty(H * W * 3, dtype=torch.uint8, device="cuda") for _ in range(3)]
    ctx.gbuffer(*bufs, camera=cam)
    torch.cuda.synchronize()
    ref = osc.gbuffer()
    for name, g, r in zip(["albedo", "normal", "worldpos"], bufs, ref):
        g = g.cpu().numpy().reshape(H, W, 3)
        print(f"gbuffer {name}: differing bytes {np.count_nonzero(g != r)}", flush=True)
        assert np.array_equal(g, r), name


@pytest.mark.parametrize("flags", [2, 0, 10], ids=["sorted", "inline", "wavefront-no-sort"])
def test_frame_graph_replays_match_eager_renderers(gpu, flags):
    """A frame (zero accumulators, trace (+ sort), accumulate) of either renderer captured in a
    HIP graph: every replay reproduces the eager frame bit for bit, with eager work and
    synchronisations between the replays."""
    import torch
    ctx, _, s = make_pair(seed=4)
    cam = small_camera(s["camera"], 160, 96)
    eager = gpu_accumulate(ctx, cam, 8, 2, flags=flags, sample0=5)
    bufs = [torch.zeros_like(t) for t in eager]
    ctx.accumulate(bufs[0], bufs[1], 8, 2, flags=flags, sample0=5, d_sumsq=bufs[2], camera=cam)  # warm-up
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        ctx.set_stream(torch.cuda.current_stream().cuda_stream)  # the renderer joins the capture stream
        for t in bufs:
            t.zero_()
        ctx.accumulate(bufs[0], bufs[1], 8, 2, flags=flags, sample0=5, d_sumsq=bufs[2], camera=cam)
    ctx.set_stream(None)
    for k in range(3):
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip(eager, bufs):
            assert torch.equal(a, b), f"replay {k}"
        _ = gpu_accumulate(ctx, cam, 2, 1)  # eager work between replays


WAVEFRONT_PATHS = 96 << 20  # paths the wavefront renderer keeps in flight per pass


@pytest.mark.parametrize("flags", [2, 10], ids=["sorted", "no_ray_sort"])
def test_wavefront_multi_pass_equals_inline(gpu, flags):
    """A render with more samples than the wavefront renderer keeps in flight runs as several
    passes (a full one, then the rest); the sums equal the inline kernel's bit for bit."""
    ctx, _, s = make_pair(seed=4)
    cam = small_camera(s["camera"], 64, 64)  # 64 tiles of 64 paths per sample
    spp = WAVEFRONT_PATHS // (64 * 64) + 40
    inline = gpu_accumulate(ctx, cam, spp, 2)
    wave = gpu_accumulate(ctx, cam, spp, 2, flags=flags)
    for a, b in zip(inline, wave):
        assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())


@pytest.mark.parametrize("flags,spp,noalloc", [(0, 8, False), (2, 8, False), (10, 8, False), (0, 8, True),
                                               (10, WAVEFRONT_PATHS // (64 * 64) + 40, False)],
                         ids=["inline", "sorted", "no_ray_sort", "noalloc", "wavefront_multi_pass"])
def test_set_sums_equals_zeroed_buffers(gpu, flags, spp, noalloc):
    """LGI_RENDER_SET_SUMS writes the sums over whatever the buffers held: bit for bit what an
    accumulation into zeroed buffers gives (the multi-pass case adds its later passes)."""
    import torch
    import lgi
    ctx, _, s = make_pair(seed=4)
    cam = small_camera(s["camera"], 64, 64)
    ref = gpu_accumulate(ctx, cam, spp, 2, flags=flags, sample0=3)
    bufs = [torch.full_like(t, float("nan")) for t in ref]
    bufs[1].fill_(123.0)
    call = ctx.accumulate_noalloc if noalloc else ctx.accumulate
    call(bufs[0], bufs[1], spp, 2, flags=flags | lgi.RENDER_SET_SUMS, sample0=3, d_sumsq=bufs[2], camera=cam)
    torch.cuda.synchronize()
    for a, b in zip(ref, bufs):
        assert torch.equal(a, b)


def test_set_sums_rejected_by_batch_and_tile_renders(gpu):
    import torch
    import lgi
    ctx, _, s = make_pair(seed=4)
    cam = small_camera(s["camera"], 32, 32)
    n = cam.width * cam.height
    bufs = [torch.zeros(n * 3, device="cuda"), torch.zeros(n * 3, device="cuda"), torch.zeros(n, device="cuda")]
    with pytest.raises(lgi.LgiError):
        ctx.render_tiles(np.array([0, 1], np.int32), 4, bufs[0], bufs[1], 2, d_sumsq=bufs[2], camera=cam,
                         flags=lgi.RENDER_SET_SUMS)
    with pytest.raises(lgi.LgiError):
        ctx.render_batch([(cam, 4, 0)], bufs[0], bufs[1], 2, d_sumsq=bufs[2], flags=lgi.RENDER_SET_SUMS)


@pytest.mark.parametrize("w,h", [(1, 1), (7, 3), (9, 17), (65, 1)])
def test_tiny_and_ragged_images(gpu, w, h):
    """Image sizes that leave partial 8x8 tiles (and a single pixel): the inline renderer, the
    wavefront renderer with and without its ray sort, and the tile renderer over every tile
    give the oracle's sums."""
    import torch
    import lgi
    ctx, osc, s = make_pair(seed=5)
    cam = small_camera(s["camera"], w, h)
    o = osc.render(8, 2, camera=cam)
    sums = [gpu_accumulate(ctx, cam, 8, 2, flags=f)[0].cpu().numpy()
            for f in (0, lgi.RENDER_SORTED, lgi.RENDER_SORTED | lgi.RENDER_NO_RAY_SORT)]
    tiles = list(range(((w + 7) // 8) * ((h + 7) // 8)))
    t_sum, t_comp = torch.zeros(w * h * 3, device="cuda"), torch.zeros(w * h * 3, device="cuda")
    ctx.render_tiles(tiles, 8, t_sum, t_comp, 2, flags=lgi.RENDER_SORTED, camera=cam)
    torch.cuda.synchronize()
    sums.append(t_sum.cpu().numpy())
    for k, g in enumerate(sums):
        assert np.array_equal(g, sums[0]), f"renderer {k} differs from the inline one"
    assert_close_images(sums[0], o["sum"], f"{w}x{h}")
