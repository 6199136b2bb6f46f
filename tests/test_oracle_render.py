"""Pins the oracle's renderer restatement (src/pathtracer.cu:61-429) with checks that do not
retype its formulas: published known-answer vectors (PCG32, xorshift32), libm (sin/cos),
closed forms (sun on an open plane, point-light falloff), moments of the cosine
distribution, FP64 sums, and the reference's chunked xorshift seeding."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as orc


def test_pcg32_known_answer():
    # pcg32-demo (pcg-random.org, pcg_basic.c): pcg32_srandom_r(&rng, 42u, 54u), first outputs
    expect = [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]
    assert orc.pcg32(42, 54, 6).tolist() == expect


def test_xorshift32_known_answer():
    # Marsaglia (2003) xorshift32 (13, 17, 5) from state 1
    assert orc.xorshift(1, 3).tolist() == [270369, 67634689, 2647435461]


def test_pixel_streams_uniform_and_distinct():
    a = orc.pixel_randf(0, 1000, 0, 4096)
    b = orc.pixel_randf(0, 1001, 0, 8)
    c = orc.pixel_randf(0, 1000, 1, 8)
    assert 0.0 <= a.min() and a.max() < 1.0
    assert abs(a.mean() - 0.5) < 0.02 and abs(a.var() - 1 / 12) < 0.01
    assert not np.array_equal(a[:8], b) and not np.array_equal(a[:8], c)
    # randf = (x >> 8) * 2^-24: every value is a multiple of 2^-24
    assert np.all((a * 2.0 ** 24) == np.floor(a * 2.0 ** 24))


def test_sincos_against_libm():
    x = np.linspace(0, 2 * np.pi, 200001, dtype=np.float32)[:-1]
    x = np.concatenate([x, (np.arange(1 << 16, dtype=np.float32) * np.float32(2 * np.pi / 2 ** 16))])
    s, c = orc.sincos(x)
    xs = x.astype(np.float64)
    ulp = np.float64(2.0 ** -24)
    assert np.max(np.abs(s - np.sin(xs))) <= 2 * ulp
    assert np.max(np.abs(c - np.cos(xs))) <= 2 * ulp


def test_cos_hemisphere_moments():
    for n in ([0, 0, 1], [0.6, 0, -0.8], [0, 0, -1], [0.267261, 0.534522, 0.801784]):
        d = orc.sample_cos_hemi(n, 12345, 7, 200000).astype(np.float64)
        N = np.asarray(n, np.float64)
        N /= np.linalg.norm(N)
        cos = d @ N
        assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
        assert cos.min() >= -1e-6
        # p(w) = cos/pi: E[cos] = 2/3, E[cos^2] = 1/2 (SE ~ 0.0005)
        assert abs(cos.mean() - 2 / 3) < 0.003 and abs((cos ** 2).mean() - 0.5) < 0.003
        # azimuth is isotropic: the tangential mean vanishes
        t = d - np.outer(cos, N)
        assert np.linalg.norm(t.mean(0)) < 0.01


def plane_scene(sun_dir=(0.3, -0.2, -1.0), sun_I=3.0, points=(), albedo=(0.25, 0.5, 0.75), W=64, H=48):
    """Ground quad (z = 0, 200 x 200 m) seen from above; nothing can shadow it."""
    import lgi
    t = np.zeros(2, orc.TRI_DTYPE)
    P = np.array([[-100, -100, 0], [100, -100, 0], [100, 100, 0], [-100, 100, 0]], np.float32)
    t["v"][0] = P[[0, 1, 2]]
    t["v"][1] = P[[0, 2, 3]]
    t["n"][:, :, 2] = 1.0
    cam = lgi.Camera()
    cam.width, cam.height, cam.fov_y_deg, cam.aspect = W, H, 60.0, W / H
    lgi.camera_look_at(cam, (0.0, -3.0, 6.0), (0.0, 0.0, 0.0))
    sun = lgi.DirLight()
    sun.dir[:] = list(sun_dir)
    sun.color[:] = [1.0, 0.9, 0.8]
    sun.intensity = sun_I
    pts = np.zeros(len(points), orc.POINT_DTYPE)
    for i, (pos, I, r) in enumerate(points):
        pts[i] = (pos, (1.0, 1.0, 1.0), I, r)
    return orc.Scene(t, np.array([albedo], np.float32), pts, sun, cam), cam, sun


def test_sun_on_open_plane_closed_form():
    sc, cam, sun = plane_scene()
    out = sc.render(1, 0)
    L = out["sum"]
    d = np.asarray(sun.dir[:], np.float64)
    ndotl = (-d / np.linalg.norm(d))[2]
    expect = np.array([0.25, 0.5, 0.75]) * np.array([1.0, 0.9, 0.8]) * 3.0 * ndotl
    hit = L.sum(-1) > 0
    assert hit.mean() > 0.9
    assert np.allclose(L[hit], expect, rtol=3e-6)


def test_point_light_falloff_closed_form():
    lpos, I, r = np.array([0.5, -0.5, 2.0]), 50.0, 4.0
    sc, cam, sun = plane_scene(sun_I=0.0, points=[(lpos, I, r)])
    hits, tt = sc.primary_hits()
    out = sc.render(1, 0)
    # reconstruct each pixel's hit point from the pixel-centre ray; render uses jittered rays,
    # so compare with spp=1 jitter recomputed from the same stream instead: use the G-buffer
    # independent check on the plane: L depends only on the hit point P.
    W, H = cam.width, cam.height
    th = np.float64(orc.tan_half_fov(cam.fov_y_deg))
    fwd, left, up = (np.asarray(cam.camera_to_world[4 * k:4 * k + 3], np.float64) for k in range(3))
    pos = np.asarray(cam.pos[:], np.float64)
    checked = 0
    for py in range(0, H, 3):
        for px in range(0, W, 3):
            u = orc.pixel_randf(0, py * W + px, 0, 2).astype(np.float64)
            ndcx = 2 * (px + u[0]) / W - 1
            ndcy = 1 - 2 * (py + u[1]) / H
            dw = fwd - left * ndcx * cam.aspect * th + up * ndcy * th
            t = -pos[2] / dw[2]
            P = pos + t * dw
            dist = np.linalg.norm(lpos - P)
            ndotl = max(0.0, (lpos - P)[2] / dist)
            e = I * ndotl * max(0.0, 1 - dist / r) ** 2 * np.array([0.25, 0.5, 0.75]) if dist < r else np.zeros(3)
            assert np.allclose(out["sum"][py, px], e, rtol=2e-5, atol=2e-5), (px, py)
            checked += e.sum() > 0
    assert checked > 20


def test_kahan_sum_matches_fp64_and_chunked_xorshift():
    import lgi
    ctx = lgi.Context()
    ctx.generate(1, 5, 48, 24)
    s = ctx.scene()
    sc = orc.Scene(s["tris"], s["materials"][:, 0, :], s["points"], s["sun"], s["camera"])
    one = sc.render(16, 2)
    assert np.all(np.abs(one["sum"] - one["sum64"]) <= 1e-6 * np.abs(one["sum64"]) + 1e-7)
    # xorshift-ref: one 16-spp call == two 8-spp launches (reseeded per 8-sample chunk)
    a = sc.render(16, 2, flags=1)
    b = sc.render(8, 2, flags=1)
    b = sc.render(8, 2, flags=1, sample0=8, sums=b)
    assert np.array_equal(a["sum"], b["sum"])
    # PCG32: any split of the sample range gives the same sums
    c = sc.render(5, 2)
    c = sc.render(11, 2, sample0=5, sums=c)
    assert np.array_equal(one["sum"], c["sum"])


def test_tonemap_closed_form():
    acc = np.array([[0, 0, 0], [4, 4, 4], [1, 0.25, 0.0625], [100, -1, 2.0]], np.float32)
    out = orc.tonemap(acc, np.float32(0.25))
    expect = [[0, 0, 0], [255, 255, 255], [127, 63, 31], [255, 0, 180]]
    assert out.tolist() == expect


def test_camera_look_at_matches_product_and_degenerate_cases():
    import lgi
    for pos, tgt in [((1, 2, 3), (4, -2, 1)), ((0, 0, 5), (0, 0, 0)), ((1, 1, 1), (1, 1, 1))]:
        a = lgi.Camera()
        a.camera_to_world[:] = list(range(16))
        b = orc.OCam()
        b.c2w[:] = list(range(16))
        lgi.camera_look_at(a, pos, tgt)
        orc.camera_look_at(b, pos, tgt)
        assert bytes(a)[:76] == bytes(b)[:76]
    # forward along -z: left falls back to (0, 1, 0) (src/main.cpp:43)
    a = lgi.Camera()
    lgi.camera_look_at(a, (0, 0, 5), (0, 0, 0))
    assert list(a.camera_to_world[4:7]) == [0.0, 1.0, 0.0]
    # degenerate target: camera unchanged (src/main.cpp:37)
    a = lgi.Camera()
    a.camera_to_world[0] = 7.0
    lgi.camera_look_at(a, (1, 1, 1), (1, 1, 1))
    assert a.camera_to_world[0] == 7.0
