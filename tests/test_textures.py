"""Diffuse textures on the CPU: the PNG reader against PIL and against hand-filtered files, and
the oracle's sampler (sample_albedo with the reference's texture object, src/pathtracer.cu:
274-282, 563-576) against closed forms: the sRGB curve, texel centres, wrap, the v flip, the
bilinear midpoint and the 1/256 weight quantisation bound."""
import struct
import zlib

import numpy as np
import pytest

import lgi
from oracle import oracle as orc


def png_bytes(rows, color_type, depth=8, filters=(0,), palette=None):
    """Minimal PNG writer that applies the given filter types row by row (cycled)."""
    h = len(rows)
    bpp = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[color_type] * depth // 8
    raw = bytearray()
    prev = bytes(len(rows[0]))
    for y, row in enumerate(rows):
        ft = filters[y % len(filters)]
        out = bytearray(len(row))
        for x in range(len(row)):
            a = row[x - bpp] if x >= bpp else 0
            b = prev[x]
            c = prev[x - bpp] if x >= bpp else 0
            if ft == 0:
                p = 0
            elif ft == 1:
                p = a
            elif ft == 2:
                p = b
            elif ft == 3:
                p = (a + b) >> 1
            else:
                pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
            out[x] = (row[x] - p) & 255
        raw += bytes([ft]) + out
        prev = row

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)
    w = len(rows[0]) // bpp
    o = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, color_type, 0, 0, 0))
    if palette is not None:
        o += chunk(b"PLTE", bytes(palette))
    return o + chunk(b"IDAT", zlib.compress(bytes(raw))) + chunk(b"IEND", b"")


@pytest.mark.parametrize("mode", ["L", "LA", "RGB", "RGBA", "P"])
def test_png_reader_matches_pil(tmp_path, mode):
    from PIL import Image
    rng = np.random.default_rng(3)
    ch = {"L": 1, "LA": 2, "RGB": 3, "RGBA": 4, "P": 1}[mode]
    a = rng.integers(0, 256, (23, 37, ch), dtype=np.uint8)
    img = Image.fromarray(a[..., 0] if ch == 1 else a, mode if mode != "P" else "L")
    if mode == "P":
        img = img.convert("P")
    img.save(tmp_path / "t.png")
    got = lgi.read_png(tmp_path / "t.png")
    want = np.asarray(Image.open(tmp_path / "t.png").convert("RGBA"))
    assert np.array_equal(got, want)


def test_png_reader_all_filters_and_16_bit(tmp_path):
    rng = np.random.default_rng(4)
    rgb = rng.integers(0, 256, (11, 9, 3), dtype=np.uint8)
    rows = [bytes(r.reshape(-1)) for r in rgb]
    (tmp_path / "f.png").write_bytes(png_bytes(rows, 2, filters=(0, 1, 2, 3, 4)))
    got = lgi.read_png(tmp_path / "f.png")
    assert np.array_equal(got[..., :3], rgb) and (got[..., 3] == 255).all()
    g16 = rng.integers(0, 65536, (5, 7), dtype=np.uint16)
    rows16 = [g16[y].astype(">u2").tobytes() for y in range(5)]
    (tmp_path / "g.png").write_bytes(png_bytes(rows16, 0, depth=16, filters=(4, 3)))
    got = lgi.read_png(tmp_path / "g.png")
    assert np.array_equal(got[..., 0], (g16 >> 8).astype(np.uint8))  # high byte, as stbi_load(..., 4)
    pal = rng.integers(0, 256, (4, 3), dtype=np.uint8)
    idx = rng.integers(0, 4, (6, 5), dtype=np.uint8)
    (tmp_path / "p.png").write_bytes(png_bytes([bytes(r) for r in idx], 3, palette=pal.reshape(-1)))
    assert np.array_equal(lgi.read_png(tmp_path / "p.png")[..., :3], pal[idx])


def test_png_reader_errors(tmp_path):
    with pytest.raises(lgi.LgiError) as e:
        lgi.read_png(tmp_path / "missing.png")
    assert e.value.status == 2
    (tmp_path / "bad.png").write_bytes(b"\x89PNG\r\n\x1a\nnot really")
    with pytest.raises(lgi.LgiError) as e:
        lgi.read_png(tmp_path / "bad.png")
    assert e.value.status == 3


def test_srgb_table_closed_form():
    lut = orc.srgb_lut()
    assert lut[0] == 0.0 and lut[255] == 1.0
    assert abs(lut[128] - 0.21586050011389926) < 1e-7      # ((128/255 + 0.055) / 1.055) ** 2.4
    assert lut[10] == np.float32(10 / 255 / 12.92)          # linear segment below 0.04045
    assert np.all(np.diff(lut) > 0)


def rgba(a):
    a = np.asarray(a, np.uint8)
    return np.concatenate([a, np.full(a.shape[:2] + (1,), 255, np.uint8)], -1)


def test_sampler_closed_forms():
    lut = orc.srgb_lut()
    rng = np.random.default_rng(5)
    W, H = 8, 4
    tex = rng.integers(0, 256, (H, W, 3), dtype=np.uint8)
    mats = np.array([[0.2, 0.3, 0.4], [0.5, 0.5, 0.5]], np.float32)
    T = {1: rgba(tex)}
    # material without texture: constant albedo
    assert np.array_equal(orc.tex_sample(T, mats, 0, 0.3, 0.7), mats[0])
    # texel centres give the texel itself; row j of the image is at v = 1 - (j + 0.5) / H
    for (i, j) in [(0, 0), (3, 1), (7, 3), (5, 2)]:
        got = orc.tex_sample(T, mats, 1, (i + 0.5) / W, 1.0 - (j + 0.5) / H)
        assert np.array_equal(got, lut[tex[j, i]]), (i, j)
    # wrap: whole-number shifts of u and v sample the same point
    u, v = 0.375, 0.625
    a = orc.tex_sample(T, mats, 1, u, v)
    for du, dv in [(1, 0), (-2, 1), (3, -3)]:
        assert np.array_equal(orc.tex_sample(T, mats, 1, u + du, v + dv), a)
    # halfway between two texel centres along x: the mean of the two texels
    i, j = 2, 1
    got = orc.tex_sample(T, mats, 1, (i + 1.0) / W, 1.0 - (j + 0.5) / H)
    want = np.float32(0.5) * lut[tex[j, i]] + np.float32(0.5) * lut[tex[j, i + 1]]
    assert np.allclose(got, want, rtol=0, atol=1e-7)
    # left edge: u = 0 is halfway between the last and the first texel (wrap addressing)
    got = orc.tex_sample(T, mats, 1, 0.0, 1.0 - (j + 0.5) / H)
    want = np.float32(0.5) * lut[tex[j, W - 1]] + np.float32(0.5) * lut[tex[j, 0]]
    assert np.allclose(got, want, rtol=0, atol=1e-7)
    # constant texture: the constant, for any coordinates
    C = {1: rgba(np.full((H, W, 3), 77, np.uint8))}
    for uv in rng.random((20, 2)) * 5 - 2:
        assert np.allclose(orc.tex_sample(C, mats, 1, *uv), lut[77], rtol=2e-7, atol=0)


def test_sampler_weight_quantisation_bound():
    """Against exact bilinear filtering: the 1/256 weights move the result by at most half a
    step of the largest neighbour difference."""
    lut = orc.srgb_lut().astype(np.float64)
    rng = np.random.default_rng(6)
    W, H = 16, 16
    tex = rng.integers(0, 256, (H, W, 3), dtype=np.uint8)
    mats = np.zeros((1, 3), np.float32)
    lin = lut[tex]
    for u, v in rng.random((200, 2)):
        x = u * W - 0.5
        y = (1 - v) * H - 0.5
        x0, y0 = int(np.floor(x)), int(np.floor(y))
        a, b = x - x0, y - y0
        x0, x1, y0, y1 = x0 % W, (x0 + 1) % W, y0 % H, (y0 + 1) % H
        exact = (1 - b) * ((1 - a) * lin[y0, x0] + a * lin[y0, x1]) + b * ((1 - a) * lin[y1, x0] + a * lin[y1, x1])
        got = orc.tex_sample({0: rgba(tex)}, mats, 0, u, v)
        spread = np.ptp([lin[y0, x0], lin[y0, x1], lin[y1, x0], lin[y1, x1]], axis=0)
        assert np.all(np.abs(got - exact) <= spread / 256 + 1e-6)
