"""Host-side checks of liblgi.so through the C ABI (no GPU needed): exports, scene.json
loading and its error codes (src/scene_loader.cpp:26-138), JSON round trip, batch files
(src/main.cpp:61-91), PNG output, and the procedural scene generator."""
import ctypes as C
import pathlib
import subprocess

import numpy as np
import pytest

import lgi
from lgi import _abi

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
PKG = pathlib.Path(lgi.__file__).resolve().parent.parent


def test_library_exports_every_header_symbol():
    syms = _abi.header_symbols()
    assert len(syms) >= 20
    l = lgi.lib()
    for s in syms:
        assert hasattr(l, s), s
        assert s in _abi.SIGNATURES, f"{s} has no ctypes signature"
    assert lgi.lib().lgi_version().startswith(b"lgi")


def test_struct_layouts():
    assert C.sizeof(lgi.Camera) == 100
    assert C.sizeof(_abi.PointLight) == 32
    assert C.sizeof(_abi.DirLight) == 28
    assert lgi.TRI_DTYPE.itemsize == 100 and lgi.NODE_DTYPE.itemsize == 40


def test_mock_scene_fixture_loads_like_reference():
    # scenes/mock_scene.json of the reference: 1920x1080, fov 60, 1 directional + 1 point
    # light, 2 materials, 2 triangles (summary format src/scene_loader.cpp:130-136)
    ctx = lgi.Context()
    ctx.load_json(GOLDEN / "mock_scene.json")
    assert ctx.counts() == (2, 2, 1, 1)
    s = ctx.scene()
    cam = s["camera"]
    assert (cam.width, cam.height, cam.fov_y_deg) == (1920, 1080, 60.0)
    assert cam.aspect == np.float32(1.777778)
    assert list(cam.camera_to_world[12:16]) == [0.0, -5.0, 2.0, 1.0]
    assert s["points"]["radius"][0] == 5.0 and s["points"]["intensity"][0] == np.float32(0.8)
    assert list(s["sun"].dir) == [1.0, -1.0, -1.0]
    t = s["tris"]
    assert t["v"][1].tolist() == [[-2, -2, 0], [2, -2, 0], [0, 2, 5]]
    assert t["uv"][1][2].tolist() == [0.5, 1.0] and t["material"].tolist() == [0, 1]


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return p


def test_load_errors_return_codes(tmp_path):
    ctx = lgi.Context()
    good = (GOLDEN / "mock_scene.json").read_text()
    cases = [
        (tmp_path / "missing.json", 2),
        (write(tmp_path, "bad.json", good[:200]), 3),
        (write(tmp_path, "v2.json", good.replace('"version": 1', '"version": 2')), 4),
        (write(tmp_path, "nov.json", good.replace('"version": 1,', '')), 4),
        (write(tmp_path, "nocam.json", good.replace('"camera"', '"kamera"')), 3),
        (write(tmp_path, "nomat.json", good.replace('"mat":1', '"matt":1')), 3),
        (write(tmp_path, "noradius.json", good.replace('"radius": 5.0', '"r": 5.0')), 3),
    ]
    for path, code in cases:
        with pytest.raises(lgi.LgiError) as e:
            ctx.load_json(path)
        assert e.value.status == code, path.name
    ctx.load_json(GOLDEN / "mock_scene.json")  # context still usable


def test_json_round_trip_bit_exact(tmp_path):
    ctx = lgi.Context()
    ctx.generate(1, 3, 320, 160)
    a = ctx.scene()
    ctx.save_json(tmp_path / "s.json")
    ctx2 = lgi.Context()
    ctx2.load_json(tmp_path / "s.json")
    b = ctx2.scene()
    assert a["tris"].tobytes() == b["tris"].tobytes()
    assert a["points"].tobytes() == b["points"].tobytes()
    assert np.array_equal(a["materials"], b["materials"])
    assert bytes(a["camera"]) == bytes(b["camera"]) and bytes(a["sun"]) == bytes(b["sun"])


def test_generator_sizes_and_determinism():
    ctx = lgi.Context()
    ctx.generate(1, 1, 512, 256)
    n1, m1, p1, _ = ctx.counts()
    assert 53000 <= n1 <= 82000 and m1 == 17 and p1 == 2
    t1 = ctx.scene()["tris"]
    ctx.generate(1, 1, 512, 256)
    assert ctx.scene()["tris"].tobytes() == t1.tobytes()
    ctx.generate(2, 1, 1728, 864)
    n2, m2, p2, _ = ctx.counts()
    assert abs(n2 - 1_050_000) < 0.01 * 1_050_000 and m2 == 63 and p2 == 2
    s = ctx.scene()
    assert np.all(s["points"]["intensity"] == 50) and s["sun"].intensity == 3.0
    assert s["sun"].dir[2] < 0  # light travels downwards
    alb = s["materials"][:, 0, :]
    assert alb.min() >= 0.1 and alb.max() <= 0.9
    with pytest.raises(lgi.LgiError):
        ctx.generate(3, 1, 16, 16)


def test_batch_file_parsing(tmp_path):
    text = ("# out_png spp bounces  camx camy camz  tgtx tgty tgtz\n"
            "out/a.png 8 2 1 2 3 4 5 6\n"
            "out/b.png 256 2 -1.5 0 0.25 0 0 0  # trailing comment\n"
            "broken line 1 2\n"
            "\n"
            "out/c.png 4 0 0 0 0 1 0 0 extra tokens ignored\n")
    jobs = lgi.parse_batch(write(tmp_path, "batch.txt", text))
    assert [j[0] for j in jobs] == ["out/a.png", "out/b.png", "out/c.png"]
    assert jobs[0][1:3] == (8, 2) and jobs[1][1] == 256 and jobs[2][2] == 0
    assert np.allclose(jobs[1][3], (-1.5, 0, 0.25))
    with pytest.raises(lgi.LgiError):
        lgi.parse_batch(tmp_path / "nope.txt")


def test_reference_batch_recipe(reference_dir):
    jobs = lgi.parse_batch(reference_dir / "ml" / "data" / "batch.txt")
    assert len(jobs) == 100
    assert sorted({j[1] for j in jobs}) == [8, 256]
    assert len({j[3] for j in jobs}) == 50


def test_png_writer_round_trip(tmp_path):
    from PIL import Image
    rng = np.random.default_rng(0)
    img = rng.integers(0, 256, (37, 53, 3), dtype=np.uint8)
    lgi.write_png(tmp_path / "x.png", img)
    back = np.asarray(Image.open(tmp_path / "x.png").convert("RGB"))
    assert np.array_equal(back, img)


def test_cli_usage_and_load_failure(tmp_path):
    cli = pathlib.Path(lgi.__file__).resolve().parent.parent / "bin" / "cuda_ref"
    r = subprocess.run([str(cli)], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr
    r = subprocess.run([str(cli), str(tmp_path / "missing.json")], capture_output=True, text=True)
    assert r.returncode in (2, 4)


def test_validate_tool_argument_errors(tmp_path):
    """validate.py names every missing input of a capture and stops before any GPU work."""
    import sys
    tool = PKG / "tools" / "validate.py"
    r = subprocess.run([sys.executable, str(tool), str(tmp_path / "none.json")], capture_output=True, text=True)
    assert r.returncode == 1 and "not found" in r.stderr
    assert "none.json" in r.stderr and "none.png" in r.stderr
    (tmp_path / "s.json").write_text("{}")
    r = subprocess.run([sys.executable, str(tool), str(tmp_path / "s.json")], capture_output=True, text=True)
    assert r.returncode == 1 and "s.png" in r.stderr and "s.json" not in r.stderr


def test_probe_rejects_bad_arguments():
    """lgi_probe_read_bw validates on the host before touching the device."""
    L = _abi.lib()
    sink = C.c_void_p(64)
    assert L.lgi_probe_read_bw(None, 1 << 20, 1, 8, sink, None, None) == 1
    assert L.lgi_probe_read_bw(C.c_void_p(4096 + 8), 1 << 20, 1, 8, sink, None, None) == 1  # not 16-byte aligned
    assert L.lgi_probe_read_bw(C.c_void_p(4096), 8, 1, 8, sink, None, None) == 1           # under one word
    assert L.lgi_probe_read_bw(C.c_void_p(4096), 1 << 20, 0, 8, sink, None, None) == 1     # no passes


def test_treelet_passes_argument_checks():
    """lgi_bvh_set_treelet_passes accepts 0..8; the traversal tree exists only after a build."""
    ctx = lgi.Context()
    for bad in (-1, 9):
        with pytest.raises(lgi.LgiError) as e:
            ctx.set_treelet_passes(bad)
        assert e.value.status == 1  # LGI_ERR_INVALID_ARG
    for ok in (0, 8, 2):
        ctx.set_treelet_passes(ok)
    with pytest.raises(lgi.LgiError) as e:
        ctx.bvh_traversal()
    assert e.value.status == 7  # LGI_ERR_NO_BVH


def test_wavefront_pass_plan_bounds_shadow_candidates():
    """The sorted renderer's samples per pass (lgi_wavefront_samples_per_pass): ~100M paths, at most
    3 x that many shadow-ray candidates (slot x light), never 2^31 or more of them; an image whose
    single sample already exceeds that is refused."""
    from lgi.tracer import wavefront_samples_per_pass as plan
    max_paths = 96 << 20
    slots = 216 * 108 * 64                     # 1728x864 in 8x8 tiles
    assert plan(slots, 2, 64) == 64            # the bench frame is one pass
    assert plan(slots, 2, 1000) == max_paths // slots
    for n_points in (0, 2, 21, 22, 44, 45, 100, 1000):
        s = plan(slots, n_points, 4096)
        assert s >= 1
        assert s * slots <= max_paths
        assert s == 1 or s * slots * (1 + n_points) <= 3 * max_paths
        assert s * slots * (1 + n_points) < 2 ** 31
    assert plan(64, 0, 7) == 7
    # one sample of a 2^30-pixel image with 2 lights exceeds 2^31 candidates: refused
    with pytest.raises(lgi.LgiError) as e:
        plan(1 << 30, 2, 1)
    assert e.value.status == 8
    with pytest.raises(lgi.LgiError) as e:
        plan(slots, 20000, 1)
    assert e.value.status == 8
    with pytest.raises(lgi.LgiError):
        plan(0, 2, 1)


def test_conv_variant_table_matches_the_header():
    """lgi_unet_conv_tune_table writes LGI_UNET_CONV_VARIANTS timings per row into the caller's
    array; the Python front end sizes that array from its variant table."""
    import re
    from lgi import unet
    hdr = (PKG.parent / "include" / "lgi.h").read_text()
    n = int(re.search(r"#define LGI_UNET_CONV_VARIANTS (\d+)", hdr).group(1))
    assert len(unet.CONV_VARIANTS) == n and sorted(unet.CONV_VARIANTS) == list(range(n))

