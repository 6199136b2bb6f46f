"""Binary scene file: bit-exact round trips against the JSON path, error codes, and the load
time next to scene.json for the 1.05M-triangle scene (the reference's JSON load is the step its
README times at 12 s for 300 MB)."""
import pathlib
import time

import numpy as np
import pytest

import lgi

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def scene_arrays(ctx):
    s = ctx.scene()
    return bytes(s["camera"]), bytes(s["sun"]), s["points"].tobytes(), s["materials"].tobytes(), s["tris"].tobytes()


def test_binary_round_trip_is_bit_exact(tmp_path):
    a = lgi.Context()
    a.generate(1, 11, 320, 200)
    a.save_bin(tmp_path / "s.lgis")
    b = lgi.Context()
    b.load_bin(tmp_path / "s.lgis")
    assert scene_arrays(a) == scene_arrays(b)
    assert a.counts() == b.counts()


def test_binary_keeps_everything_json_keeps(tmp_path):
    """mock_scene.json (directional + point lights, a textured material) -> bin -> JSON gives
    the same JSON text as JSON -> JSON."""
    a = lgi.Context()
    a.load_json(GOLDEN / "mock_scene.json")
    a.save_json(tmp_path / "a.json")
    a.save_bin(tmp_path / "a.lgis")
    b = lgi.Context()
    b.load_bin(tmp_path / "a.lgis")
    b.save_json(tmp_path / "b.json")
    assert (tmp_path / "a.json").read_text() == (tmp_path / "b.json").read_text()
    assert "stone.png" in (tmp_path / "b.json").read_text()


def test_binary_errors(tmp_path):
    ctx = lgi.Context()
    with pytest.raises(lgi.LgiError) as e:
        ctx.load_bin(tmp_path / "missing.lgis")
    assert e.value.status == 2
    (tmp_path / "junk.lgis").write_bytes(b"not a scene at all" * 8)
    with pytest.raises(lgi.LgiError) as e:
        ctx.load_bin(tmp_path / "junk.lgis")
    assert e.value.status == 3
    ctx.generate(1, 3, 64, 32)
    ctx.save_bin(tmp_path / "ok.lgis")
    raw = bytearray((tmp_path / "ok.lgis").read_bytes())
    (tmp_path / "cut.lgis").write_bytes(bytes(raw[:-100]))  # truncated triangles
    with pytest.raises(lgi.LgiError) as e:
        ctx.load_bin(tmp_path / "cut.lgis")
    assert e.value.status == 3
    raw[8] = 9  # version field
    (tmp_path / "ver.lgis").write_bytes(bytes(raw))
    with pytest.raises(lgi.LgiError) as e:
        ctx.load_bin(tmp_path / "ver.lgis")
    assert e.value.status == 4


def test_load_times_json_vs_binary(tmp_path):
    ctx = lgi.Context()
    ctx.generate(2, 1, 1728, 864)
    ctx.save_json(tmp_path / "big.json")
    ctx.save_bin(tmp_path / "big.lgis")
    times = {}
    for kind, path, load in (("json", "big.json", "load_json"), ("bin", "big.lgis", "load_bin")):
        c = lgi.Context()
        t = time.perf_counter()
        getattr(c, load)(tmp_path / path)
        times[kind] = time.perf_counter() - t
        assert c.counts()[0] == 1051462
    mb = {k: (tmp_path / p).stat().st_size / 1e6 for k, p in (("json", "big.json"), ("bin", "big.lgis"))}
    print(f"1.05M triangles: scene.json {mb['json']:.0f} MB loads in {times['json']:.2f} s; "
          f"binary {mb['bin']:.0f} MB in {times['bin']:.3f} s", flush=True)
    assert times["bin"] < times["json"]
