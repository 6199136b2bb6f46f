"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer (CPU only): the hand-written
scene.json parser, binary scene loader, PNG reader and batch parser of liblgi (scene_host.cpp)
and the CPU oracle (oracle.cpp), driven by tests/sanitize/host_harness.cpp on well-formed
inputs, every truncation and thousands of byte-flipped variants. Any sanitizer report fails the
run (-fno-sanitize-recover)."""
import os
import pathlib
import shutil
import subprocess

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
PKG = ROOT / "luminagi-cudaref_amd"
HERE = pathlib.Path(__file__).resolve().parent


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ absent")
    out = tmp_path_factory.mktemp("asan") / "host_harness"
    cmd = ["g++", "-std=c++17", "-g", "-O1", "-fno-omit-frame-pointer", "-ffp-contract=off",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-pthread",
           f"-I{ROOT / 'include'}", f"-I{PKG / 'csrc'}", f"-I{ROOT / 'oracle'}",
           str(HERE / "sanitize" / "host_harness.cpp"), str(PKG / "csrc" / "scene_host.cpp"), "-lz", "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return out


def test_host_parsers_and_oracle_under_asan_ubsan(harness, tmp_path):
    from PIL import Image
    rng = np.random.default_rng(1)
    pngs = []
    for mode, shape in (("L", (9, 7)), ("LA", (9, 7)), ("RGB", (9, 7, 3)), ("RGBA", (9, 7, 4)), ("P", (9, 7)),
                        ("I;16", (9, 7))):
        arr = rng.integers(0, 256 if mode != "I;16" else 65536, shape).astype(np.uint16 if mode == "I;16" else np.uint8)
        img = Image.fromarray(arr).convert(mode) if mode in ("LA", "P") else Image.fromarray(arr)
        p = tmp_path / f"in_{mode.replace(';', '')}.png"
        img.save(p)
        pngs.append(str(p))
    work = tmp_path / "work"
    work.mkdir()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(harness), str(work), str(HERE / "golden" / "mock_scene.json")] + pngs,
                       capture_output=True, text=True, env=env, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "harness ok" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
