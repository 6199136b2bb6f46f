"""Writes the BVH golden vectors in this folder from the reference's own BuildBVH.

Run in the build container, where /root/reference is mounted and `make -C oracle` has
compiled oracle/_ref/ref_bvh from the reference's src/bvh.cpp + src/scene_loader.cpp:

    python tests/golden/make_golden.py

Each case stores the input triangles (Triangle layout), the reference node array (BvhNode
layout, BFS order) and the reference triangle permutation (original index per sorted slot).
"""
import json
import pathlib
import subprocess
import sys
import tempfile

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "luminagi-cudaref_amd"))
sys.path.insert(0, str(ROOT))

from oracle import oracle as orc  # noqa: E402

OUT = pathlib.Path(__file__).resolve().parent / "bvh_golden.npz"


def scene_json(tris, path):
    """Minimal v1 scene.json around a triangle list (schema of src/scene_loader.cpp:53-128)."""
    doc = {
        "version": 1,
        "camera": {"pos": [0, -5, 2], "camera_to_world": [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, -5, 2, 1],
                   "fov_y_deg": 60.0, "aspect": 2.0, "near": 0.1, "far": 1000.0,
                   "image_width": 64, "image_height": 32},
        "sun": {"direction": [0.3, 0.2, -1.0], "color": [1, 1, 1], "intensity": 3.0},
        "lights": [],
        "materials": [{"albedo": [0.5, 0.5, 0.5], "emissive": [0, 0, 0]}],
        "triangles": [{"v": t["v"].astype(float).tolist(), "n": t["n"].astype(float).tolist(),
                       "uv": t["uv"].astype(float).tolist(), "mat": 0} for t in tris],
    }
    pathlib.Path(path).write_text(json.dumps(doc))


def rand_tris(rng, n, spread=10.0, size=0.3):
    t = np.zeros(n, orc.TRI_DTYPE)
    c = rng.uniform(-spread, spread, (n, 1, 3))
    t["v"] = (c + rng.normal(0, size, (n, 3, 3))).astype(np.float32)
    t["n"][:, :, 2] = 1.0
    return t


def cases():
    rng = np.random.default_rng(20261203)
    import lgi
    ctx = lgi.Context()
    ctx.generate(1, 1, 64, 32)
    s1 = ctx.scene()["tris"]
    out = {"subset2000": s1[:2000].copy()}
    base = rand_tris(rng, 60)
    dup = np.concatenate([base] + [base[:7]] * 9)  # runs of identical Morton codes (> 4 per leaf)
    rng.shuffle(dup)
    out["duplicates"] = dup
    flat = rand_tris(rng, 300)
    flat["v"][:, :, 2] = 0.0  # zero z extent -> ext = 1 (src/bvh.cpp:178)
    out["flat"] = flat
    for n in (1, 2, 3, 4, 5, 7, 9):
        out[f"tiny{n}"] = rand_tris(rng, n)
    same = rand_tris(rng, 1)
    out["all_equal"] = np.repeat(same, 12)
    return out


def main():
    if not orc.ref_bvh_available():
        sys.exit("oracle/_ref/ref_bvh missing: run `make -C oracle` with the reference mounted")
    arrays = {}
    with tempfile.TemporaryDirectory() as td:
        for name, tris in cases().items():
            js, bn = pathlib.Path(td) / f"{name}.json", pathlib.Path(td) / f"{name}.bin"
            scene_json(tris, js)
            nodes, perm = orc.ref_bvh(js, bn)
            arrays[f"{name}__tris"] = tris
            arrays[f"{name}__nodes"] = nodes
            arrays[f"{name}__perm"] = perm
            print(f"{name}: {len(tris)} tris -> {len(nodes)} nodes")
    np.savez_compressed(OUT, **arrays)
    print("wrote", OUT, OUT.stat().st_size, "bytes")


if __name__ == "__main__":
    main()
