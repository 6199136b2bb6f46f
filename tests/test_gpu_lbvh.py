"""GPU LBVH parity: node arrays bit-identical to the reference BuildBVH (golden vectors) and
to the CPU oracle; triangle order identical to the oracle's stable sort and equal to the
reference's modulo ties among equal Morton codes (std::sort order is unspecified)."""
import pathlib
import sys

import numpy as np
import pytest

from oracle import oracle as orc
from test_oracle_bvh import check_invariants, golden_cases, same_perm_modulo_ties

pytestmark = pytest.mark.gpu


def gpu_build(tris):
    import lgi
    tris = np.array(tris, copy=True)
    tris["material"] = 0
    ctx = lgi.Context()
    cam = lgi.Camera()
    cam.width, cam.height, cam.fov_y_deg, cam.aspect = 8, 8, 60.0, 1.0
    cam.camera_to_world[0], cam.camera_to_world[5], cam.camera_to_world[10] = 1, 1, 1
    sun = lgi.DirLight()
    ctx.set_scene(cam, sun, np.zeros(0, lgi.POINT_DTYPE), [((0.5, 0.5, 0.5), (0, 0, 0))], tris)
    ctx.build_bvh()
    nodes, perm, keys = ctx.bvh()
    # BuildBVH contract: the scene's triangles are now in Morton order
    assert np.array_equal(ctx.scene()["tris"].tobytes(), np.asarray(tris)[perm].tobytes())
    ctx.close()
    return nodes, perm, keys


@pytest.mark.parametrize("case", golden_cases(), ids=lambda c: c[0])
def test_gpu_lbvh_matches_reference_golden(gpu, case):
    name, tris, ref_nodes, ref_perm = case
    print(f"case {name}: {len(tris)} triangles", flush=True)
    nodes, perm, keys = gpu_build(tris)
    assert nodes.tobytes() == ref_nodes.tobytes(), f"{name}: node arrays differ"
    assert same_perm_modulo_ties(keys, perm, ref_perm)
    onodes, operm, okeys = orc.bvh_build(tris)
    assert np.array_equal(keys, okeys)
    assert np.array_equal(perm, operm)  # both sorts are stable


@pytest.mark.parametrize("kind,w,h", [(1, 512, 256), (2, 1728, 864)])
def test_gpu_lbvh_full_scenes_match_oracle(gpu, kind, w, h):
    import lgi
    print(f"scene kind {kind}", flush=True)
    ctx = lgi.Context()
    ctx.generate(kind, 1, w, h)
    tris = ctx.scene()["tris"]
    ctx.build_bvh()
    nodes, perm, keys = ctx.bvh()
    print(f"  {len(tris)} tris, {len(nodes)} nodes, GPU build {ctx.last_build_ms():.2f} ms", flush=True)
    onodes, operm, okeys = orc.bvh_build(tris)
    assert np.array_equal(keys, okeys)
    assert np.array_equal(perm, operm)
    assert nodes.tobytes() == onodes.tobytes()
    if kind == 1:
        check_invariants(tris, nodes, perm, keys)


def test_gpu_lbvh_rebuild_is_deterministic(gpu):
    import lgi
    ctx = lgi.Context()
    ctx.generate(1, 7, 64, 32)
    ctx.build_bvh()
    a = ctx.bvh()
    ctx.build_bvh()
    b = ctx.bvh()
    for x, y in zip(a, b):
        assert x.tobytes() == y.tobytes()
