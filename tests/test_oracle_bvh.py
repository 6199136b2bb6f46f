"""Pins the oracle's BuildBVH restatement (src/bvh.cpp) on the CPU.

* bit interleave: brute force over all 1024 values per axis;
* quantiser edge cases (src/bvh.cpp:25-30);
* golden vectors written by the reference's own BuildBVH (tests/golden/make_golden.py);
* the full ~56k-triangle test scene against the live reference build when it is mounted;
* structural invariants (every triangle in exactly one leaf, leaf rule, boxes, BFS order).
Triangle permutations are compared as multisets per run of equal Morton codes, because the
reference sorts with std::sort, whose tie order is unspecified.
"""
import pathlib

import numpy as np
import pytest

from oracle import oracle as orc

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden" / "bvh_golden.npz"


def naive_spread(v):
    out = 0
    for b in range(10):
        if v >> b & 1:
            out |= 1 << (3 * b)
    return out


def test_spread3_all_inputs():
    lib = orc.lib()
    for v in range(1024):
        assert lib.orc_spread3(v) == naive_spread(v)


def test_morton_interleave_and_quantiser():
    lib = orc.lib()
    # quant10: f*1024, clamp to [0, 1023], truncate
    assert lib.orc_morton3(0.0, 0.0, 0.0) == 0
    assert lib.orc_morton3(1.0, 1.0, 1.0) == (1 << 30) - 1          # clamped to 1023 on every axis
    assert lib.orc_morton3(-0.5, 2.0, 0.0) == naive_spread(1023) << 1  # negative -> 0, >1 -> 1023
    assert lib.orc_morton3(1023.5 / 1024, 0, 0) == naive_spread(1023) << 2
    assert lib.orc_morton3(1.0 / 1024, 0, 0) == naive_spread(1) << 2
    assert lib.orc_morton3(0.99 / 1024, 0, 0) == 0                   # truncation, not rounding
    rng = np.random.default_rng(3)
    for x, y, z in rng.uniform(0, 1, (200, 3)).astype(np.float32):
        qx, qy, qz = (min(1023, int(np.float32(c) * np.float32(1024))) for c in (x, y, z))
        assert lib.orc_morton3(x, y, z) == (naive_spread(qx) << 2) | (naive_spread(qy) << 1) | naive_spread(qz)


def same_perm_modulo_ties(keys, perm_a, perm_b):
    """perm_a and perm_b agree as multisets on every run of equal sorted keys."""
    if len(perm_a) != len(perm_b):
        return False
    bounds = np.flatnonzero(np.diff(keys.astype(np.int64))) + 1
    for run_a, run_b in zip(np.split(perm_a, bounds), np.split(perm_b, bounds)):
        if not np.array_equal(np.sort(run_a), np.sort(run_b)):
            return False
    return True


def golden_cases():
    g = np.load(GOLDEN)
    names = sorted({k.split("__")[0] for k in g.files})
    return [(n, g[f"{n}__tris"], g[f"{n}__nodes"], g[f"{n}__perm"]) for n in names]


@pytest.mark.parametrize("case", golden_cases(), ids=lambda c: c[0])
def test_oracle_matches_reference_golden(case):
    name, tris, ref_nodes, ref_perm = case
    nodes, perm, keys = orc.bvh_build(tris)
    assert nodes.tobytes() == ref_nodes.tobytes(), f"{name}: node arrays differ"
    assert same_perm_modulo_ties(keys, perm, ref_perm)
    # keys are the sorted Morton codes of the original triangles
    assert np.array_equal(keys, np.sort(orc.morton_codes(tris), kind="stable"))
    assert np.array_equal(orc.morton_codes(tris)[perm], keys)


def check_invariants(tris, nodes, perm, keys):
    n = len(tris)
    assert sorted(perm.tolist()) == list(range(n))
    assert np.all(np.diff(keys.astype(np.int64)) >= 0)
    covered = np.zeros(n, np.int32)
    leaf = nodes["tri_count"] > 0
    for nd in nodes[leaf]:
        f, c = nd["first_tri"], nd["tri_count"]
        covered[f:f + c] += 1
        if c > 4:  # only a run of identical codes may exceed 4 (src/bvh.cpp:92)
            assert keys[f] == keys[f + c - 1]
        v = np.asarray(tris["v"])[perm[f:f + c]].reshape(-1, 3)
        assert np.array_equal(nd["bb_min"], v.min(0)) and np.array_equal(nd["bb_max"], v.max(0))
    assert np.all(covered == 1)
    internal = np.flatnonzero(~leaf)
    for i in internal:
        l, r = nodes["left"][i], nodes["right"][i]
        assert i < l < r or (l == r - 1 and l > i)
        assert np.array_equal(nodes["bb_min"][i], np.minimum(nodes["bb_min"][l], nodes["bb_min"][r]))
        assert np.array_equal(nodes["bb_max"][i], np.maximum(nodes["bb_max"][l], nodes["bb_max"][r]))
        assert nodes["first_tri"][i] == -1 and nodes["tri_count"][i] == 0
    # BFS: children of consecutive internal nodes are consecutive pairs after the root
    kids = np.stack([nodes["left"][internal], nodes["right"][internal]], 1).reshape(-1)
    assert np.array_equal(kids, np.arange(1, len(nodes)))


@pytest.mark.parametrize("case", golden_cases(), ids=lambda c: c[0])
def test_oracle_invariants(case):
    _, tris, _, _ = case
    nodes, perm, keys = orc.bvh_build(tris)
    check_invariants(tris, nodes, perm, keys)


def test_oracle_full_scene_matches_live_reference(reference_dir, tmp_path):
    if not orc.ref_bvh_available():
        pytest.skip("oracle/_ref not built")
    import lgi
    ctx = lgi.Context()
    ctx.generate(1, 1, 512, 256)
    js = tmp_path / "s1.json"
    ctx.save_json(js)
    tris = ctx.scene()["tris"]
    nodes, perm, keys = orc.bvh_build(tris)
    ref_nodes, ref_perm = orc.ref_bvh(js, tmp_path / "s1.bin")
    assert nodes.tobytes() == ref_nodes.tobytes()
    assert same_perm_modulo_ties(keys, perm, ref_perm)
    check_invariants(tris, nodes, perm, keys)


def test_bvh_closest_equals_brute_force_small():
    """BVH traversal returns the brute-force closest triangle (ties excepted)."""
    g = np.load(GOLDEN)
    tris = g["subset2000__tris"]
    rng = np.random.default_rng(11)
    nodes, perm, _ = orc.bvh_build(tris)
    sorted_tris = np.ascontiguousarray(tris[perm])
    cent = sorted_tris["v"].mean(1)
    hits = 0
    for k in range(400):
        o = rng.uniform(-25, 25, 3).astype(np.float32)
        d = (cent[rng.integers(len(cent))] - o).astype(np.float32) if k % 4 else rng.normal(size=3).astype(np.float32)
        ib, tb = orc.bvh_closest(nodes, sorted_tris, o, d)
        ia, ta = orc.brute_closest(sorted_tris, o, d)
        assert tb == ta
        if ia >= 0:
            hits += 1
            assert ib == ia
    assert hits > 200
