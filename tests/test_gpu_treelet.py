"""The traversal tree: the reference LBVH restructured above its leaves (5-leaf treelets, after
Karras and Aila 2013). Checks it against the reference node array (same leaves in the same
slots, a proper tree, every box the exact union of its children's, depth within the traversal
stacks), its surface-area cost against the reference tree's, closest hits through it (the
oracle's traversal) against brute force, and renders through it against renders through the
reference tree itself."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

MAX_DEPTH = 31  # one traversal-stack entry per internal level, 32 entries


def levels(nodes):
    """Breadth-first levels from the root (numpy frontier expansion); fails on a cycle."""
    internal = nodes["tri_count"] == 0
    out, frontier, seen = [], np.array([0]), 0
    while frontier.size:
        out.append(frontier)
        seen += frontier.size
        assert seen <= len(nodes), "a slot is reached twice"
        f = frontier[internal[frontier]]
        frontier = np.concatenate([nodes["left"][f], nodes["right"][f]])
    return out


def sah(nodes, lv, c_node=1.2, c_tri=1.0):
    d = nodes["bb_max"].astype(np.float64) - nodes["bb_min"]
    area = d[:, 0] * d[:, 1] + d[:, 1] * d[:, 2] + d[:, 2] * d[:, 0]
    cost = np.zeros(len(nodes))
    for f in reversed(lv):
        leaf = nodes["tri_count"][f] > 0
        fl, fi = f[leaf], f[~leaf]
        cost[fl] = area[fl] * c_tri * nodes["tri_count"][fl]
        cost[fi] = area[fi] * c_node + cost[nodes["left"][fi]] + cost[nodes["right"][fi]]
    return cost[0]


def check_traversal_tree(ref, trav):
    assert len(ref) == len(trav)
    leaf = ref["tri_count"] > 0
    assert np.array_equal(trav["tri_count"], ref["tri_count"])
    assert trav[leaf].tobytes() == ref[leaf].tobytes(), "leaf records moved or changed"
    lv = levels(trav)
    assert sum(f.size for f in lv) == len(trav), "a slot is unreachable"
    assert len(lv) - 1 <= MAX_DEPTH
    internal = ~leaf
    L, R = trav["left"][internal], trav["right"][internal]
    assert np.array_equal(trav["bb_min"][internal], np.minimum(trav["bb_min"][L], trav["bb_min"][R]))
    assert np.array_equal(trav["bb_max"][internal], np.maximum(trav["bb_max"][L], trav["bb_max"][R]))
    assert np.array_equal(trav["bb_min"][0], ref["bb_min"][0]) and np.array_equal(trav["bb_max"][0], ref["bb_max"][0])
    return lv


@pytest.mark.parametrize("kind,w,h", [(1, 512, 256), (2, 1728, 864)])
def test_traversal_tree_is_the_reference_tree_restructured(gpu, kind, w, h):
    import lgi
    print(f"scene kind {kind}", flush=True)
    ctx = lgi.Context()
    ctx.generate(kind, 1, w, h)
    ctx.build_bvh()
    ref, _, _ = ctx.bvh()
    trav = ctx.bvh_traversal()
    lv = check_traversal_tree(ref, trav)
    s_ref, s_trav = sah(ref, levels(ref)), sah(trav, lv)
    print(f"  SAH cost {s_ref:.4g} -> {s_trav:.4g}, depth {len(levels(ref)) - 1} -> {len(lv) - 1}, "
          f"build {ctx.last_build_ms():.2f} ms", flush=True)
    assert s_trav < 0.8 * s_ref
    # the bottom-up passes run treelets in whatever order the waves reach them: the tree must
    # not depend on it
    for _ in range(2):
        ctx.build_bvh()
        assert ctx.bvh_traversal().tobytes() == trav.tobytes(), "restructured tree differs between builds"
    ctx.set_treelet_passes(0)
    ctx.build_bvh()
    ref0, _, _ = ctx.bvh()
    assert ref0.tobytes() == ref.tobytes()
    assert ctx.bvh_traversal().tobytes() == ref.tobytes()  # no passes: the reference tree itself


def test_closest_hits_through_the_traversal_tree_match_brute_force(gpu):
    import lgi
    ctx = lgi.Context()
    ctx.generate(1, 3, 128, 64)
    ctx.build_bvh()
    trav = ctx.bvh_traversal()
    tris = ctx.scene()["tris"]  # BVH order
    lo, hi = trav["bb_min"][0], trav["bb_max"][0]
    rng = np.random.default_rng(5)
    hits = 0
    for _ in range(300):
        o = (lo + (hi - lo) * rng.random(3)).astype(np.float32)
        d = rng.normal(size=3).astype(np.float32)
        d /= np.linalg.norm(d)
        i, t = orc.bvh_closest(trav, tris, o, d)
        bi, bt = orc.brute_closest(tris, o, d)
        assert (i, t) == (bi, bt)
        hits += i >= 0
    assert hits > 40


@pytest.mark.parametrize("flags", [0, 2, 10], ids=["inline", "sorted", "sorted-no-ray-sort"])
def test_renders_through_both_trees_are_equal(gpu, flags):
    import torch
    import lgi
    imgs, hits = [], []
    for passes in (0, 2):
        ctx = lgi.Context()
        ctx.set_treelet_passes(passes)
        ctx.generate(1, 1, 256, 128)
        ctx.build_bvh()
        imgs.append(ctx.render(8, 2, flags=flags, hdr=True)["hdr"])
        tri = torch.empty(256 * 128, dtype=torch.int32, device="cuda")
        t = torch.empty(256 * 128, dtype=torch.float32, device="cuda")
        ctx.primary_hits(tri, t)
        torch.cuda.synchronize()
        hits.append((tri.cpu().numpy(), t.cpu().numpy()))
        ctx.close()
    assert np.array_equal(imgs[0], imgs[1])
    assert np.array_equal(hits[0][0], hits[1][0]) and np.array_equal(hits[0][1], hits[1][1])
