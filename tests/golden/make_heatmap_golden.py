"""Writes heatmap_golden.npz: crops of the reference's own diff-tool output.

The reference repository ships docs/engine.png, docs/reference.png and docs/heatmap.png, the
heatmap its tools/diff.py made from the other two (README.md:11, README.md:218). This script
cuts three 48x64 crops out of each (the region of largest difference, a region of middling
difference, and a fixed one) so that the heatmap oracle and kernel can be checked against the
reference's real output without the 1728x864 images. Run in the build container, where
/root/reference is mounted:

    python tests/golden/make_heatmap_golden.py
"""
import pathlib

import numpy as np
from PIL import Image

DOCS = pathlib.Path("/root/reference/docs")
OUT = pathlib.Path(__file__).resolve().parent / "heatmap_golden.npz"
CH, CW = 48, 64


def main():
    eng = np.asarray(Image.open(DOCS / "engine.png").convert("RGB"), dtype=np.uint8)
    ref = np.asarray(Image.open(DOCS / "reference.png").convert("RGB"), dtype=np.uint8)
    heat = np.asarray(Image.open(DOCS / "heatmap.png").convert("RGB"), dtype=np.uint8)
    assert eng.shape == ref.shape == heat.shape
    d = np.abs(eng.astype(np.int16) - ref.astype(np.int16)).sum(axis=2)
    H, W = d.shape
    # block sums of the difference on the crop grid
    gh, gw = H // CH, W // CW
    blocks = d[:gh * CH, :gw * CW].reshape(gh, CH, gw, CW).sum(axis=(1, 3))
    order = np.argsort(blocks, axis=None)
    picks = [order[-1], order[len(order) // 2], (gh // 3) * gw + gw // 4]
    out = {}
    for k, b in enumerate(picks):
        by, bx = divmod(int(b), gw)
        ys, xs = slice(by * CH, (by + 1) * CH), slice(bx * CW, (bx + 1) * CW)
        out[f"c{k}_engine"], out[f"c{k}_reference"], out[f"c{k}_heatmap"] = eng[ys, xs], ref[ys, xs], heat[ys, xs]
        out[f"c{k}_origin"] = np.array([by * CH, bx * CW], np.int32)
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
