"""Shared helpers of the GPU parity tests: device accumulation through the C ABI and the image
comparison against the oracle.

Bar (north_star): 1e-5 relative per value. The kernels evaluate the reference expressions in the
oracle's order with no FP contraction, so values agree bit for bit except where the traversal
order decides the closest hit (an exact distance tie between two triangles, or a box grazed
within rounding on the restructured traversal tree). Such values are counted and must stay
within an allowance the caller states; the default is none.
"""
import numpy as np

REL_TOL = 1e-5


def gpu_accumulate(ctx, cam, spp, bounces, flags=0, sample0=0, sums=None):
    import torch
    n = cam.width * cam.height
    if sums is None:
        sums = [torch.zeros(n * 3, device="cuda"), torch.zeros(n * 3, device="cuda"), torch.zeros(n, device="cuda")]
    ctx.accumulate(sums[0], sums[1], spp, bounces, flags=flags, sample0=sample0, d_sumsq=sums[2], camera=cam)
    torch.cuda.synchronize()
    return sums


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-30)


def assert_close_images(gpu_sum, cpu_sum, label, max_beyond=0):
    """gpu_sum vs cpu_sum (same shape after reshape): at most max_beyond values may differ by more
    than REL_TOL relative (the caller's explicit allowance for traversal-order ties). Returns the
    fraction of bit-identical values."""
    g = np.asarray(gpu_sum).reshape(cpu_sum.shape)
    bad = np.abs(g.astype(np.float64) - cpu_sum) > REL_TOL * np.abs(cpu_sum.astype(np.float64)) + 1e-12
    exact = np.mean(g == cpu_sum)
    worst = float(rel_err(g, cpu_sum).max()) if g.size else 0.0
    print(f"{label}: exact {exact:.6f}, beyond 1e-5: {int(bad.sum())} of {bad.size} "
          f"(allowed {max_beyond}), max rel {worst:.3e}", flush=True)
    assert bad.sum() <= max_beyond, f"{label}: {int(bad.sum())} values beyond {REL_TOL}"
    return exact
