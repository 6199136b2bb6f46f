"""The allocation-free render entry point (lgi_render_accumulate_noalloc) against the render it
varies (lgi_render_accumulate): the same sums bit for bit on two calls in a row, the second
with a smaller image, fewer samples and another RNG policy; no device memory is allocated
by either call; and the call captured in a HIP graph replays to the eager sums."""
import ctypes as C

import pytest

pytestmark = pytest.mark.gpu


def _noalloc(ctx, cam, spp, bounces, flags, sample0, bufs):
    """Direct C-ABI call, no Python wrapper in between."""
    import lgi
    st = lgi.lib().lgi_render_accumulate_noalloc(ctx._h, C.byref(cam), spp, bounces, flags, sample0,
                                                 bufs[0].data_ptr(), bufs[1].data_ptr(), bufs[2].data_ptr())
    assert st == 0, lgi.lib().lgi_last_error(ctx._h)


def test_noalloc_render_equals_accumulate_twice(gpu):
    import torch
    import lgi
    from parity import gpu_accumulate
    ctx = lgi.Context()
    ctx.generate(1, 2, 200, 120)
    ctx.build_bvh()
    cam = ctx.camera()
    small = lgi.Camera.from_buffer_copy(cam)
    small.width, small.height = 61, 37                      # ragged tiles
    print("noalloc: 200x120 12 spp pcg32, then 61x37 5 spp xorshift-ref", flush=True)
    ref1 = gpu_accumulate(ctx, cam, 12, 2, flags=0, sample0=3)
    ref2 = gpu_accumulate(ctx, small, 5, 1, flags=lgi.RNG_XORSHIFT_REF, sample0=0)
    out1 = [torch.zeros_like(t) for t in ref1]
    out2 = [torch.zeros_like(t) for t in ref2]
    warm = [torch.zeros_like(t) for t in ref1]
    # The HIP runtime sizes a kernel's private-segment (register spill) memory at its first launch;
    # that is the runtime's, not the entry point's, so the kernel runs once before the snapshot.
    _noalloc(ctx, cam, 1, 2, 0, 0, warm)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    _noalloc(ctx, cam, 12, 2, 0, 3, out1)
    _noalloc(ctx, small, 5, 1, lgi.RNG_XORSHIFT_REF, 0, out2)
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    for a, b in zip(ref1, out1):
        assert torch.equal(a, b)
    for a, b in zip(ref2, out2):
        assert torch.equal(a, b)
    # a hipMalloc'd spill area for this image would be >= 9 MB
    assert free0 - free1 < (2 << 20), f"{(free0 - free1) / 2 ** 20:.1f} MiB of device memory went missing"


def test_noalloc_rejects_sorted(gpu):
    import torch
    import lgi
    ctx = lgi.Context()
    ctx.generate(1, 2, 32, 16)
    ctx.build_bvh()
    n = 32 * 16
    bufs = [torch.zeros(n * 3, device="cuda"), torch.zeros(n * 3, device="cuda")]
    with pytest.raises(lgi.LgiError) as e:
        ctx.accumulate_noalloc(bufs[0], bufs[1], 2, 1, flags=lgi.RENDER_SORTED)
    assert e.value.status == 8


def test_noalloc_render_in_graph_replays(gpu):
    import torch
    import lgi
    from parity import gpu_accumulate
    ctx = lgi.Context()
    ctx.generate(1, 3, 96, 64)
    ctx.build_bvh()
    cam = ctx.camera()
    eager = gpu_accumulate(ctx, cam, 6, 2, sample0=1)
    bufs = [torch.zeros_like(t) for t in eager]
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        ctx.set_stream(torch.cuda.current_stream().cuda_stream)
        for t in bufs:
            t.zero_()
        ctx.accumulate_noalloc(bufs[0], bufs[1], 6, 2, sample0=1, d_sumsq=bufs[2], camera=cam)
    ctx.set_stream(None)
    for k in range(3):
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip(eager, bufs):
            assert torch.equal(a, b), f"replay {k}"
        _ = gpu_accumulate(ctx, cam, 2, 1)           # eager work between replays
