"""Benchmark of the plain path tracer and the denoiser on one MI355X.

Workload: the 62-mesh / ~1.05M-triangle procedural scene, 1728x864, 64 spp, 2 bounces with
sun + 2 point-light NEE, PCG32 streams, ray sorting between bounces. One step = one 64-spp
frame rendered into device accumulators (inputs resident in HBM). Prints one JSON line with
Msamples/s, the traversal kernel's roofline figures (HIP events around every traversal
launch), the CPU oracle timed on the host beside it and, when the trained denoiser checkpoint
is present, the time to 40 dB PSNR of 8 spp + UNet against the plain tracer on the 16-mesh
scene at 1728x864 (error against a 4096-spp plain render).

    python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--spp 64]
"""
import argparse
import json
import os
import pathlib
import platform
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "luminagi-cudaref_amd"))
sys.path.insert(0, str(ROOT))

METRIC = "ms to 40 dB PSNR @1728x864: 8spp+UNet vs plain spp; Msamples/s, 2 bounces"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes (layout-independent minimum): 24 B box + 8 B links per node visit,
# 3 x 12 B vertices per triangle test, 40 B (normals + material) per closest hit. The
# traversal kernel's share is the node and triangle bytes; the hit bytes are read by shading.
BYTES_NODE, BYTES_TRI, BYTES_HIT = 32, 36, 40
# bench --mode -> renderer named in the bench line (all three give the same image bits)
RENDERER_NAMES = {"wavefront": "wavefront renderer, no ray sort", "sorted": "wavefront renderer with the ray sort",
                  "inline": "inline megakernel"}
DENOISER_CKPT = ROOT / "luminagi-cudaref_amd" / "ml" / "weights" / "unet_log1p_c32.pt"


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def trace_bytes_per_sample(c, samples):
    """Algorithmic traversal bytes per sample (32 B per node visit, 36 B per triangle test) and
    the frame's (plus 40 B per closest hit for its shading record)."""
    trace = (BYTES_NODE * c["node_visits"] + BYTES_TRI * c["tri_tests"]) / samples
    return trace, trace + BYTES_HIT * c["closest_hits"] / samples


def cpu_baseline(scene, cam, bounces, row_step, threads, spp=1, trav_nodes=None):
    """Oracle render (reference BVH and traversal) of every row_step-th row at spp samples, timed:
    the CPU baseline. With trav_nodes (the GPU's traversal tree: the same leaves, restructured
    above them) the same rows are rendered again through that tree, untimed, for the node-visit
    and triangle-test counts of the tree the GPU traverses. Returns (Msamples/s, counters on the
    reference tree, counters on the traversal tree, seconds, samples, BVH build seconds, the
    oracle's per-pixel sums of the rendered rows)."""
    import numpy as np
    from oracle import oracle as orc
    osc = orc.Scene(scene["tris"], scene["materials"][:, 0, :], scene["points"], scene["sun"], cam)
    t0 = time.perf_counter()
    out = osc.render(spp, bounces, row_step=row_step, threads=threads)
    dt = time.perf_counter() - t0
    rows = (cam.height + row_step - 1) // row_step
    samples = rows * cam.width * spp
    c_ref = out["counters"]
    rows_sum = out["sum"][::row_step].copy()
    c_trav = c_ref
    if trav_nodes is not None:
        osc.nodes = np.ascontiguousarray(trav_nodes)
        c_trav = osc.render(spp, bounces, row_step=row_step, threads=threads)["counters"]
    t0 = time.perf_counter()
    orc.bvh_build(scene["tris"])
    bvh_s = time.perf_counter() - t0
    return samples / dt / 1e6, c_ref, c_trav, dt, samples, bvh_s, rows_sum


def read_bw_probe(stream):
    """Read bandwidth (GB/s) the chip delivers from L2-, MALL- and HBM-resident buffers
    (lgi_probe_read_bw: 4096 workgroups, each XCD re-reading its eighth of the buffer): 16 MB
    stays in the eight 4 MB L2s, 128 MB (the size of the traversal's working set) is backed by
    the 256 MB MALL, 4 GB comes from HBM. Best of three launches after a warm-up launch, each
    timed by the library with HIP events on its stream."""
    import ctypes as C
    import torch
    import lgi
    lib = lgi._abi.lib()
    sink = torch.zeros(4, dtype=torch.int32, device="cuda")
    out = {}
    for level, mib, reps in (("l2", 16, 2000), ("mall", 128, 240), ("hbm", 4096, 6)):
        buf = torch.ones(mib << 18, dtype=torch.int32, device="cuda")
        nbytes = buf.numel() * 4

        def launch(r):
            ms = C.c_float(0.0)
            st = lib.lgi_probe_read_bw(C.c_void_p(buf.data_ptr()), nbytes, r, 4096, C.c_void_p(sink.data_ptr()),
                                       C.c_void_p(stream), C.byref(ms))
            if st != 0:
                raise lgi.LgiError(st, "lgi_probe_read_bw")
            return ms.value
        torch.cuda.synchronize()
        launch(1)
        best = min(launch(reps) for _ in range(3))
        out[level] = round(nbytes * reps / (best * 1e-3) / 1e9, 1)
        del buf
    return out


def poses_workload(n_poses, seed, bounces):
    """The multi-pose workload: 50 poses on the 15 m orbit of the 16-mesh scene at 1728x864, spp
    log-uniform in [8, 256], 2 bounces. Timed three ways: the persistent queue (one launch),
    the same replayed from its HIP graph, and one sorted-renderer frame per pose. Every pose's
    image is the same in all three (checked here on two poses)."""
    import torch
    import lgi
    from lgi import poses
    ctx = lgi.Context(0, torch.cuda.current_stream().cuda_stream)
    ctx.generate(1, 1, 1728, 864)
    ctx.build_bvh()
    base = ctx.camera()
    n = base.width * base.height
    jobs = [(poses.camera_for(base, pos, tgt), spp, 0) for pos, tgt, spp in poses.orbit_poses(n_poses, seed=seed)]
    samples = sum(j[1] for j in jobs) * n
    bs = torch.zeros(n_poses, n * 3, device="cuda")
    bc = torch.zeros_like(bs)

    def timed(f):
        torch.cuda.synchronize()
        t = time.perf_counter()
        f()
        torch.cuda.synchronize()
        return time.perf_counter() - t
    out = {}
    for key, graph in (("queue_eager", False), ("queue_graph_capture", True), ("queue_graph_replay", True)):
        bs.zero_()
        bc.zero_()
        dt = timed(lambda: ctx.render_batch(jobs, bs, bc, bounces, use_graph=graph))
        out[key] = {"ms": round(dt * 1e3, 1), "msamples_per_s": round(samples / dt / 1e6, 1)}
    # one sorted-renderer frame per pose; poses 0 and n/2 keep their sums for the comparison
    check = (0, n_poses // 2)
    keep_s = torch.zeros(len(check), n * 3, device="cuda")
    keep_c = torch.zeros_like(keep_s)
    tmp_s = torch.zeros(n * 3, device="cuda")
    tmp_c = torch.zeros_like(tmp_s)
    dt = 0.0
    for i, (cam, spp, s0) in enumerate(jobs):
        if i in check:
            d_s, d_c = keep_s[check.index(i)], keep_c[check.index(i)]
        else:
            tmp_s.zero_()
            tmp_c.zero_()
            d_s, d_c = tmp_s, tmp_c
        dt += timed(lambda: ctx.accumulate(d_s, d_c, spp, bounces, flags=lgi.RENDER_SORTED, sample0=s0, camera=cam))
    out["sorted_per_pose"] = {"ms": round(dt * 1e3, 1), "msamples_per_s": round(samples / dt / 1e6, 1)}
    same = all(torch.equal(keep_s[k], bs[i]) for k, i in enumerate(check))
    return {"workload": "multi-pose: 16-mesh procedural scene (%d triangles), %d poses on a 15 m orbit, "
                        "spp log-uniform in [8, 256] (seed %d), 1728x864, %d bounces"
                        % (ctx.counts()[0], n_poses, seed, bounces),
            "samples": samples, "images_equal": same, **out}


UNET_GFLOP = 635.8  # 2*H*W*Cin*Cout*9 summed over the 18 3x3 convolutions at 1728x864
BF16_PEAK_TFLOPS = 2500.0


def unet_workload(ckpt, reps=20):
    """The denoiser alone on one 1728x864 frame (log1p radiance in, radiance out): eager forward
    and the HIP-graph replay, best of `reps` after warm-up, against the dense BF16 MFMA peak."""
    import torch
    from lgi import denoise as D
    from lgi import train as T
    model = T.load_model(ckpt)
    x = torch.rand(864, 1728, 3, device="cuda", generator=torch.Generator("cuda").manual_seed(0)) * 4.0
    for _ in range(3):
        D.denoise(model, x)
    g = D.DenoiseGraph(model, 864, 1728)

    def best(f):
        t_best = float("inf")
        for _ in range(reps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            f()
            torch.cuda.synchronize()
            t_best = min(t_best, (time.perf_counter() - t) * 1e3)
        return t_best
    eager = best(lambda: D.denoise(model, x))
    graph = best(lambda: g(x))
    assert torch.equal(g(x), D.denoise(model, x)), "graph replay differs from the eager forward"
    out = {"workload": "UNet 32/64/128/256 + 512 bottleneck, one 1728x864 frame, BF16 MFMA convolutions",
           "checkpoint": os.path.relpath(ckpt, ROOT), "gflop": UNET_GFLOP,
           "eager_ms": round(eager, 3), "graph_ms": round(graph, 3),
           "tflops_graph": round(UNET_GFLOP / graph, 1),
           "frac_of_bf16_peak": round(UNET_GFLOP / graph / BF16_PEAK_TFLOPS, 4)}
    # the same frame with the 13 FP8-capable convolutions on E4M3 operands (activation scales
    # calibrated on a training pose of the 16-mesh scene)
    with torch.no_grad():
        model.enable_fp8(T.fp8_calibration_frame(T.make_scene(1, 1, 1728, 864)), radiance=True)
    for _ in range(3):
        D.denoise(model, x)
    g8 = D.DenoiseGraph(model, 864, 1728)
    eager8 = best(lambda: D.denoise(model, x))
    graph8 = best(lambda: g8(x))
    out["fp8"] = {"eager_ms": round(eager8, 3), "graph_ms": round(graph8, 3), "tflops_graph": round(UNET_GFLOP / graph8, 1),
                  "frac_of_bf16_peak": round(UNET_GFLOP / graph8 / BF16_PEAK_TFLOPS, 4)}
    return out


def train_workload(steps, warmup, batch=16, crop=256, n_frames=8):
    """Training steps of the denoiser at batch 16 x 256x256 crops (Adam, lr 1e-4, L1 on log1p
    radiance, dynamic loss scaling, activation recompute in the backward), on synthetic pairs of
    1728x864 frames resident on the GPU. A step is crop sampling, forward, backward, optimizer
    update. FLOP per step: the forward's 3x3 convolutions at the crop size, times 3 for forward,
    data gradient and weight gradient (the recomputed forward counted apart)."""
    import torch
    from lgi import train as T
    from lgi.unet import UNet
    g = torch.Generator("cuda").manual_seed(0)
    clean = torch.log1p(torch.rand(n_frames, 864, 1728, 3, device="cuda", generator=g) * 4.0)
    noisy = clean + 0.1 * torch.randn(n_frames, 864, 1728, 3, device="cuda", generator=g)
    torch.manual_seed(0)
    model = UNet(seed=0, recompute=True).cuda()
    params = list(model.parameters())
    opt = torch.optim.Adam(params, lr=1e-4)
    scaler = T.LossScaler()
    gen = torch.Generator().manual_seed(0)

    def step():
        xb, yb = T.random_crops(noisy, clean, batch, crop, gen)
        loss = (model(xb) - yb).abs().mean()
        (loss * scaler.scale).backward()
        scaler.step(params, opt)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t) * 1e3 / steps
    fwd = UNET_GFLOP * batch * crop * crop / (1728 * 864)
    return {"workload": f"denoiser training step, batch {batch} x {crop}x{crop} crops of synthetic 1728x864 "
                        "log1p pairs, Adam lr 1e-4, loss scaling, recompute in the backward",
            "steps": steps, "warmup": warmup, "ms_per_step": round(ms, 3),
            "gflop_forward": round(fwd, 1), "gflop_per_step": round(3 * fwd, 1),
            "gflop_per_step_with_recompute": round(4 * fwd, 1),
            "tflops": round(3 * fwd / ms, 1), "frac_of_bf16_peak": round(3 * fwd / ms / BF16_PEAK_TFLOPS, 4)}


def _summary(r):
    """The figures of one time-to-40-dB evaluation that the bench line keeps."""
    keep = ("plain_spp_40db", "plain_ms_40db", "unet_reaches_40db", "unet_ms", "unet_variant_40db",
            "speedup_40db", "plain_spp_same_psnr", "plain_ms_same_psnr", "speedup_same_psnr")
    out = {k: (round(r[k], 3) if isinstance(r[k], float) else r[k]) for k in keep if k in r}
    out["noisy8_psnr"] = round(r["noisy8"]["psnr"], 3)
    out["methods"] = {k: {kk: (round(vv, 3) if isinstance(vv, float) else vv) for kk, vv in m.items()}
                      for k, m in r["methods"].items()}
    out["renderer"] = r["renderer"]
    return out


def denoise_section(ckpt, bounces):
    """Time to 40 dB at 1728x864 with the faster renderer for each view (same image bits), error
    against a 4096-spp plain render. The headline view is the held-out orbit pose of the 16-mesh
    training scene; two more views check that the denoiser is not only a model of its training
    poses: a close-up from inside the scene, far from the 15 m orbit the training poses stand on,
    and the held-out pose of a second 16-mesh scene (seed 2) the network never saw."""
    from lgi import train as T
    ctx = T.make_scene(1, 1, 1728, 864)
    model = T.load_model(ckpt)
    r = T.evaluate(ctx, model, T.held_out_camera(ctx), ref_spp=4096, bounces=bounces, log=lambda *a: None)
    out = _summary(r)
    out["unet"] = {k: round(v, 3) for k, v in r["unet"].items()}
    if "unet_fp8" in r:
        out["unet_fp8"] = {k: round(v, 3) for k, v in r["unet_fp8"].items()}
    if "renderer_ms_8spp" in r:
        out["renderer_ms_8spp"] = {k: round(v, 3) for k, v in r["renderer_ms_8spp"].items()}
    if "unet_fallback" in r:
        out["unet_fallback"] = {k: round(v, 3) for k, v in r["unet_fallback"].items()}
    out["plain_curve"] = [(p["spp"], round(p["ms"], 2), round(p["psnr"], 3)) for p in r["plain"]]
    out["workload"] = "16-mesh procedural scene (seed 1), 1728x864, 2 bounces, held-out orbit pose; error vs 4096 spp"
    out["checkpoint"] = os.path.relpath(ckpt, ROOT)
    views = []
    r2 = T.evaluate(ctx, model, T.off_orbit_camera(ctx), ref_spp=4096, bounces=bounces, fallback=False,
                    log=lambda *a: None)
    views.append(dict(_summary(r2), view="off-orbit close-up inside the training scene (seed 1)"))
    ctx2 = T.make_scene(1, 2, 1728, 864)
    r3 = T.evaluate(ctx2, model, T.held_out_camera(ctx2), ref_spp=4096, bounces=bounces, fallback=False,
                    log=lambda *a: None)
    views.append(dict(_summary(r3), view="held-out orbit pose of an unseen 16-mesh scene (seed 2)"))
    out["other_views"] = views
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--bounces", type=int, default=2)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-row-step", type=int, default=2)
    ap.add_argument("--cpu-spp", type=int, default=64)
    ap.add_argument("--mode", choices=["sorted", "wavefront", "inline"], default="sorted",
                    help="wavefront renderer with the ray sort between bounces (default: the bench "
                         "configuration sorts its secondary rays), without it, or the inline megakernel; all "
                         "three give the same image bits")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-denoise", action="store_true", help="skip the time-to-40-dB measurement")
    ap.add_argument("--no-probe", action="store_true", help="skip the L2/MALL/HBM read-bandwidth probe")
    ap.add_argument("--ckpt", default=str(DENOISER_CKPT), help="denoiser checkpoint for the 40 dB measurement")
    ap.add_argument("--workload", choices=["frame", "poses", "unet", "train"], default="frame",
                    help="frame: the 64-spp frame of the 1.05M-triangle scene (default); poses: only time the "
                         "50-pose multi-pose workload; unet: only time the denoiser on one 1728x864 frame; "
                         "train: only time denoiser training steps (batch 16 x 256x256)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe_mib:
        import torch
        n = args.copy_probe_mib * (1 << 20) // 4
        src = torch.ones(n, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_each_way": n * 4, "launches": 3}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the tracer would shard by pose or image tile with one accumulator reduce)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "poses":
        print(json.dumps(poses_workload(50, 3, args.bounces)), flush=True)
        return
    if args.workload == "unet":
        print(json.dumps(unet_workload(args.ckpt)), flush=True)
        return
    if args.workload == "train":
        print(json.dumps(train_workload(args.steps, args.warmup)), flush=True)
        return

    import numpy as np
    import torch

    import lgi

    dev = torch.device("cuda:0")
    ctx = lgi.Context(0, torch.cuda.current_stream(dev).cuda_stream)
    ctx.generate(2, args.seed, 1728, 864)
    scene = ctx.scene()
    cam = scene["camera"]
    W, H = cam.width, cam.height
    n_px = W * H
    ctx.build_bvh()
    torch.cuda.synchronize()
    build_ms = [ctx.last_build_ms()]
    tree_ms = [ctx.last_treelet_ms()]
    for _ in range(3):
        ctx.build_bvh()
        build_ms.append(ctx.last_build_ms())
        tree_ms.append(ctx.last_treelet_ms())
    acc = torch.zeros(n_px * 3, device=dev)
    comp = torch.zeros(n_px * 3, device=dev)
    sumsq = torch.zeros(n_px, device=dev)

    flags = {"wavefront": lgi.RENDER_SORTED | lgi.RENDER_NO_RAY_SORT, "sorted": lgi.RENDER_SORTED,
             "inline": 0}[args.mode]
    sample0 = 0
    for _ in range(args.warmup):
        ctx.accumulate(acc, comp, args.spp, args.bounces, flags=flags, sample0=sample0, d_sumsq=sumsq)
        sample0 += args.spp
    torch.cuda.synchronize()
    kernel_ms, trace_ms, trace_launches = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.accumulate(acc, comp, args.spp, args.bounces, flags=flags, sample0=sample0, d_sumsq=sumsq)
        # HIP events on the context stream: whole render call, and every traversal launch in it
        kernel_ms.append(ctx.last_render_ms())
        tm, tl = ctx.last_trace_ms()
        trace_ms.append(tm)
        trace_launches.append(tl)
        sample0 += args.spp
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    samples_per_step = n_px * args.spp
    msps = samples_per_step / (wall / args.steps) / 1e6
    avg_kernel_ms = float(np.mean(kernel_ms))
    avg_trace_ms = float(np.mean(trace_ms))

    result = {
        "metric": METRIC, "value": round(msps, 3), "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "plain path tracer (%s), 62-mesh procedural scene (%d triangles), %dx%d, %d spp "
                               "per step, %d bounces, sun + 2 point lights, PCG32" % (
                                   RENDERER_NAMES[args.mode], len(scene["tris"]), W, H, args.spp, args.bounces),
                   "triangles": int(len(scene["tris"])), "width": W, "height": H, "spp_per_step": args.spp,
                   "bounces": args.bounces, "scene_seed": args.seed, "renderer": args.mode},
        "render_ms": round(avg_kernel_ms, 3),
        "trace_kernel_ms_per_frame": round(avg_trace_ms, 3),
        "trace_launches_per_frame": int(np.median(trace_launches)),
        "lbvh_build_ms": round(float(np.median(np.subtract(build_ms[1:], tree_ms[1:]))), 3),
        "treelet_restructure_ms": round(float(np.median(tree_ms[1:])), 3),
        "triangles": int(len(scene["tris"])),
    }
    if args.mode == "sorted":
        # The same frame without the ray sort (same image bits): on this frame the sort saves less
        # traversal time than its compaction and radix passes take. Reported beside the
        # configured workload, not as the bench value.
        fl = flags | lgi.RENDER_NO_RAY_SORT
        ctx.accumulate(acc, comp, args.spp, args.bounces, flags=fl, sample0=sample0, d_sumsq=sumsq)
        sample0 += args.spp
        torch.cuda.synchronize()
        n_alt = max(1, min(args.steps, 5))
        alt_trace = []
        t0 = time.perf_counter()
        for _ in range(n_alt):
            ctx.accumulate(acc, comp, args.spp, args.bounces, flags=fl, sample0=sample0, d_sumsq=sumsq)
            alt_trace.append(ctx.last_trace_ms()[0])
            sample0 += args.spp
        torch.cuda.synchronize()
        alt_s = (time.perf_counter() - t0) / n_alt
        result["without_ray_sort"] = {"renderer": RENDERER_NAMES["wavefront"], "steps": n_alt,
                                      "ms_per_step": round(alt_s * 1e3, 3),
                                      "msamples_per_s": round(samples_per_step / alt_s / 1e6, 3),
                                      "trace_kernel_ms_per_frame": round(float(np.mean(alt_trace)), 3)}
    if not args.no_cpu:
        threads = len(os.sched_getaffinity(0))
        cpu_msps, c_ref, counters, cpu_s, cpu_samples, bvh_s, cpu_rows = cpu_baseline(
            scene, cam, args.bounces, args.cpu_row_step, threads, args.cpu_spp, trav_nodes=ctx.bvh_traversal())
        trace_bps, bps = trace_bytes_per_sample(counters, cpu_samples)
        ref_trace_bps, _ = trace_bytes_per_sample(c_ref, cpu_samples)
        # Dominant kernel = BVH traversal. Its algorithmic bytes per frame are the node visits and
        # triangle tests on the tree it traverses, over its summed launch time per frame (HIP
        # events, live). About 91% of them are served by LDS, L1 and L2 without leaving the XCD,
        # so they are rated against the L2 read bandwidth the probe measures, not against HBM.
        # The HBM roof is rated with the bytes that do cross it: the calibrated FETCH_SIZE and
        # WRITE_SIZE counters of the same launches (rocprofv3, committed profile of this workload).
        algo_bytes = trace_bps * samples_per_step
        algo_gbs = algo_bytes / (avg_trace_ms * 1e-3) / 1e9
        traffic, counters_src = None, None
        rel = "profiles/trace_%s_config2.json" % args.mode
        prof = ROOT / rel
        if prof.exists():
            p = json.loads(prof.read_text())
            if (p.get("spp"), p.get("bounces"), p.get("scene_seed")) == (args.spp, args.bounces, args.seed):
                traffic = int(p["traffic_bytes_per_frame"])
                counters_src = {"file": rel, "profiled_trace_ms_per_frame": round(p["trace_ns_per_frame"] / 1e6, 3),
                                "l2_hit_rate": round(p["l2_hit_rate"], 4), "fetch_factor": round(p["fetch_factor"], 4),
                                "write_factor": round(p["write_factor"], 4),
                                "factors_from": "a 1 GiB device copy: FETCH_SIZE reads half the bytes moved, "
                                                "WRITE_SIZE all of them"}
                if "pipes" in p:
                    # what does bound the traversal: VALU issue and the vector-memory data path
                    counters_src["pipes_all_launches"] = p["pipes"]["per_launch_kind"]["all"]
        hbm_gbs = traffic / (avg_trace_ms * 1e-3) / 1e9 if traffic else None
        result["roofline"] = {"bound": "hbm", "achieved": round(hbm_gbs, 1) if hbm_gbs else None,
                              "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(hbm_gbs / HBM_PEAK_GBS, 4) if hbm_gbs else None, "traffic": traffic,
                              "achieved_from": "HBM-side bytes per frame (calibrated FETCH_SIZE + WRITE_SIZE of the "
                                               "traversal launches) over their live summed launch time",
                              "kernel": "k_wf_trace (all traversal launches of one frame)" if args.mode != "inline"
                              else "k_render",
                              "algorithmic": {"bytes_per_frame": int(algo_bytes), "achieved": round(algo_gbs, 1),
                                              "unit": "GB/s", "trace_bytes_per_sample": round(trace_bps, 1),
                                              "reference_tree_trace_bytes_per_sample": round(ref_trace_bps, 1),
                                              "frame_bytes_per_sample": round(bps, 1),
                                              "traffic_over_algorithmic": round(traffic / algo_bytes, 4)
                                              if traffic else None},
                              "counters": counters_src}
        if not args.no_probe:
            peaks = read_bw_probe(torch.cuda.current_stream(dev).cuda_stream)
            result["roofline"]["measured_read_peaks"] = dict(peaks, unit="GB/s",
                                                             source="lgi_probe_read_bw in this run")
            result["roofline"]["algorithmic"]["peak_l2_measured"] = peaks["l2"]
            result["roofline"]["algorithmic"]["frac_of_l2_measured"] = round(algo_gbs / peaks["l2"], 4)
        # parity of the bench frame itself: the same 64 samples (from sample 0) on the rows the
        # CPU baseline rendered, against the oracle's sums
        if args.cpu_spp == args.spp:
            chk = [torch.zeros(n_px * 3, device=dev), torch.zeros(n_px * 3, device=dev)]
            ctx.accumulate(chk[0], chk[1], args.spp, args.bounces, flags=flags)
            torch.cuda.synchronize()
            g = chk[0].cpu().numpy().reshape(H, W, 3)[::args.cpu_row_step].astype(np.float64)
            c = cpu_rows.astype(np.float64)
            beyond = np.abs(g - c) > 1e-5 * np.abs(c) + 1e-12
            result["parity_vs_oracle"] = {"rows": "every %dth row" % args.cpu_row_step, "values": int(c.size),
                                          "bit_exact_frac": round(float(np.mean(g == c)), 6),
                                          "beyond_1e-5_rel": int(beyond.sum())}
            del chk
        result["cpu_baseline"] = {"value": round(cpu_msps, 4), "unit": "Msamples/s", "cores": threads,
                                  "kind": "port",
                                  "sample": "CPU oracle, same scene and PCG32 streams, %d spp on every %dth row "
                                            "(%d samples, %.1f s), %s" % (args.cpu_spp, args.cpu_row_step, cpu_samples,
                                                                          cpu_s, cpu_model()),
                                  "bvh_build_ms_1thread": round(bvh_s * 1e3, 1),
                                  "gpu_lbvh_build_ms": result["lbvh_build_ms"]}
        from oracle import oracle as orc
        if orc.ref_bvh_available():
            # the reference's own BuildBVH (src/bvh.cpp, compiled from its sources into
            # oracle/_ref) on the same triangles, one thread; its node array against the GPU's
            import tempfile
            with tempfile.TemporaryDirectory() as td:
                ctx.save_bin(pathlib.Path(td) / "scene.lgis")
                ref_nodes, _, ref_ms = orc.ref_bvh(pathlib.Path(td) / "scene.lgis", pathlib.Path(td) / "ref.bin",
                                                   timed=True)
            result["cpu_baseline"]["bvh_build_ms_reference"] = round(ref_ms, 1)
            result["cpu_baseline"]["reference_bvh_nodes_equal_gpu"] = bool(ref_nodes.tobytes() ==
                                                                           ctx.bvh()[0].tobytes())
        result["oracle_counters_per_sample"] = {k: round(v / cpu_samples, 2) for k, v in counters.items()}
        result["oracle_counters_per_sample_reference_tree"] = {k: round(v / cpu_samples, 2) for k, v in c_ref.items()}
        # rays per sample (closest-hit + shadow) from the oracle's deterministic counts
        rays = (counters["rays"] + counters["shadow_rays"]) / cpu_samples
        result["mrays_per_s"] = round(msps * rays, 1)
    if not args.no_denoise and pathlib.Path(args.ckpt).exists():
        result["ms_to_40db"] = denoise_section(args.ckpt, args.bounces)
        # the denoiser's MFMA roofline: its 635.8 GFLOP of 3x3 convolutions per 1728x864 frame over
        # the graph-replayed forward inside that measurement, against the dense BF16 MFMA peak
        dn_ms = result["ms_to_40db"]["unet"]["denoise_ms"]
        result["denoiser_roofline"] = {"bound": "mfma", "achieved": round(UNET_GFLOP / dn_ms, 1),
                                       "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                                       "frac": round(UNET_GFLOP / dn_ms / BF16_PEAK_TFLOPS, 4),
                                       "gflop_per_frame": UNET_GFLOP, "ms_per_frame": dn_ms,
                                       "kernel": "UNet forward (18 BF16 MFMA 3x3 convolutions + pools, upsamples, "
                                                 "tail), one HIP graph"}
        if "unet_fp8" in result["ms_to_40db"]:
            f8 = result["ms_to_40db"]["unet_fp8"]["denoise_ms"]
            result["denoiser_roofline"]["fp8"] = {
                "ms_per_frame": f8, "achieved": round(UNET_GFLOP / f8, 1), "unit": "TFLOP/s",
                "frac_of_bf16_peak": round(UNET_GFLOP / f8 / BF16_PEAK_TFLOPS, 4),
                "layers": "13 of 18 convolutions on E4M3 operands (v_mfma_scale_f32_32x32x64_f8f6f4, dense "
                          "FP8 peak 5.0 PFLOP/s); the 5 with 32-channel inputs stay BF16"}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
