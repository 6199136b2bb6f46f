"""Python front end of liblgi: scene I/O, GPU LBVH and the path tracer.

Mirrors the reference's host interface (LoadSceneJSON, BuildBVH, RenderSceneCUDA,
RenderSceneCUDASorted; src/main.cpp:19-27, include/scene.h:58, include/bvh.h:26) on top of the
C ABI in include/lgi.h. Host-side calls (scene generation, JSON, batch files, PNG) work
without a GPU; every compute call runs the HIP kernels and raises LgiError on failure.
"""
import ctypes as C

import numpy as np

from . import _abi
from ._abi import (BatchJob, Camera, DirLight, LgiError, Material, NODE_DTYPE, POINT_DTYPE, TRI_DTYPE, lib, ptr)


def _check(ctx, status):
    if status != _abi.LGI_OK:
        msg = lib().lgi_last_error(ctx).decode() if ctx else ""
        raise LgiError(status, msg)


def camera_look_at(cam, pos, target):
    """overwrite_camera (src/main.cpp:29-59) on a Camera struct, in place."""
    p = (C.c_float * 3)(*pos)
    t = (C.c_float * 3)(*target)
    _check(None, lib().lgi_camera_look_at(C.byref(cam), p, t))
    return cam


def parse_batch(path):
    """read_batch_file (src/main.cpp:61-91): list of (out_png, spp, bounces, cam_pos, cam_tgt)."""
    n = C.c_int32(0)
    _check(None, lib().lgi_batch_parse(str(path).encode(), None, 0, C.byref(n)))
    jobs = (BatchJob * max(1, n.value))()
    _check(None, lib().lgi_batch_parse(str(path).encode(), jobs, n.value, C.byref(n)))
    return [(j.out_png.decode(), j.spp, j.bounces, tuple(j.cam_pos), tuple(j.cam_tgt)) for j in jobs[:n.value]]


def read_png(path):
    """PNG -> HxWx4 uint8 RGBA (the texture loader's decoder)."""
    w, h = C.c_int32(), C.c_int32()
    _check(None, lib().lgi_read_png(str(path).encode(), C.byref(w), C.byref(h), None))
    out = np.zeros((h.value, w.value, 4), np.uint8)
    _check(None, lib().lgi_read_png(str(path).encode(), C.byref(w), C.byref(h), ptr(out)))
    return out


def write_png(path, rgb):
    rgb = np.ascontiguousarray(rgb, dtype=np.uint8)
    h, w = rgb.shape[:2]
    _check(None, lib().lgi_write_png(str(path).encode(), ptr(rgb), w, h))


def wavefront_samples_per_pass(n_slots, n_points, spp):
    """Samples per pass of the sorted renderer for n_slots queue slots per sample (64 per 8x8
    tile) and n_points point lights (host arithmetic only)."""
    s = C.c_int32()
    _check(None, lib().lgi_wavefront_samples_per_pass(n_slots, n_points, spp, C.byref(s)))
    return s.value


class Context:
    """One liblgi context: a scene, its device copy and BVH, and the render entry points."""

    def __init__(self, device=0, stream=None):
        h = C.c_void_p()
        _check(None, lib().lgi_create(C.byref(h), device, stream))
        self._h = h
        self.stream = stream  # hipStream_t handle the context orders its work on (None: null stream)

    def close(self):
        h, self._h = getattr(self, "_h", None), None
        l = _abi._lib
        if h and l is not None:
            l.lgi_destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown: modules may already be torn down
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _c(self, status):
        _check(self._h, status)

    def set_stream(self, stream):
        self._c(lib().lgi_set_stream(self._h, stream))
        self.stream = stream

    # ------------------------------------------------------------------ scene
    def load_json(self, path, verbose=False):
        self._c(lib().lgi_scene_load_json(self._h, str(path).encode(), int(verbose)))

    def save_json(self, path):
        self._c(lib().lgi_scene_save_json(self._h, str(path).encode()))

    def load_bin(self, path, verbose=False):
        """Binary scene file written by save_bin (bulk records, no text parsing)."""
        self._c(lib().lgi_scene_load_bin(self._h, str(path).encode(), int(verbose)))

    def save_bin(self, path):
        self._c(lib().lgi_scene_save_bin(self._h, str(path).encode()))

    def generate(self, kind, seed, width, height):
        self._c(lib().lgi_scene_generate(self._h, kind, seed, width, height))

    def set_scene(self, camera, sun, points, materials, tris, dirs=()):
        pts = np.ascontiguousarray(points, dtype=POINT_DTYPE)
        tr = np.ascontiguousarray(tris, dtype=TRI_DTYPE)
        mats = (Material * max(1, len(materials)))()
        for i, (alb, emi) in enumerate(materials):
            mats[i].albedo[:] = list(alb)
            mats[i].emissive[:] = list(emi)
        dl = (DirLight * max(1, len(dirs)))(*dirs)
        self._c(lib().lgi_scene_set(self._h, C.byref(camera), C.byref(sun), dl, len(dirs), ptr(pts), len(pts),
                                    mats, len(materials), ptr(tr), len(tr)))

    def counts(self):
        nt, nm, np_, nd = C.c_int64(), C.c_int32(), C.c_int32(), C.c_int32()
        self._c(lib().lgi_scene_counts(self._h, C.byref(nt), C.byref(nm), C.byref(np_), C.byref(nd)))
        return nt.value, nm.value, np_.value, nd.value

    def scene(self):
        """Returns dict(camera, sun, points, materials (n,2,3), tris) — tris in BVH order after a build."""
        nt, nm, npt, _ = self.counts()
        cam, sun = Camera(), DirLight()
        pts = np.zeros(npt, POINT_DTYPE)
        mats = (Material * max(1, nm))()
        tris = np.zeros(nt, TRI_DTYPE)
        self._c(lib().lgi_scene_get(self._h, C.byref(cam), C.byref(sun), ptr(pts) if npt else None, mats,
                                    ptr(tris) if nt else None))
        m = np.array([[list(mats[i].albedo), list(mats[i].emissive)] for i in range(nm)], np.float32).reshape(nm, 2, 3)
        return {"camera": cam, "sun": sun, "points": pts, "materials": m, "tris": tris}

    def camera(self):
        cam = Camera()
        self._c(lib().lgi_scene_get(self._h, C.byref(cam), None, None, None, None))
        return cam

    def set_asset_root(self, path):
        """Directory the materials' texture paths are relative to (None: the scene's grandparent)."""
        self._c(lib().lgi_set_asset_root(self._h, None if path is None else str(path).encode()))

    def set_camera(self, cam):
        self._c(lib().lgi_scene_set_camera(self._h, C.byref(cam)))

    # ------------------------------------------------------------------ BVH
    def build_bvh(self):
        self._c(lib().lgi_bvh_build(self._h))

    def bvh(self):
        """(nodes[NODE_DTYPE], perm int32, sorted keys uint32) of the last GPU build."""
        n = C.c_int64()
        self._c(lib().lgi_bvh_node_count(self._h, C.byref(n)))
        nt = self.counts()[0]
        nodes = np.zeros(n.value, NODE_DTYPE)
        perm = np.zeros(nt, np.int32)
        keys = np.zeros(nt, np.uint32)
        self._c(lib().lgi_bvh_get(self._h, ptr(nodes), ptr(perm), ptr(keys)))
        return nodes, perm, keys

    def set_treelet_passes(self, passes):
        """Restructuring passes of the tree the tracers traverse (0 = the reference tree itself);
        takes effect at the next build_bvh()."""
        self._c(lib().lgi_bvh_set_treelet_passes(self._h, passes))

    def bvh_traversal(self):
        """The traversal tree: nodes[NODE_DTYPE] in the slots of bvh()'s array (same leaves)."""
        n = C.c_int64()
        self._c(lib().lgi_bvh_node_count(self._h, C.byref(n)))
        nodes = np.zeros(n.value, NODE_DTYPE)
        self._c(lib().lgi_bvh_get_traversal(self._h, ptr(nodes)))
        return nodes

    def last_build_ms(self):
        ms = C.c_float()
        self._c(lib().lgi_bvh_last_build_ms(self._h, C.byref(ms)))
        return ms.value

    def last_treelet_ms(self):
        """Device time of the last build's traversal-tree restructuring (0 when it did not run)."""
        ms = C.c_float()
        self._c(lib().lgi_bvh_last_treelet_ms(self._h, C.byref(ms)))
        return ms.value

    # ------------------------------------------------------------------ render
    def render(self, spp, bounces, flags=0, camera=None, hdr=False, gbuffer=False):
        """RenderSceneCUDA: returns dict(rgb u8 HxWx3, hdr f32 HxWx3?, albedo/normal/worldpos?)."""
        cam = camera if camera is not None else self.camera()
        W, H = cam.width, cam.height
        rgb = np.zeros((H, W, 3), np.uint8)
        out = {"rgb": rgb}
        hd = np.zeros((H, W, 3), np.float32) if hdr else None
        gb = [np.zeros((H, W, 3), np.uint8) for _ in range(3)] if gbuffer else [None, None, None]
        self._c(lib().lgi_render(self._h, C.byref(cam), spp, bounces, flags, ptr(rgb), ptr(hd), ptr(gb[0]),
                                 ptr(gb[1]), ptr(gb[2])))
        if hdr:
            out["hdr"] = hd
        if gbuffer:
            out["albedo"], out["normal"], out["worldpos"] = gb
        return out

    def accumulate(self, d_sum, d_comp, spp, bounces, flags=0, sample0=0, d_sumsq=None, camera=None):
        """Add spp samples into caller-owned device buffers (torch tensors or raw pointers)."""
        cam = camera if camera is not None else self.camera()
        self._c(lib().lgi_render_accumulate(self._h, C.byref(cam), spp, bounces, flags, sample0, _dp(d_sum),
                                            _dp(d_comp), _dp(d_sumsq)))

    def accumulate_noalloc(self, d_sum, d_comp, spp, bounces, flags=0, sample0=0, d_sumsq=None, camera=None):
        """accumulate() without any device allocation (lgi_render_accumulate_noalloc): same sums for
        the non-sorted renderers; for callers that cannot allocate (graph capture, serving loops)."""
        cam = camera if camera is not None else self.camera()
        self._c(lib().lgi_render_accumulate_noalloc(self._h, C.byref(cam), spp, bounces, flags, sample0, _dp(d_sum),
                                                    _dp(d_comp), _dp(d_sumsq)))

    def render_batch(self, jobs, d_sum, d_comp, bounces, d_sumsq=None, flags=0, use_graph=False):
        """jobs: list of (camera, spp, sample0); buffers [n][W*H*3] (sum, comp), [n][W*H] (sumsq)."""
        arr = (_abi.PoseJob * len(jobs))()
        for i, (cam, spp, s0) in enumerate(jobs):
            arr[i].cam = cam
            arr[i].spp = spp
            arr[i].sample0 = s0
        self._c(lib().lgi_render_batch(self._h, arr, len(jobs), bounces, flags, _dp(d_sum), _dp(d_comp),
                                       _dp(d_sumsq), int(use_graph)))

    def render_tiles(self, tiles, spp, d_sum, d_comp, bounces, sample0=0, d_sumsq=None, flags=0, camera=None):
        """Adds spp samples to the listed 8x8 tiles (row-major tile indices)."""
        cam = camera if camera is not None else self.camera()
        t = np.ascontiguousarray(tiles, np.int32)
        self._c(lib().lgi_render_tiles(self._h, C.byref(cam), ptr(t) if len(t) else None, len(t), spp, sample0,
                                       bounces, flags, _dp(d_sum), _dp(d_comp), _dp(d_sumsq)))

    def primary_hits(self, d_tri, d_t, camera=None):
        cam = camera if camera is not None else self.camera()
        self._c(lib().lgi_primary_hits(self._h, C.byref(cam), _dp(d_tri), _dp(d_t)))

    def gbuffer(self, d_alb, d_nrm, d_wp, camera=None):
        cam = camera if camera is not None else self.camera()
        self._c(lib().lgi_gbuffer(self._h, C.byref(cam), _dp(d_alb), _dp(d_nrm), _dp(d_wp)))

    def tonemap(self, d_sum, n_px, inv_spp, d_rgb):
        self._c(lib().lgi_tonemap(self._h, _dp(d_sum), n_px, inv_spp, _dp(d_rgb)))

    def last_render_ms(self):
        ms = C.c_float()
        self._c(lib().lgi_last_render_ms(self._h, C.byref(ms)))
        return ms.value

    def last_trace_ms(self):
        """(summed traversal-kernel ms, launches) of the last render call, from HIP events."""
        ms, n = C.c_float(), C.c_int32()
        self._c(lib().lgi_last_trace_ms(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value


def _dp(x):
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return ptr(x)
