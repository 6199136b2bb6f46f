"""ctypes front end of the CPU oracle (liblgi_oracle.so). TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module,
and only as the checker (or the timed CPU baseline) — never as a product path.
Every function here forwards to oracle.cpp, which restates the reference algorithms with
file:line citations.
"""
import contextlib
import ctypes as C
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB = HERE / "liblgi_oracle.so"
REF_BVH = HERE / "_ref" / "ref_bvh"

TRI_DTYPE = np.dtype([("v", "<f4", (3, 3)), ("n", "<f4", (3, 3)), ("uv", "<f4", (3, 2)), ("material", "<i4")])
NODE_DTYPE = np.dtype([("bb_min", "<f4", 3), ("bb_max", "<f4", 3), ("left", "<i4"), ("right", "<i4"),
                       ("first_tri", "<i4"), ("tri_count", "<i4")])
POINT_DTYPE = np.dtype([("pos", "<f4", 3), ("color", "<f4", 3), ("intensity", "<f4"), ("radius", "<f4")])


class OSun(C.Structure):
    _fields_ = [("dir", C.c_float * 3), ("color", C.c_float * 3), ("intensity", C.c_float)]


class OCam(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("c2w", C.c_float * 16), ("fov", C.c_float), ("aspect", C.c_float),
                ("nearz", C.c_float), ("farz", C.c_float), ("width", C.c_int32), ("height", C.c_int32)]


class RenderArgs(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("tris", C.c_void_p), ("mat_albedo", C.c_void_p), ("points", C.c_void_p),
                ("n_points", C.c_int32), ("sun", OSun), ("cam", OCam), ("spp", C.c_int32), ("bounces", C.c_int32),
                ("flags", C.c_uint32), ("seed", C.c_uint32), ("sample0", C.c_int32), ("row_step", C.c_int32),
                ("threads", C.c_int32), ("tex_texels", C.c_void_p), ("tex_desc", C.c_void_p), ("srgb_lut", C.c_void_p)]


def srgb_lut():
    """8-bit sRGB -> linear, evaluated in FP64 and rounded to FP32 (the IEC 61966-2-1 curve)."""
    c = np.arange(256) / 255.0
    return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4).astype(np.float32)


def pack_textures(textures, n_mats):
    """textures: {material index: HxWx4 uint8 RGBA} -> (texels u32, desc int32 [n_mats, 4])."""
    desc = np.zeros((max(1, n_mats), 4), np.int32)
    chunks, off = [], 0
    for m, img in sorted(textures.items()):
        img = np.ascontiguousarray(img, np.uint8)
        h, w = img.shape[:2]
        desc[m] = (off, w, h, 1)
        chunks.append(img.reshape(-1, 4).view("<u4").reshape(-1))
        off += w * h
    texels = np.concatenate(chunks) if chunks else np.zeros(1, np.uint32)
    return np.ascontiguousarray(texels, np.uint32), desc


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} missing: build with `make -C {HERE}`")
        _lib = _load(LIB)
    return _lib


def build_mutant(k, out_dir):
    """Builds oracle.cpp with planted mistake k (see ORC_MUTANT in oracle.cpp) into out_dir and
    returns the library path. Only the oracle's own mutation tests use this."""
    out = pathlib.Path(out_dir) / f"liblgi_oracle_mutant{k}.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-pthread",
                    "-shared", f"-DORC_MUTANT={int(k)}", str(HERE / "oracle.cpp"), "-o", str(out)], check=True)
    return out


@contextlib.contextmanager
def library(path):
    """Routes every oracle call through another build of oracle.cpp while the block runs."""
    global _lib
    saved = _lib
    _lib = _load(path)
    try:
        yield _lib
    finally:
        _lib = saved


def _load(path):
    l = C.CDLL(str(path))
    V = C.c_void_p
    sig = {
        "orc_morton3": (C.c_uint32, [C.c_float] * 3),
        "orc_tex_sample": (None, [V, V, V, V, C.c_int32, C.c_float, C.c_float, V]),
        "orc_spread3": (C.c_uint32, [C.c_uint32]),
        "orc_sincos": (None, [V, V, V, C.c_int64]),
        "orc_pcg32_stream": (None, [C.c_uint64, C.c_uint64, V, C.c_int]),
        "orc_pixel_randf": (None, [C.c_uint32, C.c_uint32, C.c_uint32, V, C.c_int]),
        "orc_xorshift_stream": (None, [C.c_uint32, V, C.c_int]),
        "orc_sample_cos_hemi": (None, [V, C.c_uint64, C.c_uint64, V, C.c_int]),
        "orc_morton_codes": (None, [V, C.c_int64, V]),
        "orc_bvh_build": (C.c_int64, [V, C.c_int64, V, V, V]),
        "orc_tan_half_fov": (C.c_float, [C.c_float]),
        "orc_camera_look_at": (None, [V, V, V]),
        "orc_render": (None, [C.POINTER(RenderArgs), V, V, V, V, V]),
        "orc_primary_hits": (None, [C.POINTER(RenderArgs), V, V]),
        "orc_brute_closest": (C.c_int32, [V, C.c_int64, V, V, V]),
        "orc_bvh_closest": (C.c_int32, [V, V, V, V, V]),
        "orc_gbuffer": (None, [C.POINTER(RenderArgs), V, V, V]),
        "orc_tonemap": (None, [V, C.c_int64, C.c_float, V]),
    }
    for k, (r, a) in sig.items():
        f = getattr(l, k)
        f.restype, f.argtypes = r, a
    return l


def _p(a):
    return None if a is None else a.ctypes.data


def morton_codes(tris):
    tris = np.ascontiguousarray(tris, TRI_DTYPE)
    out = np.zeros(len(tris), np.uint32)
    lib().orc_morton_codes(_p(tris), len(tris), _p(out))
    return out


def bvh_build(tris):
    """Restated BuildBVH: (nodes, perm, sorted keys). Stable Morton sort."""
    tris = np.ascontiguousarray(tris, TRI_DTYPE)
    n = len(tris)
    nodes = np.zeros(max(1, 2 * n), NODE_DTYPE)
    perm = np.zeros(n, np.int32)
    keys = np.zeros(n, np.uint32)
    m = lib().orc_bvh_build(_p(tris), n, _p(nodes), _p(perm), _p(keys))
    return nodes[:m].copy(), perm, keys


def ref_bvh_available():
    return REF_BVH.exists()


def ref_bvh(scene_path, tmp_bin, timed=False):
    """Reference BuildBVH compiled from the reference sources, on a scene.json or a binary .lgis
    scene file: (nodes, perm), and with timed=True also the milliseconds of the BuildBVH call."""
    r = subprocess.run([str(REF_BVH), str(scene_path), str(tmp_bin)], check=True, capture_output=True, text=True)
    raw = pathlib.Path(tmp_bin).read_bytes()
    nt, nn = np.frombuffer(raw[:16], np.int64)
    nodes = np.frombuffer(raw[16:16 + 40 * nn], NODE_DTYPE).copy()
    perm = np.frombuffer(raw[16 + 40 * nn:16 + 40 * nn + 4 * nt], np.int32).copy()
    if timed:
        ms = float(r.stdout.split("BuildBVH")[1].split("ms")[0])
        return nodes, perm, ms
    return nodes, perm


def sincos(x):
    x = np.ascontiguousarray(x, np.float32)
    s, c = np.zeros_like(x), np.zeros_like(x)
    lib().orc_sincos(_p(x), _p(s), _p(c), len(x))
    return s, c


def pcg32(initstate, initseq, n):
    out = np.zeros(n, np.uint32)
    lib().orc_pcg32_stream(initstate, initseq, _p(out), n)
    return out


def xorshift(seed, n):
    out = np.zeros(n, np.uint32)
    lib().orc_xorshift_stream(seed, _p(out), n)
    return out


def sample_cos_hemi(normal, initstate, initseq, n):
    N = np.asarray(normal, np.float32)
    out = np.zeros((n, 3), np.float32)
    lib().orc_sample_cos_hemi(_p(N), initstate, initseq, _p(out), n)
    return out


def pixel_randf(seed, pixel, sample, n):
    out = np.zeros(n, np.float32)
    lib().orc_pixel_randf(seed, pixel, sample, _p(out), n)
    return out


def tan_half_fov(fov):
    return lib().orc_tan_half_fov(fov)


def camera_look_at(cam, pos, tgt):
    pos = np.asarray(pos, np.float32)
    tgt = np.asarray(tgt, np.float32)
    lib().orc_camera_look_at(C.addressof(cam), _p(pos), _p(tgt))
    return cam


def tex_sample(textures, mat_albedo, mat, u, v):
    """sample_albedo of one material at (u, v); textures as for Scene."""
    mat_albedo = np.ascontiguousarray(mat_albedo, np.float32).reshape(-1, 3)
    texels, desc = pack_textures(textures, len(mat_albedo))
    lut = srgb_lut()
    out = np.zeros(3, np.float32)
    lib().orc_tex_sample(_p(texels), _p(desc), _p(lut), _p(mat_albedo), mat, u, v, _p(out))
    return out


class Scene:
    """Oracle-side scene in BVH order: nodes + permuted triangles + lights + camera."""

    def __init__(self, tris, mat_albedo, points, sun, camera, textures=None):
        """textures (optional): {material index: HxWx4 uint8 RGBA} diffuse textures."""
        self.nodes, self.perm, self.keys = bvh_build(tris)
        self.tex = None
        if textures:
            self.tex = pack_textures(textures, len(np.asarray(mat_albedo).reshape(-1, 3))) + (srgb_lut(),)
        self.tris = np.ascontiguousarray(np.asarray(tris, TRI_DTYPE)[self.perm])
        self.mat_albedo = np.ascontiguousarray(mat_albedo, np.float32).reshape(-1, 3)
        self.points = np.ascontiguousarray(points, POINT_DTYPE)
        self.sun = OSun()
        self.sun.dir[:] = list(sun.dir)
        self.sun.color[:] = list(sun.color)
        self.sun.intensity = sun.intensity
        self.cam = OCam()
        C.memmove(C.addressof(self.cam), C.addressof(camera), C.sizeof(OCam))

    def args(self, spp=1, bounces=0, flags=0, sample0=0, row_step=1, threads=0, camera=None, seed=0):
        a = RenderArgs()
        a.nodes, a.tris, a.mat_albedo, a.points = _p(self.nodes), _p(self.tris), _p(self.mat_albedo), _p(self.points)
        a.n_points = len(self.points)
        a.sun = self.sun
        if self.tex is not None:
            a.tex_texels, a.tex_desc, a.srgb_lut = _p(self.tex[0]), _p(self.tex[1]), _p(self.tex[2])
        if camera is not None:
            C.memmove(C.addressof(a.cam), C.addressof(camera), C.sizeof(OCam))
        else:
            a.cam = self.cam
        a.spp, a.bounces, a.flags, a.seed, a.sample0 = spp, bounces, flags, seed, sample0
        a.row_step, a.threads = row_step, threads
        return a

    def render(self, spp, bounces, flags=0, sample0=0, row_step=1, threads=0, camera=None, sums=None):
        """Returns dict(sum, comp, sum64, sumsq, counters) with sum/comp of shape (H, W, 3)."""
        a = self.args(spp, bounces, flags, sample0, row_step, threads, camera)
        H, W = a.cam.height, a.cam.width
        if sums is None:
            sums = {"sum": np.zeros((H, W, 3), np.float32), "comp": np.zeros((H, W, 3), np.float32),
                    "sum64": np.zeros((H, W, 3), np.float64), "sumsq": np.zeros((H, W), np.float32)}
        cnt = np.zeros(5, np.uint64)
        lib().orc_render(C.byref(a), _p(sums["sum"]), _p(sums["comp"]), _p(sums["sum64"]), _p(sums["sumsq"]),
                         _p(cnt))
        sums["counters"] = dict(zip(["node_visits", "tri_tests", "closest_hits", "rays", "shadow_rays"],
                                    [int(x) for x in cnt]))
        return sums

    def primary_hits(self, camera=None, threads=0):
        a = self.args(threads=threads, camera=camera)
        H, W = a.cam.height, a.cam.width
        tri = np.zeros((H, W), np.int32)
        t = np.zeros((H, W), np.float32)
        lib().orc_primary_hits(C.byref(a), _p(tri), _p(t))
        orig = np.where(tri >= 0, self.perm[np.maximum(tri, 0)], -1).astype(np.int32)
        return orig, t

    def gbuffer(self, camera=None):
        a = self.args(camera=camera)
        H, W = a.cam.height, a.cam.width
        bufs = [np.zeros((H, W, 3), np.uint8) for _ in range(3)]
        lib().orc_gbuffer(C.byref(a), *[_p(b) for b in bufs])
        return bufs


def tonemap(accum, inv_spp):
    accum = np.ascontiguousarray(accum, np.float32)
    out = np.zeros(accum.shape, np.uint8)
    lib().orc_tonemap(_p(accum), accum.size // 3, inv_spp, _p(out))
    return out


def brute_closest(tris, o, d):
    tris = np.ascontiguousarray(tris, TRI_DTYPE)
    o = np.asarray(o, np.float32)
    d = np.asarray(d, np.float32)
    t = np.zeros(1, np.float32)
    i = lib().orc_brute_closest(_p(tris), len(tris), _p(o), _p(d), _p(t))
    return i, float(t[0])


def bvh_closest(nodes, tris, o, d):
    o = np.asarray(o, np.float32)
    d = np.asarray(d, np.float32)
    t = np.zeros(1, np.float32)
    i = lib().orc_bvh_closest(_p(nodes), _p(tris), _p(o), _p(d), _p(t))
    return i, float(t[0])
