"""ctypes mirror of include/lgi.h and the loader for liblgi.so.

The library is built in-tree (``make -C luminagi-cudaref_amd``) and loaded from this
directory; a missing library is an error, never a silent fallback.
"""
import ctypes as C
import os
import pathlib

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = _HERE / "liblgi.so"

LGI_OK = 0
STATUS_NAMES = {
    0: "LGI_OK", 1: "LGI_ERR_INVALID_ARG", 2: "LGI_ERR_IO", 3: "LGI_ERR_PARSE", 4: "LGI_ERR_VERSION",
    5: "LGI_ERR_HIP", 6: "LGI_ERR_NO_SCENE", 7: "LGI_ERR_NO_BVH", 8: "LGI_ERR_UNSUPPORTED",
}

RNG_PCG32 = 0
RNG_XORSHIFT_REF = 1
RENDER_SORTED = 2
RENDER_NO_RR = 4
RENDER_NO_RAY_SORT = 8
RENDER_SET_SUMS = 16


class Camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("camera_to_world", C.c_float * 16), ("fov_y_deg", C.c_float),
                ("aspect", C.c_float), ("near_z", C.c_float), ("far_z", C.c_float),
                ("width", C.c_int32), ("height", C.c_int32)]


class DirLight(C.Structure):
    _fields_ = [("dir", C.c_float * 3), ("color", C.c_float * 3), ("intensity", C.c_float)]


class PointLight(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("color", C.c_float * 3), ("intensity", C.c_float),
                ("radius", C.c_float)]


class Material(C.Structure):
    _fields_ = [("albedo", C.c_float * 3), ("emissive", C.c_float * 3)]


class PoseJob(C.Structure):
    _fields_ = [("cam", Camera), ("spp", C.c_int32), ("sample0", C.c_int32)]


class BatchJob(C.Structure):
    _fields_ = [("out_png", C.c_char * 512), ("spp", C.c_int32), ("bounces", C.c_int32),
                ("cam_pos", C.c_float * 3), ("cam_tgt", C.c_float * 3)]


# numpy views of the packed records (byte-identical to Triangle / BvhNode)
TRI_DTYPE = np.dtype([("v", "<f4", (3, 3)), ("n", "<f4", (3, 3)), ("uv", "<f4", (3, 2)), ("material", "<i4")])
NODE_DTYPE = np.dtype([("bb_min", "<f4", 3), ("bb_max", "<f4", 3), ("left", "<i4"), ("right", "<i4"),
                       ("first_tri", "<i4"), ("tri_count", "<i4")])
POINT_DTYPE = np.dtype([("pos", "<f4", 3), ("color", "<f4", 3), ("intensity", "<f4"), ("radius", "<f4")])
assert TRI_DTYPE.itemsize == 100 and NODE_DTYPE.itemsize == 40 and POINT_DTYPE.itemsize == 32

_P = C.c_void_p
_F3 = C.POINTER(C.c_float)

# name -> (restype, argtypes)
SIGNATURES = {
    "lgi_create": (C.c_int, [C.POINTER(_P), C.c_int, _P]),
    "lgi_destroy": (None, [_P]),
    "lgi_last_error": (C.c_char_p, [_P]),
    "lgi_set_stream": (C.c_int, [_P, _P]),
    "lgi_version": (C.c_char_p, []),
    "lgi_scene_load_json": (C.c_int, [_P, C.c_char_p, C.c_int]),
    "lgi_scene_save_json": (C.c_int, [_P, C.c_char_p]),
    "lgi_scene_load_bin": (C.c_int, [_P, C.c_char_p, C.c_int]),
    "lgi_scene_save_bin": (C.c_int, [_P, C.c_char_p]),
    "lgi_scene_set": (C.c_int, [_P, C.POINTER(Camera), C.POINTER(DirLight), C.POINTER(DirLight), C.c_int32,
                                 _P, C.c_int32, C.POINTER(Material), C.c_int32, _P, C.c_int64]),
    "lgi_scene_generate": (C.c_int, [_P, C.c_int32, C.c_uint64, C.c_int32, C.c_int32]),
    "lgi_scene_counts": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                    C.POINTER(C.c_int32)]),
    "lgi_scene_get": (C.c_int, [_P, C.POINTER(Camera), C.POINTER(DirLight), _P, C.POINTER(Material), _P]),
    "lgi_scene_set_camera": (C.c_int, [_P, C.POINTER(Camera)]),
    "lgi_camera_look_at": (C.c_int, [C.POINTER(Camera), _F3, _F3]),
    "lgi_bvh_build": (C.c_int, [_P]),
    "lgi_bvh_node_count": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "lgi_bvh_get": (C.c_int, [_P, _P, _P, _P]),
    "lgi_bvh_last_build_ms": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "lgi_bvh_set_treelet_passes": (C.c_int, [_P, C.c_int32]),
    "lgi_bvh_last_treelet_ms": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "lgi_bvh_get_traversal": (C.c_int, [_P, _P]),
    "lgi_render": (C.c_int, [_P, C.POINTER(Camera), C.c_int32, C.c_int32, C.c_uint32, _P, _P, _P, _P, _P]),
    "lgi_render_accumulate": (C.c_int, [_P, C.POINTER(Camera), C.c_int32, C.c_int32, C.c_uint32, C.c_int32,
                                         _P, _P, _P]),
    "lgi_render_accumulate_noalloc": (C.c_int, [_P, C.POINTER(Camera), C.c_int32, C.c_int32, C.c_uint32,
                                                 C.c_int32, _P, _P, _P]),
    "lgi_primary_hits": (C.c_int, [_P, C.POINTER(Camera), _P, _P]),
    "lgi_tonemap": (C.c_int, [_P, _P, C.c_int32, C.c_float, _P]),
    "lgi_gbuffer": (C.c_int, [_P, C.POINTER(Camera), _P, _P, _P]),
    "lgi_last_render_ms": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "lgi_last_trace_ms": (C.c_int, [_P, C.POINTER(C.c_float), C.POINTER(C.c_int32)]),
    "lgi_render_batch": (C.c_int, [_P, C.POINTER(PoseJob), C.c_int32, C.c_int32, C.c_uint32, _P, _P, _P,
                                    C.c_int32]),
    "lgi_render_tiles": (C.c_int, [_P, C.POINTER(Camera), _P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint32,
                                    _P, _P, _P]),
    "lgi_wavefront_samples_per_pass": (C.c_int, [C.c_int64, C.c_int32, C.c_int32, C.POINTER(C.c_int32)]),
    "lgi_batch_parse": (C.c_int, [C.c_char_p, C.POINTER(BatchJob), C.c_int32, C.POINTER(C.c_int32)]),
    "lgi_write_png": (C.c_int, [C.c_char_p, _P, C.c_int32, C.c_int32]),
    "lgi_read_png": (C.c_int, [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), _P]),
    "lgi_set_asset_root": (C.c_int, [_P, C.c_char_p]),
    "lgi_image_metrics": (C.c_int, [_P, _P, C.c_int32, C.c_int32, _P, _P, _P]),
    "lgi_diff_heatmap": (C.c_int, [_P, _P, C.c_int32, C.c_int32, C.c_double, _P, _P]),
    "lgi_heat_lut": (C.c_int, [_P]),
    "lgi_probe_read_bw": (C.c_int, [_P, C.c_uint64, C.c_int32, C.c_int32, _P, _P, _P]),
    # denoiser layers (device pointers + hipStream_t)
    "lgi_unet_conv3x3_fwd": (C.c_int, [_P, C.c_int32, _P, C.c_int32, _P, _P, C.c_int32, C.c_int32, C.c_int32,
                                       C.c_int32, C.c_int32, _P, _P]),
    "lgi_unet_conv3x3_pool_fwd": (C.c_int, [_P, C.c_int32, _P, C.c_int32, _P, _P, C.c_int32, C.c_int32,
                                            C.c_int32, C.c_int32, C.c_int32, _P, _P, _P]),
    "lgi_unet_conv3x3_tail_radiance_fwd": (C.c_int, [_P, C.c_int32, _P, _P, C.c_int32, C.c_int32, C.c_int32,
                                                     C.c_int32, _P, _P, _P, _P, _P]),
    "lgi_unet_conv3x3_fp8_fwd": (C.c_int, [_P, C.c_int32, _P, C.c_int32, _P, _P, C.c_float, _P, C.c_int32, C.c_int32,
                                           C.c_int32, C.c_int32, C.c_int32, _P, _P, _P]),
    "lgi_unet_conv3x3_first_radiance_fwd": (C.c_int, [_P, _P, C.c_int32, _P, C.c_int32, C.c_int32, C.c_int32,
                                                      C.c_int32, _P, _P]),
    "lgi_unet_tail_radiance_fwd": (C.c_int, [_P, _P, _P, _P, C.c_int64, _P, _P]),
    "lgi_unet_conv3x3_first_fwd": (C.c_int, [_P, _P, C.c_int32, _P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _P,
                                             _P]),
    "lgi_unet_conv3x3_up_fwd": (C.c_int, [_P, C.c_int32, _P, C.c_int32, _P, _P, _P, C.c_int32, C.c_int32, C.c_int32,
                                          C.c_int32, C.c_int32, _P, _P]),
    "lgi_unet_up_phase_weights": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, _P, _P]),
    "lgi_unet_conv3x3_bwd_data": (C.c_int, [_P, _P, _P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                            C.c_int32, _P, _P, _P, _P]),
    "lgi_unet_conv3x3_bwd_weight": (C.c_int, [_P, _P, C.c_int32, _P, C.c_int32, _P, C.c_int32, C.c_int32, C.c_int32,
                                              C.c_int32, _P, _P, _P, C.c_int64, _P]),
    "lgi_unet_conv_tuning": (C.c_int, [C.c_int32]),
    "lgi_unet_conv_force_variant": (C.c_int, [C.c_int32]),
    "lgi_unet_conv_tune_table": (C.c_int, [_P, _P, C.c_int32, C.POINTER(C.c_int32)]),
    "lgi_unet_maxpool_fwd": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _P, _P]),
    "lgi_unet_maxpool_bwd": (C.c_int, [_P, _P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _P, _P]),
    "lgi_unet_upsample_fwd": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _P, _P]),
    "lgi_unet_upsample_bwd": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _P, _P]),
    "lgi_unet_pack_input": (C.c_int, [_P, C.c_int64, _P, _P]),
    "lgi_unet_tail_fwd": (C.c_int, [_P, _P, _P, _P, C.c_int64, _P, _P]),
    "lgi_unet_tail_bwd": (C.c_int, [_P, _P, _P, C.c_int64, _P, _P, _P, _P, C.c_int64, _P]),
    "lgi_unet_wgrad_scratch_floats": (C.c_int64, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
}

_lib = None


def header_symbols():
    """Function names declared with LGI_API in include/lgi.h."""
    import re
    hdr = _HERE.parent.parent / "include" / "lgi.h"
    text = hdr.read_text()
    return sorted(set(re.findall(r"LGI_API\s+[\w\s\*]+?\b(lgi_\w+)\s*\(", text)))


def lib():
    """Load liblgi.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C {LIB_PATH.parent.parent}`")
        # torch (when imported first) has already loaded the HIP runtime this library binds to
        l = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | os.RTLD_GLOBAL)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


class LgiError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


def ptr(a):
    """Raw pointer of a numpy array or torch tensor (None passes through)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
        return a.ctypes.data
    assert a.is_contiguous(), "tensor must be contiguous"
    return a.data_ptr()
