"""lgi — MI355X-native LuminaGI reference tracer and denoiser (Python front end of liblgi.so)."""
from ._abi import (LgiError, RENDER_NO_RAY_SORT, RENDER_NO_RR, RENDER_SET_SUMS, RENDER_SORTED, RNG_PCG32, RNG_XORSHIFT_REF,
                   NODE_DTYPE, TRI_DTYPE, POINT_DTYPE, Camera, DirLight, lib)
from .tracer import Context, camera_look_at, parse_batch, read_png, write_png

__all__ = ["LgiError", "RENDER_NO_RAY_SORT", "RENDER_NO_RR", "RENDER_SET_SUMS", "RENDER_SORTED", "RNG_PCG32", "RNG_XORSHIFT_REF",
           "NODE_DTYPE", "TRI_DTYPE", "POINT_DTYPE", "Camera", "DirLight", "lib", "Context", "camera_look_at",
           "parse_batch", "read_png", "write_png"]
