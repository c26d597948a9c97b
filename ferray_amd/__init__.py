"""ferray_amd — MI355X (gfx950) implementation of ferray-linalg's GEMM family.

The compute lives in libferray_gemm.so (HIP kernels for CDNA4 matrix cores
behind the C ABI in include/ferray_gemm.h); this package is its Python face.
"""
from ._lib import FerrayError, LIB_PATH, lib  # noqa: F401
from .gemm import *  # noqa: F401,F403
from . import grouped as grouped_gemm  # noqa: F401

__version__ = "0.1.0"
