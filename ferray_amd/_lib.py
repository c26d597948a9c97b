"""ctypes binding of libferray_gemm.so (the C ABI in include/ferray_gemm.h).

The library is loaded from the package directory.  There is no fallback: if
the shared object is missing, every call raises instead of silently computing
on the CPU.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libferray_gemm.so")

_sz = C.c_size_t
_p = C.c_void_p
_b = C.c_bool
_i = C.c_int
_d = C.c_double
_f = C.c_float
_u8 = C.c_uint8
_i8 = C.c_int8
_u64 = C.c_uint64
_i64 = C.c_int64

# name -> (restype, argtypes); every function declared in include/ferray_gemm.h
SIGNATURES: dict[str, tuple] = {
    "ferray_gemm_f64": (_b, [_sz, _sz, _sz, _d, _p, _p, _d, _p]),
    "ferray_gemm_f32": (_b, [_sz, _sz, _sz, _f, _p, _p, _f, _p]),
    "ferray_gemm_c64": (_b, [_sz, _sz, _sz, _d, _d, _p, _p, _d, _d, _p]),
    "ferray_gemm_c32": (_b, [_sz, _sz, _sz, _f, _f, _p, _p, _f, _f, _p]),
    "ferray_gemm_i8": (_b, [_sz, _sz, _sz, _p, _p, _p, _b]),
    "ferray_gemm_i8_signed": (_b, [_sz, _sz, _sz, _p, _p, _p, _b]),
    "ferray_gemm_i16": (_b, [_sz, _sz, _sz, _p, _p, _p, _b]),
    "ferray_quantized_matmul": (_b, [_sz, _sz, _sz, _p, _u8, _f, _p, _i8, _f, _p, _b]),
    "ferray_gemm_bf16_f32": (_b, [_sz, _sz, _sz, _p, _p, _p, _b]),
    "ferray_gemm_f16_f32": (_b, [_sz, _sz, _sz, _p, _p, _p, _b]),
    "ferray_gemm_e4m3_f32": (_b, [_sz, _sz, _sz, _p, _p, _p, _b]),
    "ferray_gemm_e5m2_f32": (_b, [_sz, _sz, _sz, _p, _p, _p, _b]),
    "ferray_gpu_available": (_b, []),
    "ferray_gemm_offload_min_dim": (_sz, [_i]),
    "ferray_gemm_last_status": (_i, []),
    "ferray_status_string": (C.c_char_p, [_i]),
    "ferray_gemm_f64_dev": (_i, [_sz, _sz, _sz, _d, _p, _p, _d, _p, _p]),
    "ferray_gemm_f32_dev": (_i, [_sz, _sz, _sz, _f, _p, _p, _f, _p, _p]),
    "ferray_gemm_c64_dev": (_i, [_sz, _sz, _sz, _d, _d, _p, _p, _d, _d, _p, _p]),
    "ferray_gemm_c32_dev": (_i, [_sz, _sz, _sz, _f, _f, _p, _p, _f, _f, _p, _p]),
    "ferray_gemm_f64_strided_dev": (_i, [_sz, _sz, _sz, _d, _p, _sz, _p, _sz, _d, _p, _sz, _p]),
    "ferray_gemm_f32_strided_dev": (_i, [_sz, _sz, _sz, _f, _p, _sz, _p, _sz, _f, _p, _sz, _p]),
    "ferray_gemm_c64_strided_dev": (_i, [_sz, _sz, _sz, _d, _d, _p, _sz, _p, _sz, _d, _d, _p, _sz,
                                         _p]),
    "ferray_gemm_c32_strided_dev": (_i, [_sz, _sz, _sz, _f, _f, _p, _sz, _p, _sz, _f, _f, _p, _sz,
                                         _p]),
    "ferray_gemm_i8_dev": (_i, [_sz, _sz, _sz, _p, _p, _p, _b, _p]),
    "ferray_gemm_i8_signed_dev": (_i, [_sz, _sz, _sz, _p, _p, _p, _b, _p]),
    "ferray_gemm_i8_signed_dev_noalloc": (_i, [_sz, _sz, _sz, _p, _p, _p, _b, _p]),
    "ferray_gemm_i16_dev": (_i, [_sz, _sz, _sz, _p, _p, _p, _b, _p]),
    "ferray_quantized_matmul_dev": (_i, [_sz, _sz, _sz, _p, _u8, _f, _p, _i8, _f, _p, _b, _p]),
    "ferray_gemm_bf16_f32_dev": (_i, [_sz, _sz, _sz, _p, _p, _p, _b, _p]),
    "ferray_gemm_f16_f32_dev": (_i, [_sz, _sz, _sz, _p, _p, _p, _b, _p]),
    "ferray_gemm_e4m3_f32_dev": (_i, [_sz, _sz, _sz, _p, _p, _p, _b, _p]),
    "ferray_gemm_e5m2_f32_dev": (_i, [_sz, _sz, _sz, _p, _p, _p, _b, _p]),
    "ferray_dgemm_ozaki": (_b, [_sz, _sz, _sz, _d, _p, _p, _d, _p, _i]),
    "ferray_dgemm_ozaki_dev": (_i, [_sz, _sz, _sz, _d, _p, _p, _d, _p, _i, _p]),
    "ferray_ozaki_max_moduli": (_i, []),
    "ferray_gemm_grouped_dev": (_i, [_p, _sz, _p]),
    "ferray_grouped_plan_create": (_i, [_sz, _sz, C.POINTER(_p)]),
    "ferray_grouped_plan_launch": (_i, [_p, _p, _sz, _p]),
    "ferray_grouped_plan_sync": (_i, [_p]),
    "ferray_grouped_plan_destroy": (_i, [_p]),
    "ferray_set_tile_table": (_i, [_p, _sz]),
    "ferray_get_tile_table": (_i, [_p, _sz]),
    "ferray_trsm_left_lower_f64": (_b, [_sz, _sz, _d, _p, _sz, _p, _sz, _b]),
    "ferray_trsm_left_upper_f64": (_b, [_sz, _sz, _d, _p, _sz, _p, _sz, _b]),
    "ferray_trsm_left_lower_f32": (_b, [_sz, _sz, _f, _p, _sz, _p, _sz, _b]),
    "ferray_trsm_left_upper_f32": (_b, [_sz, _sz, _f, _p, _sz, _p, _sz, _b]),
    "ferray_trsm_dev": (_i, [_i, _i, _sz, _sz, _d, _p, _sz, _p, _sz, _b, _p]),
    "ferray_sgemm_bf16x3": (_b, [_sz, _sz, _sz, _f, _p, _p, _f, _p, _i]),
    "ferray_sgemm_bf16x3_dev": (_i, [_sz, _sz, _sz, _f, _p, _p, _f, _p, _i, _p]),
    "ferray_matmul_batched": (_b, [_i, _sz, _sz, _sz, _sz, _sz, _sz, _p, _p, _p]),
    "ferray_matmul_batched_dev": (_i, [_i, _sz, _sz, _sz, _sz, _sz, _sz, _p, _p, _p, _p]),
    "ferray_server_create": (_i, [_p, C.POINTER(_p)]),
    "ferray_server_destroy": (_i, [_p]),
    "ferray_server_submit": (_i, [_p, _p, C.POINTER(_u64)]),
    "ferray_server_poll": (_i, [_p, _u64, C.POINTER(_i)]),
    "ferray_server_wait": (_i, [_p, _u64, C.POINTER(_d)]),
    "ferray_server_result": (_i, [_p, _u64, _p, _sz]),
    "ferray_server_result_dev": (_i, [_p, _u64, C.POINTER(_p)]),
    "ferray_server_release": (_i, [_p, _u64]),
    "ferray_server_stats_get": (_i, [_p, _p]),
    "ferray_serve_trace": (_i, [_p, _p, _p, _sz, _p, C.POINTER(_d), _p]),
    "ferray_dev_malloc": (_i, [_sz, C.POINTER(_p)]),
    "ferray_dev_free": (_i, [_p]),
    "ferray_dev_upload": (_i, [_p, _p, _sz, _p]),
    "ferray_dev_download": (_i, [_p, _p, _sz, _p]),
    "ferray_dev_memset": (_i, [_p, _i, _sz, _p]),
    "ferray_dev_copy": (_i, [_p, _p, _sz, _p]),
    "ferray_stream_sync": (_i, [_p]),
    "ferray_default_stream": (_p, []),
    "ferray_dev_fill": (_i, [_p, _sz, _i, _u64, _i64, _i64, _p]),
    "ferray_event_create": (_i, [C.POINTER(_p)]),
    "ferray_event_record": (_i, [_p, _p]),
    "ferray_event_elapsed_ms": (_i, [_p, _p, C.POINTER(_f)]),
    "ferray_event_destroy": (_i, [_p]),
    "ferray_device_sync": (_i, []),
}

DTYPE_F64, DTYPE_F32, DTYPE_C64, DTYPE_C32, DTYPE_I8 = 0, 1, 2, 3, 4


class GemmProblem(C.Structure):
    """struct ferray_gemm_problem (include/ferray_gemm.h)."""
    _fields_ = [("dtype", C.c_int), ("accumulate", C.c_int), ("m", C.c_size_t),
                ("n", C.c_size_t), ("k", C.c_size_t), ("a", C.c_void_p), ("b", C.c_void_p),
                ("c", C.c_void_p), ("alpha_re", C.c_double), ("alpha_im", C.c_double),
                ("beta_re", C.c_double), ("beta_im", C.c_double)]


SERVE_GROUPED, SERVE_PER_REQUEST = 0, 1


class ServerConfig(C.Structure):
    """struct ferray_server_config (include/ferray_gemm.h)."""
    _fields_ = [("hbm_bytes", C.c_size_t), ("max_resident", C.c_size_t),
                ("max_batch", C.c_size_t), ("lanes", C.c_int), ("mode", C.c_int),
                ("keep_results", C.c_int)]


class Request(C.Structure):
    """struct ferray_request (include/ferray_gemm.h)."""
    _fields_ = [("dtype", C.c_int), ("operands_on_host", C.c_int), ("m", C.c_size_t),
                ("n", C.c_size_t), ("k", C.c_size_t), ("alpha_re", C.c_double),
                ("alpha_im", C.c_double), ("a", C.c_void_p), ("b", C.c_void_p)]


class ServerStats(C.Structure):
    """struct ferray_server_stats (include/ferray_gemm.h)."""
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("launches", C.c_uint64),
                ("launched_requests", C.c_uint64), ("arena_bytes", C.c_size_t),
                ("arena_peak", C.c_size_t)]


_lock = threading.Lock()
_lib: C.CDLL | None = None


class FerrayError(RuntimeError):
    """A device entry point returned a non-zero status."""

    def __init__(self, status: int, what: str = ""):
        self.status = status
        super().__init__(f"{what}: status {status} ({status_string(status)})")


def lib() -> C.CDLL:
    """Loads and returns the library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} is missing: build it with "
                    "`python -c 'import __graft_entry__ as g; g.build()'` or "
                    "`python ferray_amd/build.py`")
            handle = C.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(handle, name)
                fn.restype = res
                fn.argtypes = args
            _lib = handle
    return _lib


def status_string(status: int) -> str:
    try:
        s = lib().ferray_status_string(status)
        return s.decode() if s else "?"
    except Exception:  # pragma: no cover - only when the library is absent
        return "?"


def check(status: int, what: str) -> None:
    if status != 0:
        raise FerrayError(status, what)
