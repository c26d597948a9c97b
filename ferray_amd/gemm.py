"""Python mirror of ferray-linalg's BLAS-style GEMM entry points.

Each function keeps the name, argument order and meaning of its counterpart
in ferray-linalg/src/gemm/mod.rs and returns ``bool`` exactly like it: True
when the device computed the product into ``c`` (a contiguous numpy array,
updated in place), False when the backend could not run.  Operand lengths are
checked (ferray only ``debug_assert``s them, mod.rs:207-209) and a short
buffer raises ValueError before anything is sent to the device.

``DeviceBuffer`` and the ``*_dev`` helpers drive the device-pointer entry
points on operands that stay resident in HBM.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib

__all__ = [
    "gpu_supports", "offload_min_dim", "gemm_f64", "gemm_f32", "gemm_c64", "gemm_c32", "gemm_i8",
    "gemm_i8_signed", "gemm_i16", "quantized_matmul", "gemm_bf16_f32",
    "gemm_f16_f32", "gemm_e4m3_f32", "gemm_e5m2_f32", "dgemm_ozaki", "last_status", "DeviceBuffer", "fill", "sync",
    "sgemm_bf16x3", "trsm_left_lower_f64", "trsm_left_upper_f64",
    "trsm_left_lower_f32", "trsm_left_upper_f32", "FILL_F64_UNIFORM", "FILL_F64_NORMAL", "FILL_F64_WIDE", "FILL_F32_UNIFORM",
    "FILL_F32_NORMAL", "FILL_I8", "FILL_I16", "FILL_U8", "Event", "device_sync",
]

FILL_F64_UNIFORM, FILL_F64_NORMAL, FILL_F64_WIDE = 0, 1, 2
FILL_F32_UNIFORM, FILL_F32_NORMAL = 3, 4
FILL_I8, FILL_I16, FILL_U8 = 5, 6, 7


def _ptr(arr: np.ndarray) -> int:
    return arr.ctypes.data


def _need(arr: np.ndarray, dtype, count: int, name: str, writable: bool = False) -> None:
    if not isinstance(arr, np.ndarray):
        raise TypeError(f"{name} must be a numpy array")
    if arr.dtype != np.dtype(dtype):
        raise TypeError(f"{name} must have dtype {np.dtype(dtype)}, got {arr.dtype}")
    if not arr.flags.c_contiguous:
        raise ValueError(f"{name} must be C-contiguous")
    if arr.size < count:
        raise ValueError(f"{name} holds {arr.size} elements, needs {count}")
    if writable and not arr.flags.writeable:
        raise ValueError(f"{name} must be writable")


def gpu_supports() -> bool:
    """Device analogue of gemm::cpu_supports_avx2_fma (mod.rs:118-188)."""
    return bool(lib().ferray_gpu_available())


def offload_min_dim(dtype: int) -> int:
    """Smallest max(m, n, k) from which a host-pointer call pays off against
    ferray's CPU GEMM (measured on MI355X; FERRAY_OFFLOAD_MIN_DIM overrides):
    the counterpart of the reference's size dispatch (products/mod.rs:445,
    :499; complex.rs:160) for callers that keep their CPU kernels below it."""
    return int(lib().ferray_gemm_offload_min_dim(int(dtype)))


def last_status() -> int:
    return int(lib().ferray_gemm_last_status())


def gemm_f64(m, n, k, alpha, a, b, beta, c) -> bool:
    """C = alpha*A@B + beta*C (mod.rs:197-328)."""
    _need(a, np.float64, m * k, "a"); _need(b, np.float64, k * n, "b")
    _need(c, np.float64, m * n, "c", True)
    return bool(lib().ferray_gemm_f64(m, n, k, alpha, _ptr(a), _ptr(b), beta, _ptr(c)))


def dgemm_ozaki(m, n, k, alpha, a, b, beta, c, moduli: int = 0) -> bool:
    """Split-precision C = alpha*A@B + beta*C on INT8 matrix cores (integer
    modular Ozaki scheme); same contract as gemm_f64 plus the modulus count."""
    _need(a, np.float64, m * k, "a"); _need(b, np.float64, k * n, "b")
    _need(c, np.float64, m * n, "c", True)
    return bool(lib().ferray_dgemm_ozaki(m, n, k, alpha, _ptr(a), _ptr(b), beta, _ptr(c),
                                         int(moduli)))


def gemm_f32(m, n, k, alpha, a, b, beta, c) -> bool:
    """f32 sibling of gemm_f64 (mod.rs:940-1062)."""
    _need(a, np.float32, m * k, "a"); _need(b, np.float32, k * n, "b")
    _need(c, np.float32, m * n, "c", True)
    return bool(lib().ferray_gemm_f32(m, n, k, alpha, _ptr(a), _ptr(b), beta, _ptr(c)))


def _trsm(fn, dtype, m, n, alpha, a, lda, b, ldb, unit_diag) -> bool:
    if lda < m or ldb < n:
        raise ValueError("lda must be >= m and ldb >= n")
    _need(a, dtype, m * lda if m else 0, "a")
    _need(b, dtype, m * ldb if m else 0, "b", True)
    return bool(fn(m, n, alpha, _ptr(a), lda, _ptr(b), ldb, bool(unit_diag)))


def trsm_left_lower_f64(m, n, alpha, a, lda, b, ldb, unit_diag) -> bool:
    """L X = alpha B in place, L lower triangular (trsm.rs:46-68)."""
    return _trsm(lib().ferray_trsm_left_lower_f64, np.float64, m, n, alpha, a, lda, b, ldb,
                 unit_diag)


def trsm_left_upper_f64(m, n, alpha, a, lda, b, ldb, unit_diag) -> bool:
    """U X = alpha B in place, U upper triangular (trsm.rs:70-92)."""
    return _trsm(lib().ferray_trsm_left_upper_f64, np.float64, m, n, alpha, a, lda, b, ldb,
                 unit_diag)


def trsm_left_lower_f32(m, n, alpha, a, lda, b, ldb, unit_diag) -> bool:
    """f32 sibling (trsm.rs:94-116)."""
    return _trsm(lib().ferray_trsm_left_lower_f32, np.float32, m, n, alpha, a, lda, b, ldb,
                 unit_diag)


def trsm_left_upper_f32(m, n, alpha, a, lda, b, ldb, unit_diag) -> bool:
    """f32 sibling (trsm.rs:118-140)."""
    return _trsm(lib().ferray_trsm_left_upper_f32, np.float32, m, n, alpha, a, lda, b, ldb,
                 unit_diag)


def sgemm_bf16x3(m, n, k, alpha, a, b, beta, c, terms: int = 3) -> bool:
    """f32 C = alpha*A@B + beta*C on BF16 matrix cores by operand splitting
    (terms 3: ~2^-16 relative error per product; terms 6: f32 level)."""
    _need(a, np.float32, m * k, "a"); _need(b, np.float32, k * n, "b")
    _need(c, np.float32, m * n, "c", True)
    return bool(lib().ferray_sgemm_bf16x3(m, n, k, alpha, _ptr(a), _ptr(b), beta, _ptr(c), terms))


def _cplx_view(arr: np.ndarray, real_dtype, count: int, name: str, writable=False) -> np.ndarray:
    if arr.dtype in (np.complex128, np.complex64):
        arr = arr.view(real_dtype)
    _need(arr, real_dtype, 2 * count, name, writable)
    return arr


def gemm_c64(m, n, k, alpha_re, alpha_im, a, b, beta_re, beta_im, c) -> bool:
    """Complex<f64> C = alpha*A@B + beta*C, interleaved (mod.rs:757-826).

    ``a``, ``b``, ``c`` may be complex128 arrays or their float64 views.
    """
    a = _cplx_view(a, np.float64, m * k, "a"); b = _cplx_view(b, np.float64, k * n, "b")
    c = _cplx_view(c, np.float64, m * n, "c", True)
    return bool(lib().ferray_gemm_c64(m, n, k, alpha_re, alpha_im, _ptr(a), _ptr(b),
                                      beta_re, beta_im, _ptr(c)))


def gemm_c32(m, n, k, alpha_re, alpha_im, a, b, beta_re, beta_im, c) -> bool:
    """Complex<f32> sibling (mod.rs:669-736)."""
    a = _cplx_view(a, np.float32, m * k, "a"); b = _cplx_view(b, np.float32, k * n, "b")
    c = _cplx_view(c, np.float32, m * n, "c", True)
    return bool(lib().ferray_gemm_c32(m, n, k, alpha_re, alpha_im, _ptr(a), _ptr(b),
                                      beta_re, beta_im, _ptr(c)))


def gemm_i8(m, n, k, a, b, c, accumulate) -> bool:
    """u8 x i8 -> i32, exact wrapping (mod.rs:601-651)."""
    _need(a, np.uint8, m * k, "a"); _need(b, np.int8, k * n, "b")
    _need(c, np.int32, m * n, "c", True)
    return bool(lib().ferray_gemm_i8(m, n, k, _ptr(a), _ptr(b), _ptr(c), bool(accumulate)))


def gemm_i8_signed(m, n, k, a, b, c, accumulate) -> bool:
    """i8 x i8 -> i32 wrapping (mod.rs:414-448)."""
    _need(a, np.int8, m * k, "a"); _need(b, np.int8, k * n, "b")
    _need(c, np.int32, m * n, "c", True)
    return bool(lib().ferray_gemm_i8_signed(m, n, k, _ptr(a), _ptr(b), _ptr(c),
                                            bool(accumulate)))


def gemm_i16(m, n, k, a, b, c, accumulate) -> bool:
    """i16 x i16 -> i32 wrapping (mod.rs:348-392)."""
    _need(a, np.int16, m * k, "a"); _need(b, np.int16, k * n, "b")
    _need(c, np.int32, m * n, "c", True)
    return bool(lib().ferray_gemm_i16(m, n, k, _ptr(a), _ptr(b), _ptr(c), bool(accumulate)))


def quantized_matmul(m, n, k, a, a_zp, scale_a, b, b_zp, scale_b, c, accumulate) -> bool:
    """Zero-point / scale quantized matmul with f32 output (mod.rs:484-575)."""
    _need(a, np.uint8, m * k, "a"); _need(b, np.int8, k * n, "b")
    _need(c, np.float32, m * n, "c", True)
    return bool(lib().ferray_quantized_matmul(m, n, k, _ptr(a), int(a_zp), scale_a, _ptr(b),
                                              int(b_zp), scale_b, _ptr(c), bool(accumulate)))


def _half_bits(arr: np.ndarray, name: str, count: int) -> np.ndarray:
    if arr.dtype == np.float16:
        arr = arr.view(np.uint16)
    _need(arr, np.uint16, count, name)
    return arr


def gemm_bf16_f32(m, n, k, a, b, c, accumulate) -> bool:
    """bf16 x bf16 -> f32; ``a``/``b`` hold raw bf16 bits as uint16 (mod.rs:851-882)."""
    a = _half_bits(a, "a", m * k); b = _half_bits(b, "b", k * n)
    _need(c, np.float32, m * n, "c", True)
    return bool(lib().ferray_gemm_bf16_f32(m, n, k, _ptr(a), _ptr(b), _ptr(c), bool(accumulate)))


def gemm_f16_f32(m, n, k, a, b, c, accumulate) -> bool:
    """f16 x f16 -> f32; float16 arrays or their uint16 bits (mod.rs:906-933)."""
    a = _half_bits(a, "a", m * k); b = _half_bits(b, "b", k * n)
    _need(c, np.float32, m * n, "c", True)
    return bool(lib().ferray_gemm_f16_f32(m, n, k, _ptr(a), _ptr(b), _ptr(c), bool(accumulate)))


def gemm_e4m3_f32(m, n, k, a, b, c, accumulate) -> bool:
    """OCP E4M3 x E4M3 -> f32; ``a``/``b`` hold the fp8 bytes as uint8."""
    _need(a, np.uint8, m * k, "a"); _need(b, np.uint8, k * n, "b")
    _need(c, np.float32, m * n, "c", True)
    return bool(lib().ferray_gemm_e4m3_f32(m, n, k, _ptr(a), _ptr(b), _ptr(c), bool(accumulate)))


def gemm_e5m2_f32(m, n, k, a, b, c, accumulate) -> bool:
    """OCP E5M2 x E5M2 -> f32; ``a``/``b`` hold the fp8 bytes as uint8."""
    _need(a, np.uint8, m * k, "a"); _need(b, np.uint8, k * n, "b")
    _need(c, np.float32, m * n, "c", True)
    return bool(lib().ferray_gemm_e5m2_f32(m, n, k, _ptr(a), _ptr(b), _ptr(c), bool(accumulate)))


# ---------------------------------------------------------------------------
# Device-resident operands
# ---------------------------------------------------------------------------

class DeviceBuffer:
    """A block of device memory owned by the library's allocator."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib().ferray_dev_malloc(max(self.nbytes, 1), C.byref(p)), "dev_malloc")
        self.ptr = int(p.value)

    @classmethod
    def from_array(cls, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        buf = cls(arr.nbytes)
        buf.upload(arr)
        return buf

    def upload(self, arr: np.ndarray) -> None:
        arr = np.ascontiguousarray(arr)
        if arr.nbytes > self.nbytes:
            raise ValueError("array larger than buffer")
        check(lib().ferray_dev_upload(self.ptr, arr.ctypes.data, arr.nbytes, None), "upload")

    def download(self, dtype, count: int | None = None, offset_bytes: int = 0) -> np.ndarray:
        dtype = np.dtype(dtype)
        if count is None:
            count = (self.nbytes - offset_bytes) // dtype.itemsize
        out = np.empty(count, dtype=dtype)
        if offset_bytes + out.nbytes > self.nbytes:
            raise ValueError("download beyond buffer end")
        check(lib().ferray_dev_download(out.ctypes.data, self.ptr + offset_bytes, out.nbytes,
                                        None), "download")
        return out

    def free(self) -> None:
        if self.ptr:
            lib().ferray_dev_free(self.ptr)
            self.ptr = 0

    def __del__(self):  # pragma: no cover - interpreter shutdown ordering
        try:
            self.free()
        except Exception:
            pass


def fill(buf: DeviceBuffer, count: int, kind: int, seed: int, lo: int = 0, hi: int = 0,
         stream=None) -> None:
    """Seeded device generator (see ferray_dev_fill)."""
    check(lib().ferray_dev_fill(buf.ptr, count, kind, seed, lo, hi, stream), "dev_fill")


def sync(stream=None) -> None:
    check(lib().ferray_stream_sync(stream), "stream_sync")


class Event:
    """A HIP event recorded on a library stream (NULL = the thread's stream)."""

    def __init__(self):
        p = C.c_void_p()
        check(lib().ferray_event_create(C.byref(p)), "event_create")
        self.ptr = p.value

    def record(self, stream=None) -> None:
        check(lib().ferray_event_record(self.ptr, stream), "event_record")

    def elapsed_ms(self, stop: "Event") -> float:
        ms = C.c_float()
        check(lib().ferray_event_elapsed_ms(self.ptr, stop.ptr, C.byref(ms)), "event_elapsed")
        return float(ms.value)

    def __del__(self):  # pragma: no cover
        try:
            if self.ptr:
                lib().ferray_event_destroy(self.ptr)
                self.ptr = None
        except Exception:
            pass


def device_sync() -> None:
    check(lib().ferray_device_sync(), "device_sync")
