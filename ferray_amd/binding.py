"""The Python binding's `matmul` (`numpy.matmul` / `@`), on the GPU GEMM family.

Mirrors ferray-python/src/linalg.rs:1219-1284 for host (numpy) operands:

* float32 / float64: `b` is coerced to `a`'s dtype; 1-D x 1-D collapses to
  the inner product returned as a numpy scalar (the gufunc signature
  (n?,k),(k,m?)->(n?,m?)); every other rank combination follows the
  product layer's matmul (2-D, 1-D x 2-D, 2-D x 1-D, batched).
* complex64 / complex128: 1-D x 1-D is the unconjugated inner product
  sum(a*b) as a complex scalar; 2-D x 2-D is matmul_complex.
* integers: numpy keeps the input dtype and wraps (the reference delegates
  to numpy for that).  Here int8, uint8 and int16 operands, after numpy's
  type promotion, run on the INT8 / INT16 matrix-core kernels.  The exact
  product mod 2^32 reduced to the input width is numpy's wrapped result.
  Integer dtypes without a device kernel (int32, int64, bool) raise
  TypeError; nothing is computed on the host.

Device operands (ferray_amd.device.DeviceArray) are forwarded without a host
round trip.  Both operands must be DeviceArrays of one dtype (there are no
implicit device casts).  The product is enqueued on the calling thread's
library stream and returned as a DeviceArray: float and complex keep their
dtype, and int8 / int16 products are the exact wrapping int32.  Other
integer dtypes raise TypeError on the device.  Shapes follow the host rules,
and 1-D x 1-D gives a 0-d DeviceArray.
"""
from __future__ import annotations

import numpy as np

from . import products as _p
from ._lib import DTYPE_C32, DTYPE_C64, DTYPE_F32, DTYPE_F64, FerrayError, lib
from .device import DeviceArray

__all__ = ["matmul"]

_FLOAT = (np.dtype(np.float32), np.dtype(np.float64))
_COMPLEX = (np.dtype(np.complex64), np.dtype(np.complex128))
_INT_KERNEL = (np.dtype(np.int8), np.dtype(np.uint8), np.dtype(np.int16))


def matmul(a, b):
    if isinstance(a, DeviceArray) or isinstance(b, DeviceArray):
        return _device_matmul(a, b)
    a = np.asarray(a)
    dt = a.dtype
    if dt in _COMPLEX:
        cdt = np.dtype(np.complex64) if dt == np.complex64 else np.dtype(np.complex128)
        a = a.astype(cdt, copy=False)
        b = np.asarray(b).astype(cdt, copy=False)
        if a.ndim == 1 and b.ndim == 1:
            return _p.dot(a, b)[()]
        return _p.matmul_complex(a, b)
    if dt.kind in "iub":
        return _int_matmul(a, np.asarray(b))
    if dt not in _FLOAT:
        raise TypeError(f"matmul: unsupported dtype {dt}")
    b = np.asarray(b).astype(dt, copy=False)
    if a.ndim == 1 and b.ndim == 1:
        return _p.dot(a, b)[()]
    return _p.matmul(a, b)


# ---------------------------------------------------------------------------
# Host integer operands: numpy's dtype rules on the integer kernels.
# ---------------------------------------------------------------------------

def _int_product_2d(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Exact wrapping int32 product of 2-D int8 / uint8 / int16 operands of
    one dtype.  uint8 x uint8 runs as uint8 x int8 with B reinterpreted
    (b and b - 256 agree mod 2^8, so the product does too)."""
    if a.dtype == np.uint8:
        return _p.matmul_int(a, b.view(np.int8))
    return _p.matmul_int(a, b)


def _int_matmul(a: np.ndarray, b: np.ndarray):
    out_dt = np.result_type(a, b)
    if out_dt not in _INT_KERNEL:
        raise TypeError(f"matmul: no device kernel for integer dtype {out_dt} "
                        "(int8, uint8 and int16 run on the matrix cores)")
    a = a.astype(out_dt, copy=False)
    b = b.astype(out_dt, copy=False)
    if a.ndim == 1 and b.ndim == 1:
        if a.shape[0] != b.shape[0]:
            raise _p.ShapeMismatch(f"dot: vectors have different lengths {a.shape[0]} and "
                                   f"{b.shape[0]}")
        return _int_product_2d(a.reshape(1, -1), b.reshape(-1, 1)).astype(out_dt)[0, 0]
    if a.ndim == 1 and b.ndim == 2:
        if a.shape[0] != b.shape[0]:
            raise _p.ShapeMismatch(f"matmul: shapes {a.shape} and {b.shape} not aligned")
        return _int_product_2d(a.reshape(1, -1), b).astype(out_dt).reshape(b.shape[1])
    if a.ndim == 2 and b.ndim == 1:
        if a.shape[1] != b.shape[0]:
            raise _p.ShapeMismatch(f"matmul: shapes {a.shape} and {b.shape} not aligned")
        return _int_product_2d(a, b.reshape(-1, 1)).astype(out_dt).reshape(a.shape[0])
    if a.ndim == 2 and b.ndim == 2:
        if a.shape[1] != b.shape[0]:
            raise _p.ShapeMismatch(f"matmul: inner dimensions don't match "
                                   f"({a.shape[0]}x{a.shape[1]} @ {b.shape[0]}x{b.shape[1]})")
        return _int_product_2d(a, b).astype(out_dt)
    if a.ndim < 2 or b.ndim < 2:
        raise _p.ShapeMismatch("matmul: need at least 2D arrays for batched matmul")
    if a.shape[-1] != b.shape[-2]:
        raise _p.ShapeMismatch(f"matmul: inner dimensions don't match "
                               f"({a.shape[-1]} != {b.shape[-2]})")
    batch_shape = _p._broadcast_batch(a.shape[:-2], b.shape[:-2])
    m, n = a.shape[-2], b.shape[-1]
    a3 = a.reshape((-1,) + a.shape[-2:])
    b3 = b.reshape((-1,) + b.shape[-2:])
    batch = int(np.prod(batch_shape, dtype=np.int64)) if batch_shape else 1
    out = np.zeros((batch, m, n), out_dt)
    for i in range(batch):  # products/mod.rs:625-650 indexing: i % batch of each side
        out[i] = _int_product_2d(a3[i % max(1, a3.shape[0])],
                                 b3[i % max(1, b3.shape[0])]).astype(out_dt)
    return out.reshape(batch_shape + (m, n))


# ---------------------------------------------------------------------------
# Device operands
# ---------------------------------------------------------------------------

_DEV_DT = {np.dtype(np.float64): DTYPE_F64, np.dtype(np.float32): DTYPE_F32,
           np.dtype(np.complex128): DTYPE_C64, np.dtype(np.complex64): DTYPE_C32}


def _check(status: int, what: str) -> None:
    if status != 0:
        raise FerrayError(status, what)


def _dev_gemm(dt: np.dtype, m: int, n: int, k: int, a: int, b: int, c: int) -> None:
    L = lib()
    if dt == np.float64:
        _check(L.ferray_gemm_f64_dev(m, n, k, 1.0, a, b, 0.0, c, None), "gemm_f64_dev")
    elif dt == np.float32:
        _check(L.ferray_gemm_f32_dev(m, n, k, 1.0, a, b, 0.0, c, None), "gemm_f32_dev")
    elif dt == np.complex128:
        _check(L.ferray_gemm_c64_dev(m, n, k, 1.0, 0.0, a, b, 0.0, 0.0, c, None), "gemm_c64_dev")
    elif dt == np.complex64:
        _check(L.ferray_gemm_c32_dev(m, n, k, 1.0, 0.0, a, b, 0.0, 0.0, c, None), "gemm_c32_dev")
    elif dt == np.int8:
        _check(L.ferray_gemm_i8_signed_dev(m, n, k, a, b, c, False, None), "gemm_i8_signed_dev")
    elif dt == np.int16:
        _check(L.ferray_gemm_i16_dev(m, n, k, a, b, c, False, None), "gemm_i16_dev")
    else:
        raise TypeError(f"matmul: no device kernel for {dt}")


def _device_matmul(a, b) -> DeviceArray:
    if not (isinstance(a, DeviceArray) and isinstance(b, DeviceArray)):
        raise TypeError("matmul: device and host operands cannot be mixed")
    if a.dtype != b.dtype:
        raise TypeError(f"matmul: device operands of one dtype required ({a.dtype}, {b.dtype})")
    dt = a.dtype
    integer = dt in (np.dtype(np.int8), np.dtype(np.int16))
    if dt not in _DEV_DT and not integer:
        raise TypeError(f"matmul: no device kernel for {dt}")
    out_dt = np.dtype(np.int32) if integer else dt
    if a.ndim == 1 and b.ndim == 1:
        if a.shape[0] != b.shape[0]:
            raise _p.ShapeMismatch(f"dot: vectors have different lengths {a.shape[0]} and "
                                   f"{b.shape[0]}")
        m, k, n, out_shape = 1, a.shape[0], 1, ()
    elif a.ndim == 1 and b.ndim == 2:
        if a.shape[0] != b.shape[0]:
            raise _p.ShapeMismatch(f"matmul: shapes {a.shape} and {b.shape} not aligned")
        m, k, n, out_shape = 1, a.shape[0], b.shape[1], (b.shape[1],)
    elif a.ndim == 2 and b.ndim == 1:
        if a.shape[1] != b.shape[0]:
            raise _p.ShapeMismatch(f"matmul: shapes {a.shape} and {b.shape} not aligned")
        m, k, n, out_shape = a.shape[0], a.shape[1], 1, (a.shape[0],)
    elif a.ndim == 2 and b.ndim == 2:
        if a.shape[1] != b.shape[0]:
            raise _p.ShapeMismatch(f"matmul: inner dimensions don't match "
                                   f"({a.shape[0]}x{a.shape[1]} @ {b.shape[0]}x{b.shape[1]})")
        m, k, n, out_shape = a.shape[0], a.shape[1], b.shape[1], (a.shape[0], b.shape[1])
    else:
        return _device_batched(a, b, out_dt)
    out = DeviceArray(out_shape, out_dt)
    if m and n:
        _dev_gemm(dt, m, n, k, a.ptr, b.ptr, out.ptr)
    return out


def _device_batched(a: DeviceArray, b: DeviceArray, out_dt: np.dtype) -> DeviceArray:
    if a.ndim < 2 or b.ndim < 2:
        raise _p.ShapeMismatch("matmul: need at least 2D arrays for batched matmul")
    m, ka = a.shape[-2:]
    kb, n = b.shape[-2:]
    if ka != kb:
        raise _p.ShapeMismatch(f"matmul: inner dimensions don't match ({ka} != {kb})")
    if a.dtype not in _DEV_DT:
        raise TypeError(f"matmul: batched device products take f32, f64, c64 or c128, "
                        f"not {a.dtype}")
    batch_shape = _p._broadcast_batch(a.shape[:-2], b.shape[:-2])
    batch = int(np.prod(batch_shape, dtype=np.int64)) if batch_shape else 1
    a_batch = max(1, int(np.prod(a.shape[:-2], dtype=np.int64)))
    b_batch = max(1, int(np.prod(b.shape[:-2], dtype=np.int64)))
    out = DeviceArray(batch_shape + (m, n), out_dt)
    if batch == 0 or m == 0 or n == 0:
        return out
    if ka == 0:  # empty contraction: zeros
        out.buffer.upload(np.zeros(out.size, out_dt))
        return out
    _check(lib().ferray_matmul_batched_dev(_DEV_DT[a.dtype], batch, a_batch, b_batch, m, ka, n,
                                           a.ptr, b.ptr, out.ptr, None), "matmul_batched_dev")
    return out
