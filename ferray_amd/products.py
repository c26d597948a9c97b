"""ferray-linalg's product layer on the GPU GEMM family.

Host-side mirror of the functions that route into the GEMM entry points in
the reference (ferray-linalg/src/products/mod.rs, complex.rs, gemm/mod.rs),
with the same shape rules and errors, on numpy arrays:

    matmul(a, b)          products/mod.rs:362-437  (2-D, 1-D x 2-D, 2-D x 1-D, batched)
    dot(a, b)             products/mod.rs:62-137   (1-D . 1-D, 2-D, N-D x 1-D, tensordot)
    matmul_raw(a, b, m, k, n)   products/mod.rs:462-577 (flat row-major product)
    matmul_batched(a, b)  products/mod.rs:579-659  (batch dims broadcast, i % batch indexing)
    matmul_complex(a, b)  complex.rs:136-264
    gemm(a, b)            gemm/mod.rs:1139-1184    (BLAS-style A @ B of 2-D arrays)

Every product runs on the device: the reference's size thresholds choose
between its CPU kernels (naive loop below 9, faer, hand-tuned above 255);
here each call is one GEMM (or one grouped launch for a batch) whatever the
size, and a failing device call raises instead of falling back.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import gemm as _g
from ._lib import DTYPE_C32, DTYPE_C64, DTYPE_F32, DTYPE_F64, FerrayError, lib

__all__ = ["ShapeMismatch", "matmul", "dot", "matmul_raw", "matmul_batched", "matmul_complex",
           "gemm"]

_DT = {np.dtype(np.float64): DTYPE_F64, np.dtype(np.float32): DTYPE_F32,
       np.dtype(np.complex128): DTYPE_C64, np.dtype(np.complex64): DTYPE_C32}


class ShapeMismatch(ValueError):
    """FerrayError::ShapeMismatch of the reference."""


def _real_or_complex(a: np.ndarray, b: np.ndarray):
    if a.dtype != b.dtype or a.dtype not in _DT:
        raise TypeError(f"unsupported dtypes {a.dtype}, {b.dtype} (f32, f64, c64, c128)")
    return a.dtype


def _raw(a: np.ndarray, b: np.ndarray, m: int, k: int, n: int) -> np.ndarray:
    """C = A @ B for flat row-major operands; fresh flat result of m*n."""
    dt = _real_or_complex(a, b)
    a = np.ascontiguousarray(a).ravel()
    b = np.ascontiguousarray(b).ravel()
    if a.size != m * k or b.size != k * n:
        raise ShapeMismatch(f"matmul_raw: operand sizes {a.size}, {b.size} != {m}x{k}, {k}x{n}")
    c = np.zeros(m * n, dtype=dt)
    if m == 0 or n == 0 or k == 0:
        return c
    if dt == np.float64:
        ok = _g.gemm_f64(m, n, k, 1.0, a, b, 0.0, c)
    elif dt == np.float32:
        ok = _g.gemm_f32(m, n, k, np.float32(1.0), a, b, np.float32(0.0), c)
    elif dt == np.complex128:
        ok = _g.gemm_c64(m, n, k, 1.0, 0.0, a, b, 0.0, 0.0, c)
    else:
        ok = _g.gemm_c32(m, n, k, 1.0, 0.0, a, b, 0.0, 0.0, c)
    if not ok:
        raise FerrayError(_g.last_status(), "matmul_raw")
    return c


def matmul_raw(a, b, m: int, k: int, n: int) -> np.ndarray:
    return _raw(np.asarray(a), np.asarray(b), m, k, n)


def _broadcast_batch(a_batch, b_batch):
    """products/mod.rs broadcast_shapes: right-aligned, 1 stretches."""
    n = max(len(a_batch), len(b_batch))
    a_full = (1,) * (n - len(a_batch)) + tuple(a_batch)
    b_full = (1,) * (n - len(b_batch)) + tuple(b_batch)
    out = []
    for da, db in zip(a_full, b_full):
        if da == db or db == 1:
            out.append(da)
        elif da == 1:
            out.append(db)
        else:
            raise ShapeMismatch(f"matmul: batch shapes {tuple(a_batch)} and {tuple(b_batch)} "
                                "cannot be broadcast")
    return tuple(out)


def matmul_batched(a, b) -> np.ndarray:
    """Leading dims broadcast; matrix i uses A[i % a_batch], B[i % b_batch]
    (products/mod.rs:625-650), all problems in one grouped launch."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.ndim < 2 or b.ndim < 2:
        raise ShapeMismatch("matmul: need at least 2D arrays for batched matmul")
    m, ka = a.shape[-2:]
    kb, n = b.shape[-2:]
    if ka != kb:
        raise ShapeMismatch(f"matmul: inner dimensions don't match ({ka} != {kb})")
    dt = _real_or_complex(a, b)
    batch_shape = _broadcast_batch(a.shape[:-2], b.shape[:-2])
    batch = int(np.prod(batch_shape, dtype=np.int64)) if batch_shape else 1
    a_batch = max(1, int(np.prod(a.shape[:-2], dtype=np.int64)))
    b_batch = max(1, int(np.prod(b.shape[:-2], dtype=np.int64)))
    out = np.zeros(batch_shape + (m, n), dtype=dt)
    if batch == 0 or m == 0 or n == 0:
        return out
    if ka == 0:
        return out
    a_c = np.ascontiguousarray(a)
    b_c = np.ascontiguousarray(b)
    ok = lib().ferray_matmul_batched(_DT[dt], batch, a_batch, b_batch, m, ka, n,
                                     a_c.ctypes.data, b_c.ctypes.data, out.ctypes.data)
    if not ok:
        raise FerrayError(_g.last_status(), "matmul_batched")
    return out


def matmul(a, b) -> np.ndarray:
    a = np.asarray(a)
    b = np.asarray(b)
    if a.ndim == 1 and b.ndim == 1:
        raise ShapeMismatch("matmul: cannot multiply two 1D arrays (use dot instead)")
    if a.ndim == 1 and b.ndim == 2:
        k, n = a.shape[0], b.shape[1]
        if k != b.shape[0]:
            raise ShapeMismatch(f"matmul: shapes ({k},) and ({b.shape[0]},{n}) not aligned")
        return _raw(a, b, 1, k, n).reshape(n)
    if a.ndim == 2 and b.ndim == 1:
        m, k = a.shape
        if k != b.shape[0]:
            raise ShapeMismatch(f"matmul: shapes ({m},{k}) and ({b.shape[0]},) not aligned")
        return _raw(a, b, m, k, 1).reshape(m)
    if a.ndim == 2 and b.ndim == 2:
        (m, k1), (k2, n) = a.shape, b.shape
        if k1 != k2:
            raise ShapeMismatch(f"matmul: inner dimensions don't match ({m}x{k1} @ {k2}x{n})")
        return _raw(a, b, m, k1, n).reshape(m, n)
    return matmul_batched(a, b)


def dot(a, b) -> np.ndarray:
    a = np.asarray(a)
    b = np.asarray(b)
    if a.ndim == 1 and b.ndim == 1:
        if a.shape[0] != b.shape[0]:
            raise ShapeMismatch(f"dot: vectors have different lengths {a.shape[0]} and "
                                f"{b.shape[0]}")
        return _raw(a, b, 1, a.shape[0], 1).reshape(())
    if a.ndim == 2 and b.ndim == 2:
        (m, k1), (k2, n) = a.shape, b.shape
        if k1 != k2:
            raise ShapeMismatch(f"dot: inner dimensions don't match ({m}x{k1} @ {k2}x{n})")
        return _raw(a, b, m, k1, n).reshape(m, n)
    if b.ndim == 1:
        k = a.shape[-1]
        if k != b.shape[0]:
            raise ShapeMismatch(f"dot: last axis of a ({k}) != length of b ({b.shape[0]})")
        out_shape = a.shape[:-1]
        rows = int(np.prod(out_shape, dtype=np.int64)) if out_shape else 1
        res = _raw(a, b, rows, k, 1)
        return res.reshape(out_shape if out_shape else (1,))
    # General case: tensordot over one axis (the last of a with the first of
    # b), products/mod.rs:131-134.
    k = a.shape[-1]
    if k != b.shape[0]:
        raise ShapeMismatch(f"tensordot: axis lengths {k} and {b.shape[0]} differ")
    rows = int(np.prod(a.shape[:-1], dtype=np.int64))
    cols = int(np.prod(b.shape[1:], dtype=np.int64))
    return _raw(a, b, rows, k, cols).reshape(a.shape[:-1] + b.shape[1:])


def matmul_complex(a, b) -> np.ndarray:
    a = np.asarray(a)
    b = np.asarray(b)
    if a.ndim != 2 or b.ndim != 2:
        raise ShapeMismatch("matmul_complex: 2-D operands required")
    if a.dtype not in (np.complex64, np.complex128) or a.dtype != b.dtype:
        raise TypeError("matmul_complex: complex64 or complex128 operands of one dtype")
    (m, k1), (k2, n) = a.shape, b.shape
    if k1 != k2:
        raise ShapeMismatch(f"matmul_complex: inner dimensions don't match ({m}x{k1} @ {k2}x{n})")
    return _raw(a, b, m, k1, n).reshape(m, n)


def gemm(a, b) -> np.ndarray:
    """BLAS-style C = A @ B of 2-D, row-major contiguous arrays."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.ndim != 2 or b.ndim != 2:
        raise ShapeMismatch("gemm: 2-D operands required")
    (m, k1), (k2, n) = a.shape, b.shape
    if k1 != k2:
        raise ShapeMismatch(f"gemm: inner dimensions don't match ({m}x{k1} @ {k2}x{n})")
    if a.dtype not in (np.float32, np.float64) or a.dtype != b.dtype:
        raise TypeError("gemm: f32 or f64 operands of one dtype")
    if m == 0 or n == 0 or k1 == 0:
        return np.zeros((m, n), dtype=a.dtype)
    if not a.flags.c_contiguous:
        raise ShapeMismatch("gemm: input `a` must be contiguous (row-major)")
    if not b.flags.c_contiguous:
        raise ShapeMismatch("gemm: input `b` must be contiguous (row-major)")
    return _raw(a, b, m, k1, n).reshape(m, n)
