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
    matmul_int(a, b)      integer operands -> exact wrapping int32
    matrix_power(a, n)    solve.rs:456-499         (repeated squaring, device resident)
    multi_dot(arrays)     products/mod.rs:776-904  (matrix-chain order)
    einsum_matmul(s, a, b) products/einsum/mod.rs:251-281 (matmul-shaped einsum)

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
           "gemm", "matmul_int", "matrix_power", "multi_dot", "einsum_matmul"]

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


# ---------------------------------------------------------------------------
# Integer products (the typed integer entry the reference lacks: its matmul is
# sealed to f32/f64, scalar.rs:10-14).  i8 x i8, u8 x i8 and i16 x i16 give
# exact wrapping int32, the semantics of gemm_i8_signed / gemm_i8 / gemm_i16.
# ---------------------------------------------------------------------------

def _int_kind(a: np.ndarray, b: np.ndarray):
    kinds = {(np.dtype(np.int8), np.dtype(np.int8)): "i8",
             (np.dtype(np.uint8), np.dtype(np.int8)): "u8",
             (np.dtype(np.int16), np.dtype(np.int16)): "i16"}
    return kinds.get((a.dtype, b.dtype))


def matmul_int(a, b) -> np.ndarray:
    """2-D integer product with an int32 result (exact, wrapping mod 2^32)."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.ndim != 2 or b.ndim != 2:
        raise ShapeMismatch("matmul_int: 2-D operands required")
    (m, k1), (k2, n) = a.shape, b.shape
    if k1 != k2:
        raise ShapeMismatch(f"matmul_int: inner dimensions don't match ({m}x{k1} @ {k2}x{n})")
    kind = _int_kind(a, b)
    if kind is None:
        raise TypeError("matmul_int: (int8, int8), (uint8, int8) or (int16, int16) operands")
    c = np.zeros(m * n, dtype=np.int32)
    if m == 0 or n == 0 or k1 == 0:
        return c.reshape(m, n)
    a = np.ascontiguousarray(a).ravel()
    b = np.ascontiguousarray(b).ravel()
    fn = {"i8": _g.gemm_i8_signed, "u8": _g.gemm_i8, "i16": _g.gemm_i16}[kind]
    if not fn(m, n, k1, a, b, c, False):
        raise FerrayError(_g.last_status(), "matmul_int")
    return c.reshape(m, n)


# ---------------------------------------------------------------------------
# GEMM consumers: matrix_power (solve.rs:456-499), multi_dot
# (products/mod.rs:776-904) and einsum's matmul shortcut
# (products/einsum/mod.rs:251-281).
# ---------------------------------------------------------------------------

def matrix_power(a, n: int) -> np.ndarray:
    """A^n by repeated squaring with the operands kept in device memory.

    n >= 0; negative powers need the inverse, which is not on this path."""
    a = np.asarray(a)
    if a.ndim != 2 or a.shape[0] != a.shape[1]:
        raise ShapeMismatch(f"matrix_power requires a square matrix, got {a.shape}")
    if a.dtype not in _DT:
        raise TypeError("matrix_power: f32, f64, c64 or c128 operand")
    if n < 0:
        raise ValueError("matrix_power: negative powers need the inverse, which is not provided")
    sz = a.shape[0]
    eye = np.eye(sz, dtype=a.dtype)
    if n == 0 or sz == 0:
        return eye
    from .gemm import DeviceBuffer, sync
    L = lib()
    dt = _DT[a.dtype]
    nbytes = a.nbytes
    res, base, tmp = DeviceBuffer.from_array(eye), DeviceBuffer.from_array(np.ascontiguousarray(a)), \
        DeviceBuffer(nbytes)

    def mm(x, y, out):
        if dt == DTYPE_F64:
            st = L.ferray_gemm_f64_dev(sz, sz, sz, 1.0, x.ptr, y.ptr, 0.0, out.ptr, None)
        elif dt == DTYPE_F32:
            st = L.ferray_gemm_f32_dev(sz, sz, sz, 1.0, x.ptr, y.ptr, 0.0, out.ptr, None)
        elif dt == DTYPE_C64:
            st = L.ferray_gemm_c64_dev(sz, sz, sz, 1.0, 0.0, x.ptr, y.ptr, 0.0, 0.0, out.ptr, None)
        else:
            st = L.ferray_gemm_c32_dev(sz, sz, sz, 1.0, 0.0, x.ptr, y.ptr, 0.0, 0.0, out.ptr, None)
        if st != 0:
            raise FerrayError(st, "matrix_power")

    try:
        p = n
        while p > 0:
            if p & 1:
                mm(res, base, tmp)
                res, tmp = tmp, res
            p >>= 1
            if p:
                mm(base, base, tmp)
                base, tmp = tmp, base
        sync()
        return res.download(a.dtype, sz * sz).reshape(sz, sz)
    finally:
        for buf in (res, base, tmp):
            buf.free()


def multi_dot(arrays) -> np.ndarray:
    """Chain product in the cost-optimal order (matrix-chain DP); the first
    and last operands may be 1-D."""
    arrays = [np.asarray(x) for x in arrays]
    if len(arrays) < 2:
        raise ValueError("multi_dot: need at least 2 matrices")
    if len(arrays) == 2:
        return matmul(arrays[0], arrays[1])
    n = len(arrays)
    dims = []
    for i, x in enumerate(arrays):
        if x.ndim != 2:
            if i == 0 and x.ndim == 1:
                dims += [1, x.shape[0]]
                continue
            if i == n - 1 and x.ndim == 1:
                dims += [x.shape[0], 1]
                continue
            raise ShapeMismatch(f"multi_dot: matrix {i} has {x.ndim} dimensions (expected 2)")
        if i == 0:
            dims.append(x.shape[0])
        dims.append(x.shape[1])
    # every 2-D operand contributed its column count, so dims has n + 1 entries
    dims = dims[:n + 1]
    for i in range(1, n):
        prev_cols = arrays[i - 1].shape[-1]
        if prev_cols != arrays[i].shape[0]:
            raise ShapeMismatch(f"multi_dot: shapes of matrices {i - 1} and {i} not aligned "
                                f"({prev_cols} != {arrays[i].shape[0]})")
    cost = [[0] * n for _ in range(n)]
    split = [[0] * n for _ in range(n)]
    for length in range(2, n + 1):
        for i in range(0, n - length + 1):
            j = i + length - 1
            cost[i][j] = None
            for k in range(i, j):
                q = cost[i][k] + cost[k + 1][j] + dims[i] * dims[k + 1] * dims[j + 1]
                if cost[i][j] is None or q < cost[i][j]:
                    cost[i][j] = q
                    split[i][j] = k

    def chain(i, j):
        if i == j:
            return arrays[i]
        k = split[i][j]
        left, right = chain(i, k), chain(k + 1, j)
        # a 1-D first and last operand can meet at the end: their product is
        # the inner product (numpy's multi_dot result), where matmul would
        # refuse two 1-D operands
        if left.ndim == 1 and right.ndim == 1:
            return dot(left, right)
        return matmul(left, right)

    return chain(0, n - 1)


def einsum_matmul(subscripts: str, a, b) -> np.ndarray:
    """The matmul-shaped two-operand einsum forms ('ij,jk->ik' with any
    labels, and the batched 'bij,bjk->bik'); other forms raise
    NotImplementedError (general einsum is not on this path)."""
    spec = subscripts.replace(" ", "")
    if "->" not in spec:
        raise NotImplementedError("einsum_matmul: explicit output required")
    ins, out = spec.split("->")
    if "," not in ins:
        raise NotImplementedError("einsum_matmul: two operands required")
    x, y = ins.split(",")
    a = np.asarray(a)
    b = np.asarray(b)
    if len(x) == 2 and len(y) == 2 and len(out) == 2 and len(set(x + y)) == 3:
        i, j = x
        j2, k = y
        if j == j2 and out == i + k and i != k:
            return matmul(a, b)
    if len(x) == 3 and len(y) == 3 and len(out) == 3 and x[0] == y[0] == out[0]:
        bl, i, j = x
        _, j2, k = y
        if j == j2 and out == bl + i + k and len({bl, i, j, k}) == 4:
            return matmul_batched(a, b)
    raise NotImplementedError(f"einsum_matmul: '{subscripts}' is not a matmul contraction")
