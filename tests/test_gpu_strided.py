"""Row-strided GEMM entries (the drivers' lda / ldb / ldc form, outer.rs:198-210).

Bars:
  * the strided result equals the tight entry's on the same operands copied
    out of their padded storage, bit for bit (same kernels, same k order),
    on every path: aligned 128 x 128 tiles, ragged shapes through the padded
    copies, unaligned strides, the 64 x 64 small-product tiles, k == 0;
  * elements of C outside its m x n block keep their values;
  * the result meets |C - C_oracle| <= 2(k + 2) u (|alpha||A||B| + |beta||C0|)
    against the oracle's tight restatement on the extracted operands;
  * a stride below its row length is refused.
"""
from __future__ import annotations

import numpy as np
import pytest

import ferray_amd as fa
import oracle
from ferray_amd._lib import lib
from tests.helpers import rng

pytestmark = pytest.mark.gpu

KINDS = {
    "f64": (np.float64, np.float64, 2.0 ** -53),
    "f32": (np.float32, np.float32, 2.0 ** -24),
    "c64": (np.complex128, np.float64, 2.0 ** -53),
    "c32": (np.complex64, np.float32, 2.0 ** -24),
}
# (m, n, k, lda, ldb, ldc): aligned 128 x 128 tiles (>= 256 tiles), ragged, odd
# strides, small, k == 0
SHAPES = [(2048, 2048, 128, 144, 2064, 2056), (300, 257, 130, 131, 300, 261),
          (513, 129, 257, 301, 133, 140), (40, 33, 17, 17, 40, 35), (64, 48, 0, 3, 48, 50)]


def _entry(kind):
    L = lib()
    return {"f64": L.ferray_gemm_f64_strided_dev, "f32": L.ferray_gemm_f32_strided_dev,
            "c64": L.ferray_gemm_c64_strided_dev, "c32": L.ferray_gemm_c32_strided_dev}[kind]


def _tight(kind):
    L = lib()
    return {"f64": L.ferray_gemm_f64_dev, "f32": L.ferray_gemm_f32_dev,
            "c64": L.ferray_gemm_c64_dev, "c32": L.ferray_gemm_c32_dev}[kind]


def _call(fn, kind, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    cplx = kind in ("c64", "c32")
    args = [m, n, k]
    args += [alpha.real, alpha.imag] if cplx else [alpha.real]
    args += [a] if lda is None else [a, lda]
    args += [b] if ldb is None else [b, ldb]
    args += [beta.real, beta.imag] if cplx else [beta.real]
    args += [c] if ldc is None else [c, ldc]
    return fn(*args, None)


def _rand(r, shape, dt):
    x = r.standard_normal(shape)
    if np.issubdtype(dt, np.complexfloating):
        x = x + 1j * r.standard_normal(shape)
    return x.astype(dt)


@pytest.mark.parametrize("kind", list(KINDS))
@pytest.mark.parametrize("shape", SHAPES)
def test_strided_equals_tight_and_oracle(gpu, kind, shape):
    m, n, k, lda, ldb, ldc = shape
    dt, real, u = KINDS[kind]
    r = rng(sum(shape))
    Ap = _rand(r, (m, lda), dt)
    Bp = _rand(r, (max(k, 1), ldb), dt)
    Cp = _rand(r, (m, ldc), dt)
    A = np.ascontiguousarray(Ap[:, :k]); B = np.ascontiguousarray(Bp[:k, :n])
    C0 = np.ascontiguousarray(Cp[:, :n])
    alpha, beta = (dt(0.75 - 0.5j), dt(-1.25 + 0.25j)) if kind[0] == "c" else (dt(0.75), dt(-1.25))
    dA, dB, dC = (fa.DeviceBuffer.from_array(x) for x in (Ap, Bp, Cp))
    st = _call(_entry(kind), kind, m, n, k, alpha, dA.ptr, lda, dB.ptr, ldb, beta, dC.ptr, ldc)
    assert st == 0, lib().ferray_status_string(st).decode()
    fa.sync()
    got = dC.download(dt, m * ldc).reshape(m, ldc)
    for d in (dA, dB, dC):
        d.free()
    # padding of C untouched
    assert got[:, n:].tobytes() == Cp[:, n:].tobytes()
    # tight entry on the extracted operands: same bits
    tA, tB, tC = (fa.DeviceBuffer.from_array(x) for x in (A if k else np.zeros(1, dt),
                                                          B if k else np.zeros(1, dt), C0))
    st = _call(_tight(kind), kind, m, n, k, alpha, tA.ptr, None, tB.ptr, None, beta, tC.ptr, None)
    assert st == 0
    fa.sync()
    tight = tC.download(dt, m * n).reshape(m, n)
    for d in (tA, tB, tC):
        d.free()
    assert got[:, :n].tobytes() == tight.tobytes()
    # the oracle on the extracted operands
    ref = C0.copy()
    if kind == "f64":
        oracle.gemm_f64(m, n, k, alpha, A.ravel(), B.ravel(), beta, ref.ravel())
    elif kind == "f32":
        oracle.gemm_f32(m, n, k, alpha, A.ravel(), B.ravel(), beta, ref.ravel())
    elif kind == "c64":
        oracle.gemm_c64(m, n, k, alpha.real, alpha.imag, A.view(np.float64).ravel(),
                        B.view(np.float64).ravel(), beta.real, beta.imag,
                        ref.view(np.float64).ravel())
    else:
        oracle.gemm_c32(m, n, k, alpha.real, alpha.imag, A.view(np.float32).ravel(),
                        B.view(np.float32).ravel(), beta.real, beta.imag,
                        ref.view(np.float32).ravel())
    mag = abs(complex(alpha)) * (np.abs(A).astype(np.float64) @ np.abs(B).astype(np.float64)) \
        if k else np.zeros((m, n))
    bound = 2 * 2 * (k + 2) * u * (mag + abs(complex(beta)) * np.abs(C0))
    assert np.all(np.abs(got[:, :n].astype(np.complex128) - ref.astype(np.complex128)) <= bound)


def test_strided_rejects_short_strides(gpu):
    L = lib()
    d = fa.DeviceBuffer(1 << 16)
    assert L.ferray_gemm_f64_strided_dev(8, 8, 8, 1.0, d.ptr, 7, d.ptr, 8, 0.0, d.ptr, 8, None) == 3
    assert L.ferray_gemm_f64_strided_dev(8, 8, 8, 1.0, d.ptr, 8, d.ptr, 7, 0.0, d.ptr, 8, None) == 3
    assert L.ferray_gemm_f32_strided_dev(8, 8, 8, 1.0, d.ptr, 8, d.ptr, 8, 0.0, d.ptr, 7, None) == 3
    d.free()
