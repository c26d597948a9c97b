"""Real and complex GEMMs whose sizes are not whole 128 x 128 tiles (or whose
operands are not 16-byte aligned) on the 128 x 128 kernel: A, B (and C when
beta != 0) are copied into zero-padded buffers of whole tiles, the unguarded
kernel runs there, and C's corner is copied back.  The padding only appends
zero products after each element's own k-sequence, so the result must equal,
bit for bit, the same problem embedded in an aligned one with explicit zero
padding (which runs the unguarded kernel directly); and it must be right
(checked against numpy in extended precision)."""
from __future__ import annotations

import numpy as np
import pytest

import ferray_amd as fa
from ferray_amd._lib import lib
from tests.helpers import rng

pytestmark = pytest.mark.gpu

DEV = {np.float64: "ferray_gemm_f64_dev", np.float32: "ferray_gemm_f32_dev",
       np.complex128: "ferray_gemm_c64_dev", np.complex64: "ferray_gemm_c32_dev"}


def _dev_gemm(dt, m, n, k, alpha, a, b, beta, c0, offset=0):
    """C = alpha A B + beta C on the device; `offset` elements of slack before
    each operand so its address is not 16-byte aligned."""
    bufs = []
    for arr in (a, b, c0):
        flat = np.zeros(arr.size + offset, dt)
        flat[offset:] = arr.ravel()
        bufs.append(fa.DeviceBuffer.from_array(flat))
    es = np.dtype(dt).itemsize
    pa, pb, pc = (buf.ptr + offset * es for buf in bufs)
    fn = getattr(lib(), DEV[dt])
    if dt in (np.complex128, np.complex64):
        st = fn(m, n, k, alpha.real, alpha.imag, pa, pb, beta.real, beta.imag, pc, None)
    else:
        st = fn(m, n, k, alpha, pa, pb, beta, pc, None)
    assert st == 0
    return bufs[2].download(dt, c0.size + offset)[offset:].reshape(m, n)


def _operands(dt, m, n, k, seed):
    r = rng(seed)
    cplx = dt in (np.complex128, np.complex64)

    def draw(*shape):
        x = r.standard_normal(shape)
        if cplx:
            x = x + 1j * r.standard_normal(shape)
        return x.astype(dt)
    return draw(m, k), draw(k, n), draw(m, n)


def _embedded(dt, m, n, k, alpha, a, b, beta, c0):
    mp, kp = -(-m // 128) * 128, -(-k // 16) * 16
    np_ = -(-n // 128) * 128 if dt in (np.float64, np.float32) else -(-n // 64) * 64
    if dt in (np.complex128, np.complex64):
        kp = -(-k // 8) * 8
    ap = np.zeros((mp, kp), dt); ap[:m, :k] = a
    bp = np.zeros((kp, np_), dt); bp[:k, :n] = b
    cp = np.zeros((mp, np_), dt); cp[:m, :n] = c0
    return _dev_gemm(dt, mp, np_, kp, alpha, ap, bp, beta, cp)[:m, :n]


CASES = [(np.float64, 2000, 2040, 1999), (np.float32, 2000, 2040, 1999),
         (np.complex128, 2000, 1020, 999), (np.complex64, 2000, 1020, 999)]


@pytest.mark.parametrize("dt,m,n,k", CASES)
@pytest.mark.parametrize("beta", [0.0, 0.5])
def test_ragged_equals_embedded_aligned(dt, m, n, k, beta):
    a, b, c0 = _operands(dt, m, n, k, 7)
    alpha = dt(1.5 - 0.25j) if dt in (np.complex128, np.complex64) else dt(1.5)
    beta = dt(beta)
    got = _dev_gemm(dt, m, n, k, alpha, a, b, beta, c0)
    assert np.array_equal(got, _embedded(dt, m, n, k, alpha, a, b, beta, c0))
    wide = np.complex128 if np.iscomplexobj(a) else np.float64
    ref = complex(alpha) * (a.astype(wide) @ b.astype(wide)) + complex(beta) * c0.astype(wide)
    if wide is np.float64:
        ref = ref.real
    mag = abs(alpha) * (np.abs(a).astype(np.float64) @ np.abs(b).astype(np.float64)) + \
        abs(beta) * np.abs(c0)
    # the device's bound 2(k+2)u|.|, plus the same again for the f64 reference
    u = np.finfo(dt).eps / 2
    assert np.all(np.abs(got - ref) <= 8 * (k + 2) * u * mag)


@pytest.mark.parametrize("dt", [np.float64, np.complex64])
def test_unaligned_operands_equal_aligned(dt):
    m, n, k = (2048, 2048, 2048) if dt == np.float64 else (2048, 1024, 1024)
    a, b, c0 = _operands(dt, m, n, k, 8)
    alpha = dt(1.0)
    beta = dt(0.0)
    aligned = _dev_gemm(dt, m, n, k, alpha, a, b, beta, c0)
    shifted = _dev_gemm(dt, m, n, k, alpha, a, b, beta, c0, offset=1)
    assert np.array_equal(aligned, shifted)


# Products that leave CUs idle on 128 x 128 tiles run on 64 x 64 or, below
# 192 of those, 32 x 32 tiles; every form accumulates each element over the
# same MFMA k-sequence.  The same rows of A stacked under 8192 - m more rows
# make a product on the 128 x 128 kernel, whose first m rows must match.
@pytest.mark.parametrize("dt,m,n,k", [
    (np.float64, 512, 512, 512), (np.float64, 768, 768, 768), (np.float64, 300, 200, 100),
    (np.float64, 900, 896, 900), (np.float32, 512, 512, 512), (np.float32, 130, 1024, 77),
    (np.complex128, 384, 384, 384), (np.complex64, 256, 512, 256)])
def test_small_tile_forms_equal_the_large_kernel(gpu, dt, m, n, k):
    r = rng(m * 7 + n)
    cplx = np.dtype(dt).kind == "c"

    def draw(*shape):
        x = r.standard_normal(shape)
        if cplx:
            x = x + 1j * r.standard_normal(shape)
        return x.astype(dt)

    a, b = draw(m, k), draw(k, n)
    big = 8192
    a_big = np.concatenate([a, draw(big - m, k)])
    one, zero = dt(1), dt(0)
    c_small = _dev_gemm(dt, m, n, k, one, a, b, zero, np.zeros((m, n), dt))
    c_big = _dev_gemm(dt, big, n, k, one, a_big, b, zero, np.zeros((big, n), dt))
    assert np.array_equal(c_small.view(np.uint8), c_big[:m].view(np.uint8))
