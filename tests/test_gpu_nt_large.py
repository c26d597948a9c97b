"""The integer, half and fp8 products at sizes that run the 256 x 256
direct-to-LDS kernel (at least 128 tiles of 256 x 256), with block tiles both
wholly inside C (the epilogue's unchecked form) and on its ragged border.

Integer results must be exact (wrapping mod 2^32); the references come from
float64 BLAS products, exact here because every partial sum stays below
2^53.  quantized_matmul's reference repeats the kernel's float32 steps
(scale_a * scale_b in f32, times the wrapped integer, plus C when
accumulating), so it is bit-exact too.  bf16 is held to the worst-case bound of
f32 accumulation (its products are exact in f32); fp8 with integer-valued codes is exact.
"""
from __future__ import annotations

import numpy as np
import pytest

import ferray_amd as fa
import oracle
from tests.helpers import rng, wrap_i32

pytestmark = pytest.mark.gpu

# (m, n, k): aligned (A read in place or flipped in one pass) and ragged
# (A packed, border tiles checked); both have >= 128 tiles of 256 x 256.
SHAPES = [(3072, 2816, 640), (3000, 2900, 700)]


def _same(c, want, n):
    """Exact equality; on a mismatch, name the count, the first element (row,
    column) and its got / wanted values."""
    bad = np.flatnonzero(c != want)
    assert bad.size == 0, (f"{bad.size} wrong, first at {divmod(int(bad[0]), n)}: "
                           f"got {c[bad[0]]}, want {want[bad[0]]}")


def _exact(a, b, m, n, k):
    return wrap_i32(np.rint(a.reshape(m, k).astype(np.float64) @
                            b.reshape(k, n).astype(np.float64)).astype(np.int64)).ravel()


@pytest.mark.parametrize("m,n,k", SHAPES)
def test_i8_signed_large(gpu, m, n, k):
    r = rng(301)
    a = r.integers(-128, 128, m * k).astype(np.int8)
    b = r.integers(-128, 128, k * n).astype(np.int8)
    ref = _exact(a, b, m, n, k)
    c = np.zeros(m * n, np.int32)
    assert gpu.gemm_i8_signed(m, n, k, a, b, c, False)
    _same(c, ref, n)
    c0 = r.integers(-2 ** 31, 2 ** 31, m * n).astype(np.int32)
    c = c0.copy()
    assert gpu.gemm_i8_signed(m, n, k, a, b, c, True)
    _same(c, wrap_i32(ref.astype(np.int64) + c0), n)


@pytest.mark.parametrize("m,n,k", SHAPES)
def test_u8_i8_large(gpu, m, n, k):
    r = rng(302)
    a = r.integers(0, 256, m * k).astype(np.uint8)
    b = r.integers(-128, 128, k * n).astype(np.int8)
    c = np.zeros(m * n, np.int32)
    assert gpu.gemm_i8(m, n, k, a, b, c, False)
    _same(c, _exact(a, b, m, n, k), n)


@pytest.mark.parametrize("m,n,k", SHAPES)
def test_i16_large(gpu, m, n, k):
    r = rng(303)
    a = r.integers(-32768, 32768, m * k).astype(np.int16)
    b = r.integers(-32768, 32768, k * n).astype(np.int16)
    ref = _exact(a, b, m, n, k)
    c = np.zeros(m * n, np.int32)
    assert gpu.gemm_i16(m, n, k, a, b, c, False)
    _same(c, ref, n)
    c0 = r.integers(-2 ** 31, 2 ** 31, m * n).astype(np.int32)
    c = c0.copy()
    assert gpu.gemm_i16(m, n, k, a, b, c, True)
    _same(c, wrap_i32(ref.astype(np.int64) + c0), n)


@pytest.mark.parametrize("m,n,k", SHAPES)
@pytest.mark.parametrize("acc", [False, True])
def test_quantized_large(gpu, m, n, k, acc):
    azp, sa, bzp, sb = 128, 0.0078, -3, 0.015
    r = rng(304)
    a = r.integers(0, 256, m * k).astype(np.uint8)
    b = r.integers(-128, 128, k * n).astype(np.int8)
    c0 = (r.standard_normal(m * n) * 100).astype(np.float32)
    corrected = wrap_i32(np.rint((a.reshape(m, k).astype(np.float64) - azp) @
                                 (b.reshape(k, n).astype(np.float64) - bzp)).astype(np.int64))
    out = np.float32(np.float32(sa) * np.float32(sb)) * corrected.ravel().astype(np.float32)
    ref = (c0 + out) if acc else out
    c = c0.copy()
    assert gpu.quantized_matmul(m, n, k, a, azp, sa, b, bzp, sb, c, acc)
    _same(c, ref.astype(np.float32), n)


@pytest.mark.parametrize("m,n,k", SHAPES)
def test_bf16_large(gpu, m, n, k):
    r = rng(305)
    af = r.standard_normal(m * k).astype(np.float32)
    bf = r.standard_normal(k * n).astype(np.float32)
    a = (af.view(np.uint32) >> 16).astype(np.uint16)
    b = (bf.view(np.uint32) >> 16).astype(np.uint16)
    A = (a.astype(np.uint32) << 16).view(np.float32).astype(np.float64).reshape(m, k)
    B = (b.astype(np.uint32) << 16).view(np.float32).astype(np.float64).reshape(k, n)
    for acc in (False, True):
        c0 = r.standard_normal(m * n).astype(np.float32)
        c = c0.copy()
        assert gpu.gemm_bf16_f32(m, n, k, a, b, c, acc)
        ref = (A @ B).ravel() + (c0.astype(np.float64) if acc else 0.0)
        # Products of bf16 values are exact in f32, so only the k + 1 f32
        # additions round: |error| <= (k + 2) u (|A| |B| + |C0|), u = 2^-24.
        mag = (np.abs(A) @ np.abs(B)).ravel() + (np.abs(c0) if acc else 0.0)
        err = np.abs(c - ref)
        bound = (k + 2) * 2.0 ** -24 * mag
        assert np.all(err <= bound), (int((err > bound).sum()), float(err.max()))


@pytest.mark.parametrize("m,n,k", SHAPES)
def test_f16_large(gpu, m, n, k):
    r = rng(307)
    a16 = r.standard_normal(m * k).astype(np.float16)
    b16 = r.standard_normal(k * n).astype(np.float16)
    a, b = a16.view(np.uint16), b16.view(np.uint16)
    A = a16.astype(np.float64).reshape(m, k)
    B = b16.astype(np.float64).reshape(k, n)
    for acc in (False, True):
        c0 = r.standard_normal(m * n).astype(np.float32)
        c = c0.copy()
        assert gpu.gemm_f16_f32(m, n, k, a, b, c, acc)
        ref = (A @ B).ravel() + (c0.astype(np.float64) if acc else 0.0)
        # f16 products are exact in f32 (11-bit significands), so the bound
        # is the one for bf16 above.
        mag = (np.abs(A) @ np.abs(B)).ravel() + (np.abs(c0) if acc else 0.0)
        err = np.abs(c - ref)
        bound = (k + 2) * 2.0 ** -24 * mag
        assert np.all(err <= bound), (int((err > bound).sum()), float(err.max()))


@pytest.mark.parametrize("m,n,k", SHAPES)
def test_e4m3_integer_valued_large(gpu, m, n, k):
    table = np.array([oracle.e4m3_to_f64(i) for i in range(256)])
    ok = np.flatnonzero(np.isfinite(table) & (table == np.rint(table)) & (np.abs(table) <= 8))
    r = rng(306)
    a = r.choice(ok, m * k).astype(np.uint8)
    b = r.choice(ok, k * n).astype(np.uint8)
    ref = (table[a].reshape(m, k) @ table[b].reshape(k, n)).ravel()
    c = np.zeros(m * n, np.float32)
    assert gpu.gemm_e4m3_f32(m, n, k, a, b, c, False)
    _same(c.astype(np.float64), ref, n)
