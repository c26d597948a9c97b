"""fp8 x fp8 -> f32 (OCP E4M3 and E5M2 bytes) on the f8f6f4 matrix instruction,
through the C ABI, against the oracle's exact product.

Bars:
  * integer-valued operands (|x| <= 8, so every partial sum is an integer far
    below 2^24): bit-exact to the oracle whatever the summation order, which
    pins indexing, the k pairing of the two operands, padding and accumulate;
  * subnormal operands (exponent field 0) against 1.0: exact, so denormal
    inputs are neither flushed nor rounded;
  * operands whose products stay within a factor 16 of each other (values in
    [0.5, 2]): bit-exact, like the integer case;
  * every finite code at random (products spanning ~2^36): the instruction
    sums each group of 64 products in a fixed-point window tied to the
    group's largest product (measured on MI355X: an unbiased error of at most
    2^-11.2 of that product per group), so the bar is
        |C_gpu - C| <= 2^-10 G + 2 (k/64 + 2) 2^-24 (|A||B| + |C0|),
    G_ij = sum over 64-groups g of max_{p in g} |a_ip b_pj|, elementwise
    against the double-summed product, and the normwise error
    ||C_gpu - C||_F / (||A||_F ||B||_F) is reported.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from tests.helpers import rng

pytestmark = pytest.mark.gpu
FORMATS = ["e4m3", "e5m2"]


def _lut(fmt):
    dec = oracle.e4m3_to_f64 if fmt == "e4m3" else oracle.e5m2_to_f64
    return np.array([dec(i) for i in range(256)])


def _codes(fmt, pred):
    lut = _lut(fmt)
    ok = np.isfinite(lut) & pred(lut)
    return np.flatnonzero(ok).astype(np.uint8)


def _run(gpu, fmt, m, n, k, a, b, c, accumulate):
    fn = gpu.gemm_e4m3_f32 if fmt == "e4m3" else gpu.gemm_e5m2_f32
    assert fn(m, n, k, a, b, c, accumulate), gpu.last_status()


def _exact(fmt, m, n, k, a, b, c0=None):
    c = np.zeros(m * n) if c0 is None else c0.astype(np.float64).copy()
    oracle.gemm_fp8_f64(m, n, k, a, b, fmt, c, c0 is not None)
    return c


SHAPES = [(1, 1, 1), (7, 11, 9), (33, 65, 130), (64, 64, 64), (256, 256, 256), (300, 257, 129),
          (512, 384, 1024), (130, 1000, 77)]


@pytest.mark.parametrize("fmt", FORMATS)
@pytest.mark.parametrize("m,n,k", SHAPES)
def test_integer_values_bit_exact(gpu, fmt, m, n, k):
    r = rng(m * 7 + n * 3 + k)
    codes = _codes(fmt, lambda v: (np.abs(v) <= 8) & (v == np.round(v)))
    a = r.choice(codes, m * k)
    b = r.choice(codes, k * n)
    c = np.full(m * n, np.nan, np.float32)
    _run(gpu, fmt, m, n, k, a, b, c, False)
    assert np.array_equal(c, _exact(fmt, m, n, k, a, b).astype(np.float32))
    c0 = r.integers(-1000, 1000, m * n).astype(np.float32)
    c = c0.copy()
    _run(gpu, fmt, m, n, k, a, b, c, True)
    assert np.array_equal(c, _exact(fmt, m, n, k, a, b, c0).astype(np.float32))


@pytest.mark.parametrize("fmt", FORMATS)
def test_subnormal_inputs_are_exact(gpu, fmt):
    m, n, k = 64, 96, 256
    r = rng(11)
    # positive subnormals only, so no sum cancels to zero
    sub = _codes(fmt, lambda v: (v > 0) & (v < (2.0 ** -6 if fmt == "e4m3" else 2.0 ** -14)))
    one = 0x38 if fmt == "e4m3" else 0x3C
    a = r.choice(sub, m * k)
    b = np.full(k * n, one, np.uint8)
    c = np.zeros(m * n, np.float32)
    _run(gpu, fmt, m, n, k, a, b, c, False)
    ref = _exact(fmt, m, n, k, a, b)
    assert np.array_equal(c, ref.astype(np.float32))
    assert np.all(ref != 0)


@pytest.mark.parametrize("fmt", FORMATS)
@pytest.mark.parametrize("m,n,k", [(64, 96, 64), (300, 257, 129), (512, 384, 1024)])
def test_narrow_range_values_bit_exact(gpu, fmt, m, n, k):
    r = rng(3 * m + k)
    codes = _codes(fmt, lambda v: (np.abs(v) >= 0.5) & (np.abs(v) <= 2))
    a = r.choice(codes, m * k)
    b = r.choice(codes, k * n)
    c = np.zeros(m * n, np.float32)
    _run(gpu, fmt, m, n, k, a, b, c, False)
    assert np.array_equal(c, _exact(fmt, m, n, k, a, b).astype(np.float32))


def _group_max(A, B):
    """G_ij = sum over 64-wide k groups of max_p |a_ip b_pj|."""
    m, k = A.shape
    G = np.zeros((m, B.shape[1]))
    for g in range(0, k, 64):
        Ag, Bg = np.abs(A[:, g:g + 64]), np.abs(B[g:g + 64])
        for r0 in range(0, m, 64):
            G[r0:r0 + 64] += (Ag[r0:r0 + 64, :, None] * Bg[None]).max(axis=1)
    return G


@pytest.mark.parametrize("fmt", FORMATS)
@pytest.mark.parametrize("m,n,k", [(97, 131, 300), (256, 512, 640), (512, 512, 1024)])
def test_all_finite_codes_within_stated_bound(gpu, fmt, m, n, k):
    r = rng(m + n + k)
    codes = _codes(fmt, lambda v: np.ones_like(v, bool))
    a = r.choice(codes, m * k)
    b = r.choice(codes, k * n)
    lut = _lut(fmt)
    A, B = lut[a].reshape(m, k), lut[b].reshape(k, n)
    mag = np.abs(A) @ np.abs(B)
    G = _group_max(A, B)
    c0 = r.standard_normal(m * n).astype(np.float32)
    for accumulate in (False, True):
        c = c0.copy()
        _run(gpu, fmt, m, n, k, a, b, c, accumulate)
        ref = _exact(fmt, m, n, k, a, b, c0 if accumulate else None).reshape(m, n)
        c0m = np.abs(c0.reshape(m, n)) if accumulate else 0
        bound = 2.0 ** -10 * G + 2 * (k / 64 + 2) * 2.0 ** -24 * (mag + c0m)
        err = np.abs(c.reshape(m, n).astype(np.float64) - ref)
        assert np.all(err <= bound), float(np.max(err / bound))
        normwise = np.linalg.norm(err) / (np.linalg.norm(A) * np.linalg.norm(B))
        print(f"{fmt} {m}x{n}x{k} accumulate={accumulate}: normwise {normwise:.2e}")
        assert normwise < 2e-5


def test_k_zero_and_empty(gpu):
    c = np.full(12, 3.0, np.float32)
    empty = np.zeros(0, np.uint8)
    assert gpu.gemm_e4m3_f32(3, 4, 0, empty, empty, c, True)
    assert np.all(c == 3.0)
    assert gpu.gemm_e5m2_f32(3, 4, 0, empty, empty, c, False)
    assert np.all(c == 0.0)
    assert gpu.gemm_e4m3_f32(0, 4, 5, empty, np.zeros(20, np.uint8), np.zeros(0, np.float32), False)


def test_device_entry_matches_host_entry(gpu):
    m, n, k = 384, 320, 512
    r = rng(5)
    codes = _codes("e4m3", lambda v: np.ones_like(v, bool))
    a = r.choice(codes, m * k)
    b = r.choice(codes, k * n)
    host = np.zeros(m * n, np.float32)
    _run(gpu, "e4m3", m, n, k, a, b, host, False)
    dA, dB = gpu.DeviceBuffer.from_array(a), gpu.DeviceBuffer.from_array(b)
    dC = gpu.DeviceBuffer(host.nbytes)
    assert gpu.lib().ferray_gemm_e4m3_f32_dev(m, n, k, dA.ptr, dB.ptr, dC.ptr, False, None) == 0
    gpu.sync()
    assert dC.download(np.float32, m * n).tobytes() == host.tobytes()
