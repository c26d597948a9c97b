"""GPU parity: every C-ABI entry point against the oracle on the same inputs.

Bars (stated per test):
  * integer GEMMs and quantized_matmul: bit-exact;
  * f64 / f32 / complex: elementwise |C_gpu - C_oracle| <= 2*(k+2)*u*(|alpha||A||B| + |beta||C0|)
    with u the unit roundoff (2^-53 for f64, 2^-24 for f32): both results are
    within (k+1)*u*(...) of the exact product, so their distance is bounded by
    twice that.  The normwise error ||C^-C||_F/(||A||_F ||B||_F) is also
    reported for the config shapes;
  * bf16 / f16 -> f32: the reference's own tolerance, |d|/max(|ref|,1) < 1e-5
    (integer_tests.rs:458-466).
"""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

import oracle
from tests.helpers import exact_int_matmul, hexf, normwise_error, rng, wrap_i32

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
U64, U32 = 2.0 ** -53, 2.0 ** -24


def _bound(a, b, m, n, k, alpha, beta, c0, u, cplx=False):
    if cplx:
        A = np.abs(a.reshape(m, k)); B = np.abs(b.reshape(k, n)); C0 = np.abs(c0.reshape(m, n))
        mag = abs(alpha) * (A @ B) * 2 + abs(beta) * C0 * 2
    else:
        mag = abs(alpha) * (np.abs(a.reshape(m, k)) @ np.abs(b.reshape(k, n))) + \
            abs(beta) * np.abs(c0.reshape(m, n))
    return 2 * (k + 2) * u * mag + 1e-300


def _check_real(gpu, dtype, m, n, k, alpha, beta, seed, scale_c0=1.0):
    r = rng(seed)
    a = r.standard_normal(m * k).astype(dtype)
    b = r.standard_normal(k * n).astype(dtype)
    c0 = (r.standard_normal(m * n) * scale_c0).astype(dtype)
    c_gpu = c0.copy(); c_orc = c0.copy()
    if dtype == np.float64:
        ok = gpu.gemm_f64(m, n, k, alpha, a, b, beta, c_gpu)
        oracle.gemm_f64(m, n, k, alpha, a, b, beta, c_orc)
        u = U64
    else:
        ok = gpu.gemm_f32(m, n, k, np.float32(alpha), a, b, np.float32(beta), c_gpu)
        oracle.gemm_f32(m, n, k, np.float32(alpha), a, b, np.float32(beta), c_orc)
        u = U32
    assert ok, gpu.last_status()
    diff = np.abs(c_gpu.astype(np.float64) - c_orc.astype(np.float64)).reshape(m, n)
    bound = _bound(a.astype(np.float64), b.astype(np.float64), m, n, k, alpha, beta,
                   c0.astype(np.float64), u)
    assert np.all(diff <= bound), f"max excess {np.max(diff - bound)}"


SHAPES = [(1, 1, 1), (2, 3, 4), (7, 5, 3), (17, 13, 9), (64, 64, 64), (128, 128, 16),
          (128, 256, 48), (129, 130, 17), (300, 300, 300), (513, 257, 129), (256, 128, 1000)]


@pytest.mark.parametrize("shape", SHAPES)
def test_dgemm_shapes(gpu, shape):
    _check_real(gpu, np.float64, *shape, 1.0, 0.0, seed=sum(shape))


@pytest.mark.parametrize("ab", [(1.0, 1.0), (2.5, -0.75), (-1.0, 0.0), (0.0, 3.0)])
def test_dgemm_alpha_beta(gpu, ab):
    _check_real(gpu, np.float64, 200, 136, 70, ab[0], ab[1], seed=11)
    _check_real(gpu, np.float64, 256, 256, 64, ab[0], ab[1], seed=12)


@pytest.mark.parametrize("shape", SHAPES)
def test_sgemm_shapes(gpu, shape):
    _check_real(gpu, np.float32, *shape, 1.0, 0.0, seed=100 + sum(shape))


@pytest.mark.parametrize("ab", [(1.0, 1.0), (2.5, -0.75), (0.0, 3.0)])
def test_sgemm_alpha_beta(gpu, ab):
    _check_real(gpu, np.float32, 200, 136, 70, ab[0], ab[1], seed=21)
    _check_real(gpu, np.float32, 256, 256, 64, ab[0], ab[1], seed=22)


def test_dgemm_known_answers(gpu):
    with open(os.path.join(GOLD, "ferray_kats.json")) as f:
        kats = json.load(f)
    k = kats["gemm_f64_2x2"]
    c = np.zeros(4)
    assert gpu.gemm_f64(2, 2, 2, 1.0, np.array(k["a"]), np.array(k["b"]), 0.0, c)
    assert c.tolist() == k["c"]
    g = kats["numpy_cancellation_2x3x2"]   # exact in any summation order
    a = np.array([hexf(s) for s in g["a"]]); b = np.array([hexf(s) for s in g["b"]])
    c = np.zeros(4)
    assert gpu.gemm_f64(2, 2, 3, 1.0, a, b, 0.0, c)
    assert c.tolist() == [hexf(s) for s in g["c"]]
    g = kats["numpy_random_4x3x5"]
    a = np.array([hexf(s) for s in g["a"]]); b = np.array([hexf(s) for s in g["b"]])
    want = np.array([hexf(s) for s in g["c"]])
    c = np.zeros(20)
    assert gpu.gemm_f64(4, 5, 3, 1.0, a, b, 0.0, c)
    assert np.all(np.abs(c - want) <= 2 * 5 * U64 * (np.abs(a.reshape(4, 3)) @ np.abs(b.reshape(3, 5))).ravel())


def test_dgemm_edges(gpu):
    # m == 0 / n == 0: no-op; k == 0: C = beta*C (outer.rs:211-231)
    c = np.arange(6, dtype=np.float64)
    assert gpu.gemm_f64(0, 3, 4, 1.0, np.zeros(12), np.zeros(12), 0.0, c)
    assert gpu.gemm_f64(2, 0, 4, 1.0, np.zeros(12), np.zeros(12), 0.0, c)
    assert c.tolist() == list(range(6))
    assert gpu.gemm_f64(2, 3, 0, 1.0, np.zeros(1), np.zeros(1), 2.0, c)
    assert c.tolist() == [0.0, 2.0, 4.0, 6.0, 8.0, 10.0]
    assert gpu.gemm_f64(2, 3, 0, 1.0, np.zeros(1), np.zeros(1), 1.0, c)
    assert c.tolist() == [0.0, 2.0, 4.0, 6.0, 8.0, 10.0]
    c[:] = np.nan
    assert gpu.gemm_f64(2, 3, 0, 1.0, np.zeros(1), np.zeros(1), 0.0, c)
    assert c.tolist() == [0.0] * 6
    # beta == 0 never reads C, so NaN in C does not propagate
    c[:] = np.nan
    assert gpu.gemm_f64(2, 2, 2, 1.0, np.array([1., 2, 3, 4]), np.array([5., 6, 7, 8]), 0.0, c)
    assert c[:4].tolist() == [19.0, 22.0, 43.0, 50.0]


def test_dgemm_config1_lcg_512(gpu):
    """Config 1: M=N=K=512, make_matrix seeds 1 and 2 (bench_matmul.rs:21-34)."""
    n = 512
    a = oracle.lcg_matrix(n, 1); b = oracle.lcg_matrix(n, 2)
    c = np.zeros(n * n); ref = np.zeros(n * n)
    assert gpu.gemm_f64(n, n, n, 1.0, a, b, 0.0, c)
    oracle.gemm_f64(n, n, n, 1.0, a, b, 0.0, ref)
    bound = _bound(a, b, n, n, n, 1.0, 0.0, np.zeros(n * n), U64)
    assert np.all(np.abs(c - ref).reshape(n, n) <= bound)
    # make_matrix yields values in [-1, 0), so ||C|| ~ 0.75 ||A|| ||B|| here and
    # both norms measure the same few-ulp relative error.
    e_ab, e_c = normwise_error(c, ref, a, b)
    assert e_ab < 4e-15 and e_c < 4e-15, (e_ab, e_c)


# ---------------------------------------------------------------- complex

def _check_cplx(gpu, single, m, n, k, al, be, seed):
    r = rng(seed)
    rdt = np.float32 if single else np.float64
    cdt = np.complex64 if single else np.complex128
    A = (r.standard_normal((m, k)) + 1j * r.standard_normal((m, k))).astype(cdt)
    B = (r.standard_normal((k, n)) + 1j * r.standard_normal((k, n))).astype(cdt)
    C0 = (r.standard_normal((m, n)) + 1j * r.standard_normal((m, n))).astype(cdt)
    cg = C0.copy(); co = C0.copy()
    fn = gpu.gemm_c32 if single else gpu.gemm_c64
    of = oracle.gemm_c32 if single else oracle.gemm_c64
    assert fn(m, n, k, al.real, al.imag, A, B, be.real, be.imag, cg), gpu.last_status()
    of(m, n, k, al.real, al.imag, A.view(rdt).ravel(), B.view(rdt).ravel(), be.real, be.imag,
       co.view(rdt).ravel())
    u = U32 if single else U64
    mag = abs(al) * (np.abs(A).astype(np.float64) @ np.abs(B).astype(np.float64)) + \
        abs(be) * np.abs(C0).astype(np.float64)
    bound = 2 * (2 * k + 4) * u * mag * 2 + 1e-300
    d = np.abs(cg.astype(np.complex128) - co.astype(np.complex128))
    assert np.all(d <= bound), f"max excess {np.max(d - bound)}"


CSHAPES = [(1, 1, 1), (2, 2, 2), (7, 5, 3), (17, 13, 9), (64, 64, 64), (128, 64, 8),
           (130, 70, 33), (256, 128, 200)]


@pytest.mark.parametrize("shape", CSHAPES)
@pytest.mark.parametrize("single", [False, True])
def test_complex_shapes(gpu, shape, single):
    _check_cplx(gpu, single, *shape, 1 + 0j, 0j, seed=31 + sum(shape))


@pytest.mark.parametrize("single", [False, True])
def test_complex_alpha_beta(gpu, single):
    _check_cplx(gpu, single, 96, 80, 40, 0.5 - 2j, -1.5 + 0.25j, seed=41)
    _check_cplx(gpu, single, 128, 64, 32, 1 + 0j, 1 + 0j, seed=42)


def test_complex_known_answers_and_k0(gpu):
    with open(os.path.join(GOLD, "ferray_kats.json")) as f:
        kats = json.load(f)
    for name in ("complex_2x2_identity", "complex_2x2_nontrivial"):
        g = kats[name]
        a = np.array(g["a"], np.float64).ravel(); b = np.array(g["b"], np.float64).ravel()
        c = np.zeros(8)
        assert gpu.gemm_c64(2, 2, 2, 1.0, 0.0, a, b, 0.0, 0.0, c)
        assert np.all(np.abs(c - np.array(g["c"], np.float64).ravel()) <= g["tol"])
    C0 = np.array([1 + 2j, 3 - 1j, -2 + 0.5j, 4j])
    c = C0.copy()
    assert gpu.gemm_c64(2, 2, 0, 1.0, 0.0, np.zeros(2), np.zeros(2), 0.5, -1.0, c)
    assert np.allclose(c, (0.5 - 1j) * C0, rtol=1e-15, atol=0)


# ---------------------------------------------------------------- integers

def _ints(r, lo, hi, count, dt):
    return r.integers(lo, hi + 1, count).astype(dt)


ISHAPES = [(1, 1, 1), (7, 11, 9), (12, 16, 20), (13, 17, 19), (64, 64, 64), (128, 128, 64),
           (129, 200, 65), (256, 384, 320), (300, 100, 1000)]


@pytest.mark.parametrize("shape", ISHAPES)
def test_i8_signed_bit_exact(gpu, shape):
    m, n, k = shape
    r = rng(50 + m)
    a = _ints(r, -128, 127, m * k, np.int8); b = _ints(r, -128, 127, k * n, np.int8)
    c = np.zeros(m * n, np.int32); ref = np.zeros(m * n, np.int32)
    assert gpu.gemm_i8_signed(m, n, k, a, b, c, False)
    oracle.gemm_i8_signed(m, n, k, a, b, ref, False)
    assert np.array_equal(c, ref)
    c0 = _ints(r, -2 ** 31, 2 ** 31 - 1, m * n, np.int32)
    c = c0.copy(); ref = c0.copy()
    assert gpu.gemm_i8_signed(m, n, k, a, b, c, True)
    oracle.gemm_i8_signed(m, n, k, a, b, ref, True)
    assert np.array_equal(c, ref)


@pytest.mark.parametrize("shape", ISHAPES)
def test_u8_i8_bit_exact(gpu, shape):
    m, n, k = shape
    r = rng(60 + m)
    a = _ints(r, 0, 255, m * k, np.uint8); b = _ints(r, -128, 127, k * n, np.int8)
    c = np.zeros(m * n, np.int32); ref = np.zeros(m * n, np.int32)
    assert gpu.gemm_i8(m, n, k, a, b, c, False)
    oracle.gemm_i8(m, n, k, a, b, ref, False)
    assert np.array_equal(c, ref)
    c0 = _ints(r, -1000, 1000, m * n, np.int32)
    c = c0.copy(); ref = c0.copy()
    assert gpu.gemm_i8(m, n, k, a, b, c, True)
    oracle.gemm_i8(m, n, k, a, b, ref, True)
    assert np.array_equal(c, ref)


@pytest.mark.parametrize("shape", ISHAPES)
def test_i16_bit_exact(gpu, shape):
    m, n, k = shape
    r = rng(70 + m)
    a = _ints(r, -32768, 32767, m * k, np.int16); b = _ints(r, -32768, 32767, k * n, np.int16)
    c = np.zeros(m * n, np.int32); ref = np.zeros(m * n, np.int32)
    assert gpu.gemm_i16(m, n, k, a, b, c, False)
    oracle.gemm_i16(m, n, k, a, b, ref, False)
    assert np.array_equal(c, ref)
    c0 = _ints(r, -2 ** 31, 2 ** 31 - 1, m * n, np.int32)
    c = c0.copy(); ref = c0.copy()
    assert gpu.gemm_i16(m, n, k, a, b, c, True)
    oracle.gemm_i16(m, n, k, a, b, ref, True)
    assert np.array_equal(c, ref)


def test_integer_reference_kats(gpu):
    """The reference's own index-formula inputs (integer_tests.rs:65-263)."""
    from tests.test_oracle import _int_kats
    for kind, m, n, k, araw, braw in _int_kats():
        if kind == "i16":
            a = np.array(araw, np.uint16).view(np.int16); b = np.array(braw, np.uint16).view(np.int16)
            fn = gpu.gemm_i16
        else:
            a = np.array(araw, np.uint8).view(np.int8); b = np.array(braw, np.uint8).view(np.int8)
            fn = gpu.gemm_i8_signed
        c = np.zeros(m * n, np.int32)
        assert fn(m, n, k, a, b, c, False)
        assert np.array_equal(c, wrap_i32(exact_int_matmul(a.reshape(m, k), b.reshape(k, n))).ravel())


def test_integer_extremes_and_k0(gpu):
    m, n, k = 33, 65, 130
    a = np.full(m * k, -128, np.int8); b = np.full(k * n, -128, np.int8)
    c = np.zeros(m * n, np.int32)
    assert gpu.gemm_i8_signed(m, n, k, a, b, c, False)
    assert np.all(c == k * 16384)
    a16 = np.full(m * k, -32768, np.int16); b16 = np.full(k * n, 32767, np.int16)
    assert gpu.gemm_i16(m, n, k, a16, b16, c, False)
    assert np.array_equal(c, wrap_i32(exact_int_matmul(a16.reshape(m, k), b16.reshape(k, n))).ravel())
    c0 = np.arange(m * n, dtype=np.int32)
    c = c0.copy()
    assert gpu.gemm_i8_signed(m, n, 0, a, b, c, True)
    assert np.array_equal(c, c0)
    assert gpu.gemm_i8_signed(m, n, 0, a, b, c, False)
    assert np.all(c == 0)


def test_i8_accumulator_wraps_mod_2_32(gpu):
    """k = 2^18 products of (-128)(-128) = 2^14 sum to 2^32 = 0 mod 2^32.
    Exact wrapping semantics (faer_fallback.rs:278-341) require the device
    accumulation to wrap rather than saturate."""
    m, n, k = 16, 16, 1 << 18
    a = np.full(m * k, -128, np.int8); b = np.full(k * n, -128, np.int8)
    c = np.zeros(m * n, np.int32)
    assert gpu.gemm_i8_signed(m, n, k, a, b, c, False)
    assert np.all(c == 0), np.unique(c)[:5]


@pytest.mark.parametrize("params", [
    (8, 12, 16, 0, 0.01, 0, 0.02), (10, 14, 18, 64, 0.0078, -3, 0.015),
    (6, 10, 12, 64, 0.05, 0, 0.05), (130, 129, 300, 200, 0.003, 77, 0.7),
    (256, 256, 128, 128, 0.01, -128, 0.01)])
def test_quantized_matmul_bit_exact(gpu, params):
    m, n, k, azp, sa, bzp, sb = params
    r = rng(80 + m)
    a = _ints(r, 0, 255, m * k, np.uint8); b = _ints(r, -128, 127, k * n, np.int8)
    for acc in (False, True):
        c0 = (np.arange(m * n) * 0.1).astype(np.float32)
        c = c0.copy(); ref = c0.copy()
        assert gpu.quantized_matmul(m, n, k, a, azp, sa, b, bzp, sb, c, acc)
        oracle.quantized_matmul(m, n, k, a, azp, sa, b, bzp, sb, ref, acc)
        assert np.array_equal(c, ref)


# ---------------------------------------------------------------- halves

@pytest.mark.parametrize("shape", [(1, 1, 1), (7, 5, 3), (64, 64, 64), (129, 130, 70),
                                   (256, 256, 512)])
@pytest.mark.parametrize("fmt", ["bf16", "f16"])
def test_half_to_f32(gpu, shape, fmt):
    m, n, k = shape
    r = rng(90 + m)
    af = r.standard_normal(m * k).astype(np.float32); bf = r.standard_normal(k * n).astype(np.float32)
    if fmt == "bf16":
        a = (af.view(np.uint32) >> 16).astype(np.uint16); b = (bf.view(np.uint32) >> 16).astype(np.uint16)
        fn, of = gpu.gemm_bf16_f32, oracle.gemm_bf16_f32
    else:
        a = af.astype(np.float16).view(np.uint16); b = bf.astype(np.float16).view(np.uint16)
        fn, of = gpu.gemm_f16_f32, oracle.gemm_f16_f32
    for acc in (False, True):
        c0 = r.standard_normal(m * n).astype(np.float32)
        c = c0.copy(); ref = c0.copy()
        assert fn(m, n, k, a, b, c, acc)
        of(m, n, k, a, b, ref, acc)
        aw = a.astype(np.uint32) << 16 if fmt == "bf16" else None
        A = (aw.view(np.float32) if fmt == "bf16" else a.view(np.float16).astype(np.float32))
        bw = b.astype(np.uint32) << 16 if fmt == "bf16" else None
        B = (bw.view(np.float32) if fmt == "bf16" else b.view(np.float16).astype(np.float32))
        bound = _bound(A.astype(np.float64), B.astype(np.float64), m, n, k, 1.0,
                       1.0 if acc else 0.0, c0.astype(np.float64), U32)
        d = np.abs(c.astype(np.float64) - ref.astype(np.float64)).reshape(m, n)
        assert np.all(d <= bound), np.max(d - bound)
        if k <= 70:   # the reference's own bar at its test sizes (integer_tests.rs:458-466)
            rel = d.ravel() / np.maximum(np.abs(ref), 1.0)
            assert rel.max() < 1e-5, rel.max()


def test_i16_wide_and_flat_shapes(gpu):
    """Shapes whose padded column count is many times the padded row count
    (the column sums are cleared by the row pass, several words per row)."""
    for m, n, k in ((1, 1500, 70), (3, 40000, 33), (130, 300, 129)):
        r = rng(m * 7 + n)
        a = _ints(r, -32768, 32767, m * k, np.int16); b = _ints(r, -32768, 32767, k * n, np.int16)
        c = np.zeros(m * n, np.int32); ref = np.zeros(m * n, np.int32)
        assert gpu.gemm_i16(m, n, k, a, b, c, False)
        oracle.gemm_i16(m, n, k, a, b, ref, False)
        assert np.array_equal(c, ref), (m, n, k)
