"""Property tests over random shapes, scalars and flags, through the
host-pointer drop-in entries (the reference fuzzes matmul shapes and values,
fuzz/fuzz_targets/linalg_ops.rs:52-102, and property-tests products with
proptest, tests/property_tests.rs:67-68).

Hypothesis draws M, N from 1..600 (with the tile edges 127-129, 255-257,
511-513 weighted in), K from 0..700, alpha / beta (or accumulate) and the
data seed; examples are derandomized, so every run checks the same cases.

Bars, against the oracle on the same inputs:
  * f64 / f32: |C - C_oracle| <= 2(k + 2) u (|alpha||A||B| + |beta||C0|);
    complex: the same bound with the complex magnitudes doubled;
  * i8 x i8, u8 x i8, i16 -> i32 and quantized_matmul: bit-exact;
  * bf16 / f16 -> f32: 2(k + 2) 2^-24 (|A||B| + |C0|) against the oracle's
    widen-then-SGEMM (the widening is exact);
  * E4M3 / E5M2 -> f32 on entries in {-2, -1, 0, 1, 2}: every product is a
    power of two within a factor 4 of the others and every sum below 2^24, so
    the device result equals the exact product;
  * Ozaki DGEMM (8-16 moduli): |C - C_dd| <= |alpha| (2^(8 - g) ||a_i||
    ||b_j|| + 4 u |C_dd|) + 4 u (|C| + |beta||C0|) against the compensated
    reference (g the scaling exponent of the modulus count; the rounding of
    the scaled operands to integers gives at most sqrt(k) 2^(2 - g) ||a_i||
    ||b_j|| for k <= 700, and a wrong residue or weight would be off by far
    more).
"""
from __future__ import annotations

import math

import numpy as np
import pytest
from hypothesis import HealthCheck, given, settings
from hypothesis import strategies as st

import oracle
from tests.helpers import rng

pytestmark = pytest.mark.gpu

U64, U32 = 2.0 ** -53, 2.0 ** -24

EDGES = [127, 128, 129, 255, 256, 257, 511, 512, 513]
DIM = st.one_of(st.integers(1, 40), st.integers(41, 600), st.sampled_from(EDGES))
DEPTH = st.one_of(st.integers(0, 40), st.integers(41, 700), st.sampled_from(EDGES))
SCALAR = st.sampled_from([0.0, 1.0, -1.0, 0.5, -2.25, 3.0])
SEED = st.integers(0, 2 ** 31 - 1)
FAST = settings(max_examples=150, deadline=None, derandomize=True, database=None,
                suppress_health_check=[HealthCheck.too_slow, HealthCheck.data_too_large])


def _mag(a, b, m, n, k):
    return np.abs(a.reshape(m, k).astype(np.float64)) @ np.abs(b.reshape(k, n).astype(np.float64))


@FAST
@given(m=DIM, n=DIM, k=DEPTH, alpha=SCALAR, beta=SCALAR, seed=SEED,
       f64=st.booleans())
def test_real_random_shapes(gpu, m, n, k, alpha, beta, seed, f64):
    dt, u = (np.float64, U64) if f64 else (np.float32, U32)
    r = rng(seed)
    a = r.standard_normal(m * k).astype(dt)
    b = r.standard_normal(k * n).astype(dt)
    c0 = r.standard_normal(m * n).astype(dt)
    c, ref = c0.copy(), c0.copy()
    if f64:
        assert gpu.gemm_f64(m, n, k, alpha, a, b, beta, c), gpu.last_status()
        oracle.gemm_f64(m, n, k, alpha, a, b, beta, ref)
    else:
        assert gpu.gemm_f32(m, n, k, dt(alpha), a, b, dt(beta), c), gpu.last_status()
        oracle.gemm_f32(m, n, k, dt(alpha), a, b, dt(beta), ref)
    bound = 2 * (k + 2) * u * (abs(alpha) * _mag(a, b, m, n, k) +
                               abs(beta) * np.abs(c0.reshape(m, n).astype(np.float64)))
    err = np.abs(c.astype(np.float64) - ref.astype(np.float64)).reshape(m, n)
    assert np.all(err <= bound), (m, n, k, float(np.max(err - bound)))


@FAST
@given(m=DIM, n=DIM, k=DEPTH, alpha=SCALAR, ai=SCALAR, beta=SCALAR, bi=SCALAR, seed=SEED,
       c64=st.booleans())
def test_complex_random_shapes(gpu, m, n, k, alpha, ai, beta, bi, seed, c64):
    rt, u = (np.float64, U64) if c64 else (np.float32, U32)
    r = rng(seed)
    a = r.standard_normal(2 * m * k).astype(rt)
    b = r.standard_normal(2 * k * n).astype(rt)
    c0 = r.standard_normal(2 * m * n).astype(rt)
    c, ref = c0.copy(), c0.copy()
    run = gpu.gemm_c64 if c64 else gpu.gemm_c32
    orc = oracle.gemm_c64 if c64 else oracle.gemm_c32
    assert run(m, n, k, alpha, ai, a, b, beta, bi, c), gpu.last_status()
    orc(m, n, k, alpha, ai, a, b, beta, bi, ref)
    A = np.abs(a.astype(np.float64).view(np.complex128)).reshape(m, k)
    B = np.abs(b.astype(np.float64).view(np.complex128)).reshape(k, n)
    C0 = np.abs(c0.astype(np.float64).view(np.complex128)).reshape(m, n)
    bound = 4 * (k + 2) * u * (abs(complex(alpha, ai)) * 2 * (A @ B) + abs(complex(beta, bi)) * 2 * C0)
    err = np.abs(c.astype(np.float64).view(np.complex128) -
                 ref.astype(np.float64).view(np.complex128)).reshape(m, n)
    assert np.all(err <= bound + 1e-300), (m, n, k, float(np.max(err - bound)))


@FAST
@given(m=DIM, n=DIM, k=DEPTH, seed=SEED, accumulate=st.booleans(),
       kind=st.sampled_from(["i8s", "u8i8", "i16", "quant"]))
def test_integer_random_shapes(gpu, m, n, k, seed, accumulate, kind):
    r = rng(seed)
    if kind == "i16":
        a = r.integers(-32768, 32768, m * k).astype(np.int16)
        b = r.integers(-32768, 32768, k * n).astype(np.int16)
    else:
        a = (r.integers(0, 256, m * k).astype(np.uint8) if kind in ("u8i8", "quant")
             else r.integers(-128, 128, m * k).astype(np.int8))
        b = r.integers(-128, 128, k * n).astype(np.int8)
    if kind == "quant":
        a_zp, b_zp = int(r.integers(0, 256)), int(r.integers(-128, 128))
        sa, sb = np.float32(r.uniform(0.001, 0.1)), np.float32(r.uniform(0.001, 0.1))
        c0 = r.standard_normal(m * n).astype(np.float32)
        c, ref = c0.copy(), c0.copy()
        assert gpu.quantized_matmul(m, n, k, a, a_zp, sa, b, b_zp, sb, c, accumulate), \
            gpu.last_status()
        oracle.quantized_matmul(m, n, k, a, a_zp, sa, b, b_zp, sb, ref, accumulate)
    else:
        c0 = r.integers(-2 ** 31, 2 ** 31, m * n).astype(np.int32)
        c, ref = c0.copy(), c0.copy()
        run, orc = {"i8s": (gpu.gemm_i8_signed, oracle.gemm_i8_signed),
                    "u8i8": (gpu.gemm_i8, oracle.gemm_i8),
                    "i16": (gpu.gemm_i16, oracle.gemm_i16)}[kind]
        assert run(m, n, k, a, b, c, accumulate), gpu.last_status()
        orc(m, n, k, a, b, ref, accumulate)
    bad = np.flatnonzero(c.view(np.uint32) != ref.view(np.uint32))
    assert bad.size == 0, (kind, m, n, k, accumulate, bad.size, divmod(int(bad[0]), n))


def _bf16_bits(x):
    return (np.asarray(x, np.float32).view(np.uint32) >> 16).astype(np.uint16)


@FAST
@given(m=DIM, n=DIM, k=DEPTH, seed=SEED, accumulate=st.booleans(), bf16=st.booleans())
def test_half_random_shapes(gpu, m, n, k, seed, accumulate, bf16):
    r = rng(seed)
    if bf16:
        a, b = _bf16_bits(r.standard_normal(m * k)), _bf16_bits(r.standard_normal(k * n))
        run, orc = gpu.gemm_bf16_f32, oracle.gemm_bf16_f32
    else:
        a = r.standard_normal(m * k).astype(np.float16).view(np.uint16)
        b = r.standard_normal(k * n).astype(np.float16).view(np.uint16)
        run, orc = gpu.gemm_f16_f32, oracle.gemm_f16_f32
    c0 = r.standard_normal(m * n).astype(np.float32)
    c, ref = c0.copy(), c0.copy()
    assert run(m, n, k, a, b, c, accumulate), gpu.last_status()
    orc(m, n, k, a, b, ref, accumulate)
    if bf16:
        wa = (a.astype(np.uint32) << 16).view(np.float32)
        wb = (b.astype(np.uint32) << 16).view(np.float32)
    else:
        wa, wb = a.view(np.float16).astype(np.float32), b.view(np.float16).astype(np.float32)
    bound = 2 * (k + 2) * U32 * (_mag(wa, wb, m, n, k) +
                                 (np.abs(c0.reshape(m, n)) if accumulate else 0.0))
    err = np.abs(c.astype(np.float64) - ref.astype(np.float64)).reshape(m, n)
    assert np.all(err <= bound + 1e-30), (m, n, k, float(np.max(err - bound)))


# OCP fp8 codes of -2, -1, 0, 1, 2
FP8_CODES = {"e4m3": [0xC0, 0xB8, 0x00, 0x38, 0x40], "e5m2": [0xC0, 0xBC, 0x00, 0x3C, 0x40]}


@FAST
@given(m=DIM, n=DIM, k=DEPTH, seed=SEED, accumulate=st.booleans(),
       fmt=st.sampled_from(["e4m3", "e5m2"]))
def test_fp8_random_shapes(gpu, m, n, k, seed, accumulate, fmt):
    r = rng(seed)
    codes = np.array(FP8_CODES[fmt], np.uint8)
    ia, ib = r.integers(0, 5, m * k), r.integers(0, 5, k * n)
    a, b = codes[ia], codes[ib]
    c0 = r.integers(-1000, 1000, m * n).astype(np.float32)
    c = c0.copy()
    run = gpu.gemm_e4m3_f32 if fmt == "e4m3" else gpu.gemm_e5m2_f32
    assert run(m, n, k, a, b, c, accumulate), gpu.last_status()
    va = (ia - 2).astype(np.float64).reshape(m, k)
    vb = (ib - 2).astype(np.float64).reshape(k, n)
    want = va @ vb + (c0.reshape(m, n).astype(np.float64) if accumulate else 0.0)
    ref = np.zeros(m * n)
    oracle.gemm_fp8_f64(m, n, k, a, b, fmt, ref, False)
    assert np.array_equal(ref.reshape(m, n), va @ vb)  # the oracle's decoding agrees
    assert np.array_equal(c.reshape(m, n).astype(np.float64), want), (fmt, m, n, k)


OZ_MODULI = [256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197, 193]


@settings(max_examples=60, deadline=None, derandomize=True, database=None,
          suppress_health_check=[HealthCheck.too_slow, HealthCheck.data_too_large])
@given(m=DIM, n=DIM, k=DEPTH.filter(lambda x: x > 0), alpha=SCALAR, beta=SCALAR, seed=SEED,
       moduli=st.integers(8, 16))
def test_ozaki_random_shapes(gpu, m, n, k, alpha, beta, seed, moduli):
    r = rng(seed)
    a = r.standard_normal(m * k)
    b = r.standard_normal(k * n)
    c0 = r.standard_normal(m * n)
    c = c0.copy()
    assert gpu.dgemm_ozaki(m, n, k, alpha, a, b, beta, c, moduli), gpu.last_status()
    rows = np.arange(m) if m <= 24 else np.sort(rng(seed + 1).choice(m, 24, replace=False))
    hi, lo = oracle.dd_gemm_f64_sampled(m, n, k, a, b, rows, np.arange(n))
    want = alpha * hi + beta * c0.reshape(m, n)[rows]
    d = np.abs(c.reshape(m, n)[rows] - want)
    M = math.prod(OZ_MODULI[:moduli])
    g = (M.bit_length() - 1 - 2) // 2
    ra = np.linalg.norm(a.reshape(m, k)[rows], axis=1)[:, None]
    cb = np.linalg.norm(b.reshape(k, n), axis=0)[None, :]
    lim = abs(alpha) * (2.0 ** (8 - g) * ra * cb + 4 * U64 * np.abs(hi) + np.abs(lo)) + \
        4 * U64 * (np.abs(want) + abs(beta) * np.abs(c0.reshape(m, n)[rows]))
    assert np.all(d <= lim + 1e-300), (m, n, k, moduli, float(np.max(d - lim)))
