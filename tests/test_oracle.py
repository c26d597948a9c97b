"""Pins the oracle (oracle/ferray_oracle.c) before it is trusted as the checker.

Each restated function is checked against something independent of its own
formula: the reference's golden vectors and known-answer tests
(tests/golden/), numpy computing the same closed form a different way, exact
arithmetic on hand-built inputs whose answer is fixed by IEEE rounding, and
brute force on tiny inputs.  Several checks are built so that a plausible
mistake (no KC blocking, beta applied on every block, alpha applied once,
saturating instead of wrapping, a transposed operand) fails them.
"""
from __future__ import annotations

import json
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle
from tests.helpers import exact_int_matmul, hexf, rng, ulp_diff, wrap_i32

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _kats():
    with open(os.path.join(GOLD, "ferray_kats.json")) as f:
        return json.load(f)


# ---------------------------------------------------------------- real f64/f32

def test_gemm_f64_2x2_kat():
    k = _kats()["gemm_f64_2x2"]
    c = np.zeros(4)
    oracle.gemm_f64(2, 2, 2, 1.0, np.array(k["a"]), np.array(k["b"]), 0.0, c)
    assert c.tolist() == k["c"]


@pytest.mark.parametrize("case", ["numpy_cancellation_2x3x2", "numpy_random_4x3x5"])
def test_gemm_f64_numpy_hex_goldens(case):
    g = _kats()[case]
    a = np.array([hexf(s) for s in g["a"]])
    b = np.array([hexf(s) for s in g["b"]])
    want = np.array([hexf(s) for s in g["c"]])
    c = np.zeros(g["m"] * g["n"])
    oracle.gemm_f64(g["m"], g["n"], g["k"], 1.0, a, b, 0.0, c)
    assert ulp_diff(c, want).max() <= g["max_ulp"]


@pytest.mark.parametrize("name", ["matmul.json", "dot.json"])
def test_reference_fixtures(name):
    """fixtures/linalg/{matmul,dot}.json (numpy 2.4.5), ferray's tolerance
    max(tolerance_ulps, MIN_ULP_TOLERANCE=10) (ferray-test-oracle/src/lib.rs:59)."""
    with open(os.path.join(GOLD, name)) as f:
        fx = json.load(f)
    for tc in fx["test_cases"]:
        a, b, e = tc["inputs"]["a"], tc["inputs"]["b"], tc["expected"]
        sa, sb = a["shape"], b["shape"]
        m = sa[0] if len(sa) == 2 else 1
        k = sa[-1]
        n = sb[1] if len(sb) == 2 else 1
        tol = max(tc.get("tolerance_ulps", 0), 10)
        if e["dtype"] == "float32":
            A = np.array(a["data"], np.float32); B = np.array(b["data"], np.float32)
            C = np.zeros(m * n, np.float32)
            oracle.gemm_f32(m, n, k, 1.0, A, B, 0.0, C)
            want = np.array(e["data"], np.float32).ravel()
            ulps = np.abs(C.view(np.int32).astype(np.int64) - want.view(np.int32).astype(np.int64))
            assert ulps.max() <= tol, tc["name"]
        else:
            A = np.array(a["data"], np.float64); B = np.array(b["data"], np.float64)
            C = np.zeros(m * n)
            oracle.gemm_f64(m, n, k, 1.0, A, B, 0.0, C)
            want = np.array(e["data"], np.float64).ravel()
            assert ulp_diff(C, want).max() <= tol, tc["name"]


def _kc_block_row():
    """One row of A over k = 512 where the KC=256 split decides the answer.

    x0 = 2^53, x256 = 1, x257 = -2^53, every other product 0 (b = ones).
    Blocked (ferray): acc0 = 2^53; acc1 = 1 - 2^53 = -(2^53-1) exactly;
    C = acc0 + acc1 = 1.  One unblocked chain: 2^53 + 1 rounds to 2^53
    (tie to even), minus 2^53 gives 0."""
    a = np.zeros(512)
    a[0] = 2.0 ** 53
    a[256] = 1.0
    a[257] = -(2.0 ** 53)
    return a


def test_gemm_f64_kc_blocking_decides_rounding():
    a = _kc_block_row()
    b = np.ones(512)
    c = np.zeros(1)
    oracle.gemm_f64(1, 1, 512, 1.0, a, b, 0.0, c)
    assert c[0] == 1.0
    # the Goto restatement must agree bit for bit
    c2 = np.zeros(1)
    oracle.gemm_f64_goto(1, 1, 512, 1.0, a, b, 0.0, c2)
    assert c2[0] == 1.0


def test_gemm_f64_alpha_per_block_beta_first_block():
    """alpha scales each block partial, beta only the first block's store:
    alpha=2, beta=0.5, C0=4: block 0 stores fma(4, .5, 2^54) = 2^54 (2^54+2
    is a tie, rounds to even); block 1 adds 2*(1-2^53) = -(2^54-2) exactly, so
    C = 2.  Beta on both blocks gives 2 - 2^53; alpha applied once to the
    unscaled total gives 6."""
    a = _kc_block_row()
    b = np.ones(512)
    c = np.array([4.0])
    oracle.gemm_f64(1, 1, 512, 2.0, a, b, 0.5, c)
    assert c[0] == 2.0
    # k = 300: acc0 = sum_{0..255} 1 = 256, acc1 = 44; alpha = 3, beta = -1, C0 = 5
    c = np.array([5.0])
    oracle.gemm_f64(1, 1, 300, 3.0, np.ones(300), np.ones(300), -1.0, c)
    assert c[0] == -5.0 + 3 * 256 + 3 * 44


def test_gemm_f64_k0_and_beta_edges():
    c = np.array([1.0, np.nan, 3.0, 4.0])
    oracle.gemm_f64(2, 2, 0, 1.0, np.zeros(1), np.zeros(1), 0.0, c)
    assert c.tolist() == [0.0, 0.0, 0.0, 0.0]          # beta == 0 writes zeros
    c = np.array([1.0, 2.0, 3.0, 4.0])
    oracle.gemm_f64(2, 2, 0, 1.0, np.zeros(1), np.zeros(1), 2.0, c)
    assert c.tolist() == [2.0, 4.0, 6.0, 8.0]
    c = np.array([np.nan] * 4)                          # beta == 0 never reads C
    oracle.gemm_f64(2, 2, 2, 1.0, np.array([1., 2, 3, 4]), np.array([5., 6, 7, 8]), 0.0, c)
    assert c.tolist() == [19.0, 22.0, 43.0, 50.0]


def test_gemm_f64_exact_on_small_integers_and_transposition():
    """Integer-valued operands with exact partial sums: the formula must equal
    numpy's product exactly; a transposed or swapped operand would not."""
    r = rng(1)
    for (m, n, k) in [(7, 5, 3), (17, 13, 9), (33, 65, 300), (64, 64, 513)]:
        a = r.integers(-8, 9, size=(m, k)).astype(np.float64)
        b = r.integers(-8, 9, size=(k, n)).astype(np.float64)
        c = np.zeros(m * n)
        oracle.gemm_f64(m, n, k, 1.0, a.ravel(), b.ravel(), 0.0, c)
        assert np.array_equal(c.reshape(m, n), a @ b)
        cf = np.zeros(m * n, np.float32)
        oracle.gemm_f32(m, n, k, 1.0, a.ravel().astype(np.float32), b.ravel().astype(np.float32),
                        0.0, cf)
        assert np.array_equal(cf.reshape(m, n).astype(np.float64), a @ b)


def test_gemm_f64_error_bound_random():
    """|C - exact| <= gamma_256 * sum|a||b| elementwise (each KC chain is a
    recursive sum of <= 256 terms plus the block adds)."""
    r = rng(2)
    m, n, k = 9, 11, 700
    a = r.standard_normal((m, k)); b = r.standard_normal((k, n))
    c = np.zeros(m * n)
    oracle.gemm_f64(m, n, k, 1.0, a.ravel(), b.ravel(), 0.0, c)
    exact = np.array([[float(sum(Fraction(x) * Fraction(y) for x, y in zip(a[i], b[:, j])))
                       for j in range(n)] for i in range(m)])
    u = 2.0 ** -53
    bound = (260 * u / (1 - 260 * u)) * (np.abs(a) @ np.abs(b))
    assert np.all(np.abs(c.reshape(m, n) - exact) <= bound)


@pytest.mark.parametrize("shape", [(8, 6, 4), (17, 13, 9), (300, 300, 300), (513, 257, 129),
                                   (64, 1100, 600)])
@pytest.mark.parametrize("ab", [(1.0, 0.0), (1.5, 1.0), (0.5, -2.0)])
def test_goto_restatement_bit_identical(shape, ab):
    """The threaded AVX2 Goto restatement (the timed CPU baseline) and the
    scalar formula are two independent codings of the same arithmetic; they
    must agree bit for bit.  Inputs follow the reference driver tests
    (outer.rs:3313-3315)."""
    m, n, k = shape
    alpha, beta = ab
    a = np.arange(m * k) * 0.013 - 0.7
    b = np.arange(k * n) * 0.011 + 0.3
    c0 = np.arange(m * n) * 0.005 - 1.0
    c1 = c0.copy(); c2 = c0.copy()
    oracle.gemm_f64(m, n, k, alpha, a, b, beta, c1)
    oracle.gemm_f64_goto(m, n, k, alpha, a, b, beta, c2)
    if beta in (0.0, 1.0):
        assert np.array_equal(c1.view(np.int64), c2.view(np.int64))
    else:  # the scalar edge tile rounds beta*C separately (kernel_avx2.rs:280-287)
        assert np.allclose(c1, c2, rtol=1e-13, atol=1e-13)
    # and the driver-test tolerance against a plain f64 loop (outer.rs:3335-3345)
    ref = alpha * (a.reshape(m, k) @ b.reshape(k, n)) + beta * c0.reshape(m, n)
    assert np.all(np.abs(c1.reshape(m, n) - ref) / np.maximum(np.abs(ref), 1.0) < 1e-10)
    af, bf, cf = a.astype(np.float32), b.astype(np.float32), c0.astype(np.float32)
    s1, s2 = cf.copy(), cf.copy()
    oracle.gemm_f32(m, n, k, np.float32(alpha), af, bf, np.float32(beta), s1)
    oracle.gemm_f32_goto(m, n, k, np.float32(alpha), af, bf, np.float32(beta), s2)
    if beta in (0.0, 1.0):
        assert np.array_equal(s1.view(np.int32), s2.view(np.int32))


def test_config1_lcg_generator():
    """make_matrix (bench_matmul.rs:21-34) recomputed with numpy uint64."""
    out = oracle.lcg_matrix(8, 1)
    st = np.uint64(1) * np.uint64(6364136223846793005) + np.uint64(1)
    vals = []
    with np.errstate(over="ignore"):
        for _ in range(64):
            st = st * np.uint64(6364136223846793005) + np.uint64(1)
            vals.append(float(st >> np.uint64(33)) / float(1 << 31) - 1.0)
    assert out.tolist() == vals


# ---------------------------------------------------------------- complex

@pytest.mark.parametrize("case", ["complex_2x2_identity", "complex_2x2_nontrivial"])
def test_complex_kats(case):
    g = _kats()[case]
    a = np.array(g["a"], np.float64).ravel(); b = np.array(g["b"], np.float64).ravel()
    want = np.array(g["c"], np.float64).ravel()
    c = np.zeros(8)
    oracle.gemm_c64(2, 2, 2, 1.0, 0.0, a, b, 0.0, 0.0, c)
    assert np.all(np.abs(c - want) <= g["tol"])
    c32 = np.zeros(8, np.float32)
    oracle.gemm_c32(2, 2, 2, 1.0, 0.0, a.astype(np.float32), b.astype(np.float32), 0.0, 0.0, c32)
    assert np.all(np.abs(c32 - want) <= max(g["tol"], 1e-6))


def test_complex_matches_numpy_and_alpha_beta():
    r = rng(3)
    m, n, k = 13, 7, 300
    A = r.standard_normal((m, k)) + 1j * r.standard_normal((m, k))
    B = r.standard_normal((k, n)) + 1j * r.standard_normal((k, n))
    C0 = r.standard_normal((m, n)) + 1j * r.standard_normal((m, n))
    al, be = 0.5 - 2j, -1.5 + 0.25j
    c = C0.copy()
    oracle.gemm_c64(m, n, k, al.real, al.imag, A.view(np.float64).ravel(),
                    B.view(np.float64).ravel(), be.real, be.imag, c.view(np.float64).ravel())
    ref = al * (A @ B) + be * C0
    assert np.max(np.abs(c - ref)) < 1e-12 * np.max(np.abs(ref))
    # exact small-integer complex case: no rounding anywhere
    Ai = r.integers(-4, 5, (m, k)) + 1j * r.integers(-4, 5, (m, k))
    Bi = r.integers(-4, 5, (k, n)) + 1j * r.integers(-4, 5, (k, n))
    ci = np.zeros((m, n), np.complex128)
    oracle.gemm_c64(m, n, k, 1.0, 0.0, Ai.view(np.float64).ravel(), Bi.view(np.float64).ravel(),
                    0.0, 0.0, ci.view(np.float64).ravel())
    assert np.array_equal(ci, Ai @ Bi)
    # k == 0 with complex beta (outer.rs:1249-1266)
    c = C0.copy()
    oracle.gemm_c64(m, n, 0, 1.0, 0.0, np.zeros(2), np.zeros(2), be.real, be.imag,
                    c.view(np.float64).ravel())
    assert np.allclose(c, be * C0, rtol=1e-15, atol=0)


# ---------------------------------------------------------------- integers

def _int_kats():
    """Index-formula inputs of the reference's integer tests
    (integer_tests.rs:65-263)."""
    out = []
    m, n, k = 12, 24, 16
    out.append(("i16", m, n, k, [((i * 17 - 200) & 0xFFFF) for i in range(m * k)],
                [((i * 13 - 100) & 0xFFFF) for i in range(k * n)]))
    m, n, k = 7, 11, 9
    out.append(("i16", m, n, k, [((i * 23 + 5) & 0xFFFF) for i in range(m * k)],
                [((i * 19 - 60) & 0xFFFF) for i in range(k * n)]))
    m, n, k = 12, 16, 20
    out.append(("i8s", m, n, k, [((i * 7 - 60) & 0xFF) for i in range(m * k)],
                [((i * 11 - 80) & 0xFF) for i in range(k * n)]))
    m, n, k = 13, 17, 19

    def ea(i):
        return [-128, 127, -127, 1, -1, (i * 5 - 80)][i % 6] if i % 6 != 5 else (i * 5 - 80)

    def eb(i):
        return [127, -128, 0][i % 5] if i % 5 < 3 else (i * 3 - 50)
    out.append(("i8s", m, n, k, [ea(i) & 0xFF for i in range(m * k)],
                [eb(i) & 0xFF for i in range(k * n)]))
    m, n, k = 9, 13, 11
    out.append(("i8s", m, n, k,
                [(-(i % 100) if i % 2 == 0 else (i % 100)) & 0xFF for i in range(m * k)],
                [(-(i % 80) if i % 3 == 0 else (i % 80)) & 0xFF for i in range(k * n)]))
    return out


@pytest.mark.parametrize("idx", range(5))
def test_integer_kats_against_numpy(idx):
    kind, m, n, k, araw, braw = _int_kats()[idx]
    if kind == "i16":
        a = np.array(araw, np.uint16).view(np.int16); b = np.array(braw, np.uint16).view(np.int16)
        fn = oracle.gemm_i16
    else:
        a = np.array(araw, np.uint8).view(np.int8); b = np.array(braw, np.uint8).view(np.int8)
        fn = oracle.gemm_i8_signed
    c = np.zeros(m * n, np.int32)
    fn(m, n, k, a, b, c, False)
    want = wrap_i32(exact_int_matmul(a.reshape(m, k), b.reshape(k, n))).ravel()
    assert np.array_equal(c, want)
    # accumulate adds mod 2^32 (integer_tests.rs:112-135)
    c0 = (np.arange(m * n, dtype=np.int64) * 1000).astype(np.int32)
    c = c0.copy()
    fn(m, n, k, a, b, c, True)
    assert np.array_equal(c, wrap_i32(c0.astype(np.int64) + want.astype(np.int64)))


def test_integer_wraps_mod_2_32():
    """i16 at full range overflows i32; the oracle must wrap, not saturate
    (the reference output is the low 32 bits)."""
    m, n, k = 3, 4, 4096
    a = np.full(m * k, -32768, np.int16); b = np.full(k * n, -32768, np.int16)
    c = np.zeros(m * n, np.int32)
    oracle.gemm_i16(m, n, k, a, b, c, False)
    assert np.all(c == 0)                   # k * 2^30 = 2^42 = 0 mod 2^32
    a[:3] = 3                               # row 0: 3*(-32768)*3 + ... perturbs the low bits
    oracle.gemm_i16(m, n, k, a, b, c, False)
    assert np.array_equal(c, wrap_i32(exact_int_matmul(a.reshape(m, k), b.reshape(k, n))).ravel())
    r = rng(4)
    m, n, k = 6, 5, 4096
    a = r.integers(-32768, 32768, m * k).astype(np.int16)
    b = r.integers(-32768, 32768, k * n).astype(np.int16)
    c = np.zeros(m * n, np.int32)
    oracle.gemm_i16(m, n, k, a, b, c, False)
    assert np.array_equal(c, wrap_i32(exact_int_matmul(a.reshape(m, k), b.reshape(k, n))).ravel())


def test_u8_gemm_exact_and_avx2_saturation_divergence():
    """gemm_i8 is exact (VNNI/NEON/fallback).  The default-AVX2 emulation
    saturates adjacent pair sums at i16: with a = 255 and b = 127 every pair
    is 64770 -> 32767 (kernel_avx2_i8.rs:9-13)."""
    m, n, k = 2, 3, 8
    a = np.full(m * k, 255, np.uint8); b = np.full(k * n, 127, np.int8)
    c = np.zeros(m * n, np.int32)
    oracle.gemm_i8(m, n, k, a, b, c, False)
    assert np.all(c == 8 * 255 * 127)
    cs = np.zeros(m * n, np.int32)
    oracle.gemm_i8_avx2_saturating(m, n, k, a, b, cs, False)
    assert np.all(cs == 4 * 32767)
    # inside the reference tests' [0,127] range the two agree (integer_tests.rs:320-328)
    r = rng(5)
    a = r.integers(0, 128, 10 * 18).astype(np.uint8); b = r.integers(-128, 128, 18 * 14).astype(np.int8)
    c1 = np.zeros(140, np.int32); c2 = np.zeros(140, np.int32)
    oracle.gemm_i8(10, 14, 18, a, b, c1, False)
    oracle.gemm_i8_avx2_saturating(10, 14, 18, a, b, c2, False)
    assert np.array_equal(c1, c2)
    assert np.array_equal(c1, wrap_i32(exact_int_matmul(a.reshape(10, 18), b.reshape(18, 14))).ravel())


def _naive_quantized(m, n, k, a, azp, sa, b, bzp, sb):
    """The reference test's own formula (integer_tests.rs:39-62): subtract the
    zero points first, multiply, wrap, then scale — a different expansion
    from the product path's row/column-sum correction."""
    av = a.reshape(m, k).astype(np.int64) - int(azp)
    bv = b.reshape(k, n).astype(np.int64) - int(bzp)
    acc = wrap_i32(av @ bv)
    return (np.float32(sa) * np.float32(sb)) * acc.astype(np.float32)


@pytest.mark.parametrize("params", [
    (8, 12, 16, lambda i: ((i * 7 + 3) & 0x7f), 0, 0.01, lambda i: ((i * 11 - 50)), 0, 0.02),
    (10, 14, 18, lambda i: ((i * 13) & 0x7f), 64, 0.0078, lambda i: ((i * 7 - 40)), -3, 0.015),
    (6, 10, 12, lambda i: ((i * 5) & 0x3f), 64, 0.05, lambda i: (i - 30), 0, 0.05),
    (9, 7, 33, lambda i: ((i * 29 + 11) & 0xff), 200, 0.003, lambda i: ((i * 5 - 90)), 77, 0.7),
])
def test_quantized_matmul_against_reference_formula(params):
    m, n, k, fa, azp, sa, fb, bzp, sb = params
    a = np.array([fa(i) & 0xFF for i in range(m * k)], np.uint8)
    b = np.array([fb(i) & 0xFF for i in range(k * n)], np.uint8).view(np.int8)
    c = np.zeros(m * n, np.float32)
    oracle.quantized_matmul(m, n, k, a, azp, sa, b, bzp, sb, c, False)
    want = _naive_quantized(m, n, k, a, azp, sa, b, bzp, sb).ravel()
    assert np.array_equal(c, want)       # same integer, same two f32 roundings
    c0 = (np.arange(m * n) * 0.1).astype(np.float32)
    c = c0.copy()
    oracle.quantized_matmul(m, n, k, a, azp, sa, b, bzp, sb, c, True)
    assert np.array_equal(c, c0 + want)


# ---------------------------------------------------------------- halves

def test_half_widening_all_bit_patterns():
    bits = np.arange(65536, dtype=np.uint32)
    f16 = np.array([oracle.f16_to_f32(int(h)) for h in bits[::7]], np.float32)
    ref = bits[::7].astype(np.uint16).view(np.float16).astype(np.float32)
    fin = ~np.isnan(ref)
    assert np.array_equal(np.isnan(f16), ~fin)
    assert np.array_equal(f16[fin].view(np.uint32), ref[fin].view(np.uint32))
    bf = np.array([oracle.bf16_to_f32(int(h)) for h in bits[::97]], np.float32)
    refb = (bits[::97] << 16).astype(np.uint32).view(np.float32)
    # NaN payloads are quieted on the way back through ctypes; compare the rest bitwise
    fin = ~np.isnan(refb)
    assert np.array_equal(np.isnan(bf), ~fin)
    assert np.array_equal(bf[fin].view(np.uint32), refb[fin].view(np.uint32))


def test_bf16_gemm_is_widen_then_sgemm():
    r = rng(6)
    m, n, k = 9, 10, 300
    a = r.standard_normal(m * k).astype(np.float32).view(np.uint32) >> 16
    b = r.standard_normal(k * n).astype(np.float32).view(np.uint32) >> 16
    a16, b16 = a.astype(np.uint16), b.astype(np.uint16)
    c = np.zeros(m * n, np.float32)
    oracle.gemm_bf16_f32(m, n, k, a16, b16, c, False)
    aw = (a16.astype(np.uint32) << 16).view(np.float32)
    bw = (b16.astype(np.uint32) << 16).view(np.float32)
    c2 = np.zeros(m * n, np.float32)
    oracle.gemm_f32(m, n, k, 1.0, aw, bw, 0.0, c2)
    assert np.array_equal(c, c2)
    ref = aw.reshape(m, k).astype(np.float64) @ bw.reshape(k, n).astype(np.float64)
    assert np.all(np.abs(c.reshape(m, n) - ref) <= 1e-5 * np.maximum(np.abs(ref), 1.0))


# ---------------------------------------------------------------- fp8

def test_fp8_decode_matches_the_ocp_spec_anchors():
    # values the OCP FP8 spec tabulates for each format
    e4 = {0x00: 0.0, 0x38: 1.0, 0xb8: -1.0, 0x7e: 448.0, 0x08: 2.0 ** -6, 0x07: 7 * 2.0 ** -9,
          0x01: 2.0 ** -9, 0x3c: 1.5, 0x40: 2.0}
    for x, v in e4.items():
        assert oracle.e4m3_to_f64(x) == v, hex(x)
    assert np.isnan(oracle.e4m3_to_f64(0x7f)) and np.isnan(oracle.e4m3_to_f64(0xff))
    e5 = {0x00: 0.0, 0x3c: 1.0, 0xbc: -1.0, 0x7b: 57344.0, 0x04: 2.0 ** -14, 0x01: 2.0 ** -16,
          0x03: 3 * 2.0 ** -16, 0x3e: 1.5}
    for x, v in e5.items():
        assert oracle.e5m2_to_f64(x) == v, hex(x)
    assert oracle.e5m2_to_f64(0x7c) == np.inf and oracle.e5m2_to_f64(0xfc) == -np.inf
    assert all(np.isnan(oracle.e5m2_to_f64(x)) for x in (0x7d, 0x7e, 0x7f, 0xfd))


def test_fp8_decode_matches_torch_on_every_code():
    torch = pytest.importorskip("torch")
    codes = np.arange(256, dtype=np.uint8)
    for fmt, dt, dec in (("e4m3", torch.float8_e4m3fn, oracle.e4m3_to_f64),
                         ("e5m2", torch.float8_e5m2, oracle.e5m2_to_f64)):
        ref = torch.from_numpy(codes).view(dt).to(torch.float64).numpy()
        mine = np.array([dec(int(x)) for x in codes])
        nan = np.isnan(ref)
        assert np.array_equal(np.isnan(mine), nan), fmt
        assert np.array_equal(mine[~nan], ref[~nan]), fmt


@pytest.mark.parametrize("fmt", ["e4m3", "e5m2"])
def test_fp8_gemm_is_decode_then_exact_product(fmt):
    r = rng(8)
    m, n, k = 6, 7, 50
    dec = oracle.e4m3_to_f64 if fmt == "e4m3" else oracle.e5m2_to_f64
    lut = np.array([dec(i) for i in range(256)])
    finite = np.flatnonzero(np.isfinite(lut)).astype(np.uint8)
    a = r.choice(finite, m * k)
    b = r.choice(finite, k * n)
    c = np.zeros(m * n)
    oracle.gemm_fp8_f64(m, n, k, a, b, fmt, c, False)
    A = lut[a].reshape(m, k); B = lut[b].reshape(k, n)
    exact = [[sum(Fraction(x) * Fraction(y) for x, y in zip(A[i], B[:, j])) for j in range(n)]
             for i in range(m)]
    scale = np.abs(A) @ np.abs(B)
    for i in range(m):
        for j in range(n):
            assert abs(Fraction(c[i * n + j]) - exact[i][j]) <= Fraction(k * 2.0 ** -53 * scale[i, j])
    c2 = np.full(m * n, 0.5)
    oracle.gemm_fp8_f64(m, n, k, a, b, fmt, c2, True)
    assert np.array_equal(c2, c + 0.5) or np.allclose(c2, c + 0.5, rtol=1e-15, atol=0)


# ---------------------------------------------------------------- dd reference

def test_double_double_reference_is_correctly_rounded():
    r = rng(7)
    m, n, k = 4, 3, 200
    a = r.standard_normal(m * k) * np.exp2(r.integers(-30, 30, m * k))
    b = r.standard_normal(k * n) * np.exp2(r.integers(-30, 30, k * n))
    hi, lo = oracle.dd_gemm_f64_sampled(m, n, k, a, b, np.arange(m), np.arange(n))
    A = a.reshape(m, k); B = b.reshape(k, n)
    for i in range(m):
        for j in range(n):
            ex = sum(Fraction(x) * Fraction(y) for x, y in zip(A[i], B[:, j]))
            err = abs(Fraction(hi[i, j]) - ex)
            assert err <= abs(ex) * Fraction(2) ** -52 + Fraction(2) ** -1000


# ------------------------------------------------------------------ trsm.rs

def _ref_lower(m):
    """make_lower_triangular of trsm.rs's tests."""
    a = np.zeros((m, m))
    for i in range(m):
        for j in range(i + 1):
            a[i, j] = 2.0 + i * 0.5 if i == j else 0.1 * (i + j + 1)
    return a


def _ref_upper(m):
    a = np.zeros((m, m))
    for i in range(m):
        for j in range(i, m):
            a[i, j] = 2.0 + i * 0.5 if i == j else 0.1 * (i + j + 1)
    return a


@pytest.mark.parametrize("lower,m,n,tol,gen", [
    (True, 5, 3, 1e-10, lambda i: i * 0.7 + 1.0),                 # trsm_lln_f64_small
    (True, 80, 7, 1e-9, lambda i: np.sin(i * 0.13) + 1.0),        # trsm_lln_f64_large_blocked
    (False, 5, 3, 1e-10, lambda i: i * 0.7 + 1.0),                # trsm_lun_f64_small
    (False, 80, 7, 1e-9, lambda i: np.sin(i * 0.13) + 1.0),       # trsm_lun_f64_large_blocked
])
def test_trsm_reference_cases(lower, m, n, tol, gen):
    a = _ref_lower(m) if lower else _ref_upper(m)
    b = np.array([gen(i) for i in range(m * n)], dtype=np.float64)
    x = b.copy()
    oracle.trsm(lower, m, n, 1.0, a, m, x, n, False)
    assert np.max(np.abs(a @ x.reshape(m, n) - b.reshape(m, n))) < tol


def test_trsm_reference_alpha_unit_and_f32():
    a = _ref_lower(4)
    b = np.arange(8, dtype=np.float64) + 1.0
    x = b.copy()
    oracle.trsm(True, 4, 2, 3.0, a, 4, x, 2, False)                 # trsm_alpha_scales_b
    assert np.max(np.abs(a @ x.reshape(4, 2) - 3.0 * b.reshape(4, 2))) < 1e-10
    a = _ref_lower(6)
    np.fill_diagonal(a, 1.0)
    b = np.arange(24) * 0.3 + 0.5
    xu, xe = b.copy(), b.copy()
    oracle.trsm(True, 6, 4, 1.0, a, 6, xu, 4, True)                 # trsm_lln_f64_unit_diag
    oracle.trsm(True, 6, 4, 1.0, a, 6, xe, 4, False)
    assert np.max(np.abs(xu - xe)) < 1e-12
    for lower, m, n in ((True, 65, 4), (False, 70, 3)):             # trsm_l{l,u}n_f32_blocked
        a = (_ref_lower(m) if lower else _ref_upper(m)).astype(np.float32)
        b = ((np.arange(m * n) * (0.13 if lower else 0.07)) + (0.5 if lower else -0.5)).astype(np.float32)
        x = b.copy()
        oracle.trsm(lower, m, n, np.float32(1.0), a, m, x, n, False)
        assert np.max(np.abs(a @ x.reshape(m, n) - b.reshape(m, n))) < 1e-3


def test_trsm_exact_on_integer_systems():
    r = rng(12)
    m, n = 70, 5
    for lower in (True, False):
        a = np.tril(r.integers(-3, 4, (m, m))).astype(np.float64) if lower else \
            np.triu(r.integers(-3, 4, (m, m))).astype(np.float64)
        np.fill_diagonal(a, 1.0)
        xs = r.integers(-5, 6, (m, n)).astype(np.float64)
        # keep |B| small enough that every intermediate is an exact integer
        b = a @ xs
        if np.max(np.abs(b)) > 2 ** 40:
            continue
        x = b.ravel().copy()
        oracle.trsm(lower, m, n, 1.0, a, m, x, n, True)
        assert np.array_equal(x.reshape(m, n), xs)


def test_trsm_lower_recursion_keeps_the_substitution_order():
    # Per element the blocked lower solve subtracts a_ip x_p for p ascending
    # and then scales, exactly like one unblocked sweep, so the recursion
    # (split at multiples of 32, scalar updates) must not change any bit.
    r = rng(13)
    m, n = 150, 6
    a = np.tril(r.uniform(-1, 1, (m, m))) + np.diag(np.full(m, 3.0))
    b = r.standard_normal(m * n)
    x = b.copy()
    oracle.trsm(True, m, n, 1.0, a, m, x, n, False)
    y = b.reshape(m, n).copy()
    for i in range(m):
        for p in range(i):
            y[i] = y[i] - a[i, p] * y[p]
        y[i] = y[i] * (1.0 / a[i, i])
    assert x.reshape(m, n).tobytes() == y.tobytes()


# ------------------------------------------- helpers the config-size checks use

def test_dd_complex_reference_meets_the_dot2_bound():
    """orc_dd_gemm_c64_sampled against exact rational sums.  Dot2 (Ogita,
    Rump, Oishi 2005, Prop. 5.5) returns s with |s - x.y| <= u|x.y| +
    gamma_n^2 |x|.|y|, here per part over n = 2k products.  Inputs span
    2^+-30 so the lo parts matter; a sign slip in either cross term (re uses
    -xi yi, im uses +xi yr) is off by 2|xi yi| and fails."""
    r = rng(21)
    m, n, k = 3, 4, 150
    sc = lambda s: np.exp2(r.integers(-30, 30, s))  # noqa: E731
    A = (r.standard_normal((m, k)) * sc((m, k)) + 1j * r.standard_normal((m, k)) * sc((m, k)))
    B = (r.standard_normal((k, n)) * sc((k, n)) + 1j * r.standard_normal((k, n)) * sc((k, n)))
    out = oracle.dd_gemm_c64_sampled(m, n, k, A.view(np.float64).ravel(),
                                     B.view(np.float64).ravel(), np.arange(m), np.arange(n))
    u = Fraction(2) ** -53
    g = 2 * k * u / (1 - 2 * k * u)
    for i in range(m):
        for j in range(n):
            xs = [(Fraction(A[i, p].real), Fraction(A[i, p].imag)) for p in range(k)]
            ys = [(Fraction(B[p, j].real), Fraction(B[p, j].imag)) for p in range(k)]
            re = sum(xr * yr - xi * yi for (xr, xi), (yr, yi) in zip(xs, ys))
            im = sum(xr * yi + xi * yr for (xr, xi), (yr, yi) in zip(xs, ys))
            mag = sum(abs(xr) * abs(yr) + abs(xi) * abs(yi) + abs(xr) * abs(yi) + abs(xi) * abs(yr)
                      for (xr, xi), (yr, yi) in zip(xs, ys))
            for got, ex in ((out[i, j].real, re), (out[i, j].imag, im)):
                assert abs(Fraction(got) - ex) <= u * abs(ex) + g * g * mag + Fraction(2) ** -1000


def test_dd_complex_reference_on_exact_cancellation():
    """Small integers: every sum is exact, so the result must equal numpy's
    integer product exactly, including rows that cancel to zero."""
    r = rng(22)
    m, n, k = 5, 6, 64
    A = r.integers(-7, 8, (m, k)) + 1j * r.integers(-7, 8, (m, k))
    B = r.integers(-7, 8, (k, n)) + 1j * r.integers(-7, 8, (k, n))
    A[2] = 1j * A[1]          # row 2 = i row 1: C[2] = i C[1]
    out = oracle.dd_gemm_c64_sampled(m, n, k, A.view(np.float64).ravel(),
                                     B.view(np.float64).ravel(), np.arange(m), np.arange(n))
    assert np.array_equal(out, A @ B)
    assert np.array_equal(out[2], 1j * out[1])


@pytest.mark.parametrize("cplx", [False, True])
def test_dd_reference_row_split_is_invisible(cplx):
    """The rows are split over threads; any subset of rows, asked for alone,
    gives the same bits as in one call (the arithmetic per element is one
    sequential loop)."""
    r = rng(23)
    m, n, k = 37, 9, 333
    w = 2 if cplx else 1
    a = r.standard_normal(m * k * w); b = r.standard_normal(k * n * w)
    fn = oracle.dd_gemm_c64_sampled if cplx else (lambda *x: oracle.dd_gemm_f64_sampled(*x)[0])
    rows = np.arange(m); cols = np.array([0, 3, 8])
    full = fn(m, n, k, a, b, rows, cols)
    for i in (0, 17, 36):
        one = fn(m, n, k, a, b, np.array([i]), cols)
        assert one.tobytes() == full[i:i + 1].tobytes()


@pytest.mark.parametrize("m,n,k,workers", [(1, 5, 7, 8), (3, 1, 300, 8), (17, 13, 9, 4),
                                           (64, 33, 129, 0), (9, 4, 0, 3)])
def test_i8_signed_threaded_is_the_closed_form(m, n, k, workers):
    """The threaded integer product splits rows over workers (fewer rows
    than workers, one column, k = 0 included) and equals orc_gemm_i8_signed
    and numpy's exact product mod 2^32."""
    r = rng(24)
    a = r.integers(-128, 128, max(m * k, 1)).astype(np.int8)
    b = r.integers(-128, 128, max(k * n, 1)).astype(np.int8)
    c1 = np.full(m * n, 7, np.int32); c2 = np.zeros(m * n, np.int32)
    oracle.gemm_i8_signed_threaded(m, n, k, a, b, c1, workers)
    oracle.gemm_i8_signed(m, n, k, a, b, c2, False)
    assert np.array_equal(c1, c2)
    want = wrap_i32(exact_int_matmul(a[:m * k].reshape(m, k), b[:k * n].reshape(k, n))).ravel()
    assert np.array_equal(c1, want)


def test_i8_signed_threaded_wraps():
    """k = 140000 products of -128 * -128 = 2^14 sum to 2293760000 > 2^31 - 1:
    the output is that value mod 2^32, not a saturated i32."""
    m, n, k = 2, 3, 140000
    a = np.full(m * k, -128, np.int8); b = np.full(k * n, -128, np.int8)
    c = np.zeros(m * n, np.int32)
    oracle.gemm_i8_signed_threaded(m, n, k, a, b, c, 2)
    assert np.all(c == np.int64(k * 2 ** 14 - 2 ** 32))


def test_gemm_c32_shapes_alpha_beta_and_blocking():
    """The c64 checks repeated on the f32 restatement: non-square with k past
    one KC block, complex alpha / beta against numpy in double within the
    f32 accumulation bound, exact small-integer products, k == 0 with
    complex beta, and the same per-block formula as c64 (on f32-exact inputs
    whose partial sums are exact in f32 the two agree bit for bit)."""
    r = rng(25)
    m, n, k = 13, 7, 300
    A = (r.standard_normal((m, k)) + 1j * r.standard_normal((m, k))).astype(np.complex64)
    B = (r.standard_normal((k, n)) + 1j * r.standard_normal((k, n))).astype(np.complex64)
    C0 = (r.standard_normal((m, n)) + 1j * r.standard_normal((m, n))).astype(np.complex64)
    al, be = np.complex64(0.5 - 2j), np.complex64(-1.5 + 0.25j)
    c = C0.copy()
    oracle.gemm_c32(m, n, k, al.real, al.imag, A.view(np.float32).ravel(),
                    B.view(np.float32).ravel(), be.real, be.imag, c.view(np.float32).ravel())
    A64, B64, C64 = A.astype(np.complex128), B.astype(np.complex128), C0.astype(np.complex128)
    ref = complex(al) * (A64 @ B64) + complex(be) * C64
    mag = abs(complex(al)) * (np.abs(A64.real) + np.abs(A64.imag)) @ (np.abs(B64.real) +
                                                                     np.abs(B64.imag))
    bound = 2 * (k + 4) * 2.0 ** -24 * (mag + abs(complex(be)) * np.abs(C64))
    assert np.all(np.abs(c - ref) <= bound)
    # a transposed or conjugated operand is far outside the bound
    assert np.max(np.abs(c - (complex(al) * (A64 @ B64.conj()) + complex(be) * C64))) > 10 * bound.max()
    # exact small-integer products
    Ai = (r.integers(-4, 5, (m, k)) + 1j * r.integers(-4, 5, (m, k))).astype(np.complex64)
    Bi = (r.integers(-4, 5, (k, n)) + 1j * r.integers(-4, 5, (k, n))).astype(np.complex64)
    ci = np.zeros((m, n), np.complex64)
    oracle.gemm_c32(m, n, k, 1.0, 0.0, Ai.view(np.float32).ravel(), Bi.view(np.float32).ravel(),
                    0.0, 0.0, ci.view(np.float32).ravel())
    assert np.array_equal(ci, (Ai.astype(np.complex128) @ Bi.astype(np.complex128)))
    c64 = np.zeros((m, n), np.complex128)
    oracle.gemm_c64(m, n, k, 1.0, 0.0, Ai.astype(np.complex128).view(np.float64).ravel(),
                    Bi.astype(np.complex128).view(np.float64).ravel(), 0.0, 0.0,
                    c64.view(np.float64).ravel())
    assert np.array_equal(ci.astype(np.complex128), c64)
    # k == 0 with complex beta (outer.rs:1502-1519)
    c = C0.copy()
    oracle.gemm_c32(m, n, 0, 1.0, 0.0, np.zeros(2, np.float32), np.zeros(2, np.float32),
                    be.real, be.imag, c.view(np.float32).ravel())
    assert np.allclose(c, complex(be) * C64, rtol=2e-7, atol=0)


def test_gemm_c32_alpha_applies_per_block():
    """With alpha = 1 + 0i the product is unscaled; with alpha = 3 each KC
    block is scaled before it is added (k = 257 = one block + one element),
    which on integers is still exact: alpha*(A B)."""
    r = rng(26)
    m, n, k = 4, 5, 257
    A = (r.integers(-3, 4, (m, k)) + 1j * r.integers(-3, 4, (m, k))).astype(np.complex64)
    B = (r.integers(-3, 4, (k, n)) + 1j * r.integers(-3, 4, (k, n))).astype(np.complex64)
    c = np.zeros((m, n), np.complex64)
    oracle.gemm_c32(m, n, k, 3.0, 0.0, A.view(np.float32).ravel(), B.view(np.float32).ravel(),
                    0.0, 0.0, c.view(np.float32).ravel())
    assert np.array_equal(c, 3 * (A.astype(np.complex128) @ B.astype(np.complex128)))


def test_f16_gemm_is_widen_then_sgemm():
    """gemm_f16_f32 (mod.rs:906-933) = widen every binary16 to f32, then the
    f32 block formula; numpy's float16 -> float32 conversion does the
    widening independently."""
    r = rng(27)
    m, n, k = 9, 10, 300
    a16 = r.standard_normal(m * k).astype(np.float16)
    b16 = r.standard_normal(k * n).astype(np.float16)
    c = np.zeros(m * n, np.float32)
    oracle.gemm_f16_f32(m, n, k, a16.view(np.uint16), b16.view(np.uint16), c, False)
    aw, bw = a16.astype(np.float32), b16.astype(np.float32)
    c2 = np.zeros(m * n, np.float32)
    oracle.gemm_f32(m, n, k, 1.0, aw, bw, 0.0, c2)
    assert np.array_equal(c, c2)
    c3 = np.full(m * n, 0.25, np.float32)
    oracle.gemm_f16_f32(m, n, k, a16.view(np.uint16), b16.view(np.uint16), c3, True)
    c4 = np.full(m * n, 0.25, np.float32)
    oracle.gemm_f32(m, n, k, 1.0, aw, bw, 1.0, c4)
    assert np.array_equal(c3, c4)


# ------------------------------------------------------------------ generators

def _splitmix_py(state: int, count: int):
    """SplitMix64 as published by Steele, Lea and Flood (2014) / Vigna's
    splitmix64.c: state += 0x9E3779B97F4A7C15, then the two xor-shift-
    multiply rounds and a final xor-shift."""
    out = []
    for _ in range(count):
        state = (state + 0x9E3779B97F4A7C15) & (2 ** 64 - 1)
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2 ** 64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2 ** 64 - 1)
        out.append(z ^ (z >> 31))
    return out


def test_splitmix64_is_the_published_generator():
    # first outputs of the generator started at state 0
    assert _splitmix_py(0, 3) == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    for seed in (0, 1, 7, 2 ** 63 + 5):
        want = _splitmix_py(seed, 50)
        assert [oracle.lib().orc_splitmix64(seed, i) for i in range(50)] == want


def test_uniform_generators():
    """uniform_f64: the top 53 bits of each draw, r * 2^-52 - 1, so exact
    multiples of 2^-52 in [-1, 1); uniform_int: lo + floor(r32 * span / 2^32)
    of the draw's top 32 bits."""
    draws = _splitmix_py(9, 4000)
    f = oracle.uniform_f64(4000, 9)
    assert f.tolist() == [(d >> 11) * 2.0 ** -52 - 1.0 for d in draws]
    assert f.min() >= -1.0 and f.max() < 1.0
    i = oracle.uniform_int(4000, 9, -128, 127)
    assert i.tolist() == [-128 + (((d >> 32) * 256) >> 32) for d in draws]
    assert i.min() == -128 and i.max() == 127
    j = oracle.uniform_int(4000, 9, -32768, 32767)
    assert j.min() < -32000 and j.max() > 32000


def test_cpu_feature_probe_matches_the_kernel_flags():
    flags = set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    flags = set(line.split(":", 1)[1].split())
                    break
    except OSError:
        pytest.skip("no /proc/cpuinfo")
    assert oracle.cpu_has_avx2_fma() == ({"avx2", "fma"} <= flags)
