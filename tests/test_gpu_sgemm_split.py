"""SGEMM on BF16 matrix cores by operand splitting (terms 3 and 6).

No reference counterpart exists (ferray's SGEMM is plain f32), so the bars
come from the arithmetic:
  * operands exactly representable in bf16 with small integer values give
    the exact integer product (every piece but p0 is zero and every partial
    sum is an integer below 2^24);
  * for general f32 operands the result is within the piece-sum bound
        |C - AB| <= (d_T + (T*k + 2) * 2^-24) * (|A||B|)   elementwise,
    d_3 = 3.1 * 2^-16 (the dropped p1p1 and p2 terms), d_6 = 2.1 * 2^-24;
  * the result matches the exact sum of the selected piece products (pieces
    computed independently here with numpy) to f32 accumulation accuracy,
    which pins which piece products are summed.
"""
from __future__ import annotations

import numpy as np
import pytest

import ferray_amd as fa
from tests.helpers import rng

pytestmark = pytest.mark.gpu
U = 2.0 ** -24
D = {3: 3.1 * 2.0 ** -16, 6: 2.1 * 2.0 ** -24}
PAIRS = {3: [(0, 0), (0, 1), (1, 0)], 6: [(0, 0), (0, 1), (1, 0), (0, 2), (1, 1), (2, 0)]}


def bf16_round(x: np.ndarray) -> np.ndarray:
    """Round f32 to bf16 (nearest, ties to even), returned as f32 values."""
    u = x.astype(np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def pieces(x: np.ndarray):
    p0 = bf16_round(x)
    r1 = (x - p0).astype(np.float32)
    p1 = bf16_round(r1)
    p2 = bf16_round((r1 - p1).astype(np.float32))
    return [p0, p1, p2]


def _run(m, n, k, a, b, terms, alpha=1.0, beta=0.0, c0=None):
    c = np.zeros(m * n, np.float32) if c0 is None else c0.copy()
    assert fa.sgemm_bf16x3(m, n, k, np.float32(alpha), a, b, np.float32(beta), c, terms), \
        fa.last_status()
    return c.reshape(m, n)


def test_pieces_restatement_is_exact():
    x = rng(0).standard_normal(100000).astype(np.float32)
    p = pieces(x)
    assert np.array_equal(p[0].astype(np.float64) + p[1] + p[2], x.astype(np.float64))


@pytest.mark.parametrize("terms", [3, 6])
@pytest.mark.parametrize("shape", [(1, 1, 1), (7, 5, 3), (129, 130, 77), (300, 257, 64),
                                   (256, 256, 200)])
def test_exact_on_small_integers(gpu, terms, shape):
    m, n, k = shape
    r = rng(1)
    a = r.integers(-16, 17, m * k).astype(np.float32)
    b = r.integers(-16, 17, k * n).astype(np.float32)
    got = _run(m, n, k, a, b, terms)
    ref = a.reshape(m, k).astype(np.int64) @ b.reshape(k, n).astype(np.int64)
    assert np.array_equal(got, ref.astype(np.float32))


@pytest.mark.parametrize("terms", [3, 6])
@pytest.mark.parametrize("shape", [(256, 256, 512), (100, 37, 1000), (33, 300, 64)])
def test_error_bound(gpu, terms, shape):
    m, n, k = shape
    r = rng(2)
    a = r.standard_normal(m * k).astype(np.float32)
    b = r.standard_normal(k * n).astype(np.float32)
    got = _run(m, n, k, a, b, terms).astype(np.float64)
    A, B = a.reshape(m, k).astype(np.float64), b.reshape(k, n).astype(np.float64)
    mag = np.abs(A) @ np.abs(B)
    assert np.all(np.abs(got - A @ B) <= (D[terms] + (terms * k + 2) * U) * mag)


def test_six_terms_reach_f32_accuracy(gpu):
    m = n = k = 256
    r = rng(3)
    a = r.standard_normal(m * k).astype(np.float32)
    b = r.standard_normal(k * n).astype(np.float32)
    ref = a.reshape(m, k).astype(np.float64) @ b.reshape(k, n).astype(np.float64)
    c_native = np.zeros(m * n, np.float32)
    assert fa.gemm_f32(m, n, k, np.float32(1), a, b, np.float32(0), c_native)
    e = {t: np.linalg.norm(_run(m, n, k, a, b, t) - ref) for t in (3, 6)}
    e_native = np.linalg.norm(c_native.reshape(m, n) - ref)
    assert e[6] < e[3] / 5
    assert e[6] < 4 * e_native


@pytest.mark.parametrize("terms", [3, 6])
def test_matches_selected_piece_products(gpu, terms):
    m, n, k = 64, 48, 4
    r = rng(4)
    a = r.standard_normal(m * k).astype(np.float32)
    b = r.standard_normal(k * n).astype(np.float32)
    pa = [p.reshape(m, k).astype(np.float64) for p in pieces(a)]
    pb = [p.reshape(k, n).astype(np.float64) for p in pieces(b)]
    exact = sum(pa[i] @ pb[j] for i, j in PAIRS[terms])
    mag = sum(np.abs(pa[i]) @ np.abs(pb[j]) for i, j in PAIRS[terms])
    got = _run(m, n, k, a, b, terms).astype(np.float64)
    assert np.all(np.abs(got - exact) <= (terms * k + 2) * U * mag + 1e-30)


@pytest.mark.parametrize("terms", [3, 6])
def test_alpha_beta_and_k0(gpu, terms):
    m, n, k = 130, 70, 90
    r = rng(5)
    a = r.standard_normal(m * k).astype(np.float32)
    b = r.standard_normal(k * n).astype(np.float32)
    c0 = r.standard_normal(m * n).astype(np.float32)
    got = _run(m, n, k, a, b, terms, alpha=-0.5, beta=2.0, c0=c0).astype(np.float64)
    A, B = a.reshape(m, k).astype(np.float64), b.reshape(k, n).astype(np.float64)
    ref = -0.5 * (A @ B) + 2.0 * c0.reshape(m, n)
    mag = 0.5 * (np.abs(A) @ np.abs(B)) + 2.0 * np.abs(c0.reshape(m, n))
    assert np.all(np.abs(got - ref) <= (D[terms] + (terms * k + 4) * U) * mag)
    k0 = _run(m, n, 0, a, b, terms, alpha=1.0, beta=3.0, c0=c0)
    assert np.array_equal(k0, (np.float32(3.0) * c0).reshape(m, n))


def test_invalid_terms(gpu):
    c = np.zeros(4, np.float32)
    a = np.zeros(4, np.float32)
    assert not fa.sgemm_bf16x3(2, 2, 2, np.float32(1), a, a, np.float32(0), c, 4)
