"""Split-precision DGEMM (integer-modular Ozaki scheme on INT8 MFMA).

Bars:
  * exactness: products whose operands scale to small integers are
    reproduced exactly (residue GEMMs and Chinese remaindering are exact);
  * accuracy: normwise error ||C^ - C||_F / (||A||_F ||B||_F) <= 1e-15 against
    the compensated (double-double) oracle with the default 16 moduli, and
    elementwise |C^ - C| <= 2^(6-g) * ||a_i|| * ||b_j|| where g is the scaling
    exponent of the modulus count (rounding each operand row/column to
    integers on a 2^-g grid relative to its 2-norm);
  * error decreases as moduli are added;
  * the native entry's edge semantics (k == 0, beta == 0 ignores C, alpha /
    beta) and NaN for rows/columns holding non-finite values.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from tests.helpers import rng

pytestmark = pytest.mark.gpu

MODULI = [256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197, 193]


def gamma(nmod: int) -> int:
    M = math.prod(MODULI[:nmod])
    return (M.bit_length() - 1 - 2) // 2


def dd_ref(a, b, m, n, k):
    hi, lo = oracle.dd_gemm_f64_sampled(m, n, k, a, b, np.arange(m), np.arange(n))
    return hi, lo


def test_exact_on_integer_operands(gpu):
    r = rng(1)
    m, n, k = 97, 130, 300
    a = r.integers(-1000, 1001, m * k).astype(np.float64) * 2.0 ** -7
    b = r.integers(-1000, 1001, k * n).astype(np.float64) * 2.0 ** 5
    c = np.zeros(m * n)
    assert gpu.dgemm_ozaki(m, n, k, 1.0, a, b, 0.0, c, 16), gpu.last_status()
    exact = (a.reshape(m, k) @ b.reshape(k, n))   # all partial sums exact in f64
    assert np.array_equal(c.reshape(m, n), exact)


@pytest.mark.parametrize("shape", [(1, 1, 1), (7, 5, 3), (64, 64, 64), (256, 256, 256),
                                   (300, 257, 700), (513, 260, 1024)])
def test_accuracy_normal(gpu, shape):
    m, n, k = shape
    r = rng(sum(shape))
    a = r.standard_normal(m * k); b = r.standard_normal(k * n)
    c = np.zeros(m * n)
    assert gpu.dgemm_ozaki(m, n, k, 1.0, a, b, 0.0, c, 16), gpu.last_status()
    hi, lo = dd_ref(a, b, m, n, k)
    d = (c.reshape(m, n) - hi) - lo
    nab = np.linalg.norm(a) * np.linalg.norm(b)
    assert np.linalg.norm(d) / nab <= 1e-15
    ra = np.linalg.norm(a.reshape(m, k), axis=1)[:, None]
    cb = np.linalg.norm(b.reshape(k, n), axis=0)[None, :]
    assert np.all(np.abs(d) <= 2.0 ** (6 - gamma(16)) * ra * cb + 4 * 2.0 ** -53 * np.abs(hi))


def test_accuracy_wide_range_and_moduli_sweep(gpu):
    m, n, k = 128, 160, 512
    r = rng(7)
    a = r.uniform(-1, 1, m * k) * np.exp2(r.integers(-20, 21, m * k))
    b = r.uniform(-1, 1, k * n) * np.exp2(r.integers(-20, 21, k * n))
    hi, lo = dd_ref(a, b, m, n, k)
    nab = np.linalg.norm(a) * np.linalg.norm(b)
    errs = []
    for nm in (8, 10, 12, 14, 16):
        c = np.zeros(m * n)
        assert gpu.dgemm_ozaki(m, n, k, 1.0, a, b, 0.0, c, nm)
        errs.append(np.linalg.norm((c.reshape(m, n) - hi) - lo) / nab)
    assert errs[-1] <= 1e-15, errs
    assert all(errs[i + 1] <= errs[i] * 1.01 for i in range(len(errs) - 1)), errs


def test_alpha_beta_and_edges(gpu):
    r = rng(9)
    m, n, k = 70, 90, 110
    a = r.standard_normal(m * k); b = r.standard_normal(k * n); c0 = r.standard_normal(m * n)
    c = c0.copy()
    assert gpu.dgemm_ozaki(m, n, k, -2.5, a, b, 0.75, c, 16)
    ref = -2.5 * (a.reshape(m, k) @ b.reshape(k, n)) + 0.75 * c0.reshape(m, n)
    assert np.max(np.abs(c.reshape(m, n) - ref)) < 1e-12
    # beta == 0 never reads C
    c = np.full(m * n, np.nan)
    assert gpu.dgemm_ozaki(m, n, k, 1.0, a, b, 0.0, c, 16)
    assert np.all(np.isfinite(c))
    # k == 0: C = beta * C
    c = c0.copy()
    assert gpu.dgemm_ozaki(m, n, 0, 1.0, np.zeros(1), np.zeros(1), 2.0, c, 16)
    assert np.array_equal(c, 2.0 * c0)
    # zero rows / columns give zeros
    a2 = a.copy().reshape(m, k); a2[3] = 0.0
    c = np.zeros(m * n)
    assert gpu.dgemm_ozaki(m, n, k, 1.0, a2.ravel(), b, 0.0, c, 16)
    assert np.all(c.reshape(m, n)[3] == 0.0)


def test_nonfinite_rows_and_columns_give_nan(gpu):
    r = rng(10)
    m, n, k = 40, 50, 60
    a = r.standard_normal((m, k)); b = r.standard_normal((k, n))
    a[5, 7] = np.inf
    b[3, 11] = np.nan
    c = np.zeros(m * n)
    assert gpu.dgemm_ozaki(m, n, k, 1.0, a.ravel(), b.ravel(), 0.0, c, 16)
    c = c.reshape(m, n)
    assert np.all(np.isnan(c[5])) and np.all(np.isnan(c[:, 11]))
    mask = np.ones((m, n), bool); mask[5] = False; mask[:, 11] = False
    assert np.all(np.isfinite(c[mask]))


def test_invalid_arguments(gpu):
    c = np.zeros(4)
    assert not gpu.dgemm_ozaki(2, 2, 2, 1.0, np.ones(4), np.ones(4), 0.0, c, 17)
    assert not gpu.dgemm_ozaki(2, 2, 2, 1.0, np.ones(4), np.ones(4), 0.0, c, 1)
