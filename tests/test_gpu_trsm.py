"""Triangular solve on the device against the oracle's restatement of trsm.rs.

Bars:
  * solves that fit one leaf (lower m <= 64, upper m <= 32) are bit-identical
    to the reference's scalar substitution (same operation order), including
    the lower leaf's split of each column over 4 or 8 threads;
  * larger solves (the trailing updates run on MFMA GEMMs) agree with the
    oracle to 1e-10 relative (f64) / 2e-4 (f32) on well-conditioned systems,
    and their normwise backward error ||A X - alpha B|| / (||A|| ||X||) is
    below 4 m u;
  * unit_diag, alpha, row strides lda > m and ldb > n behave as the
    reference's entry points (the padding of B is left untouched).
"""
from __future__ import annotations

import numpy as np
import pytest

import ferray_amd as fa
import oracle
from tests.helpers import rng

pytestmark = pytest.mark.gpu


def _system(lower, m, n, dt, seed):
    r = rng(seed)
    a = r.uniform(-1, 1, (m, m)) / max(1, m) ** 0.5
    a = np.tril(a) if lower else np.triu(a)
    np.fill_diagonal(a, 2.0 + r.uniform(0, 1, m))
    b = r.standard_normal((m, n))
    return a.astype(dt), b.astype(dt)


def _gpu(lower, m, n, alpha, a, lda, b, ldb, unit):
    f = {(True, np.float64): fa.trsm_left_lower_f64, (False, np.float64): fa.trsm_left_upper_f64,
         (True, np.float32): fa.trsm_left_lower_f32, (False, np.float32): fa.trsm_left_upper_f32}
    assert f[(lower, a.dtype.type)](m, n, a.dtype.type(alpha), a, lda, b, ldb, unit), \
        fa.last_status()


# Single-leaf solves whose operation order is the reference's: every lower
# solve up to 64 rows (n >= 16384 runs the lower leaf with 4 threads per
# column, smaller n with 8), upper solves up to 32 rows (above that the
# reference splits at 32 while the device leaf runs on to 64 rows).
LEAF_EXACT = ([(True, m, n) for m, n in [(1, 1), (5, 3), (32, 9), (64, 300), (47, 16400),
                                         (64, 16500)]] +
              [(False, m, n) for m, n in [(1, 1), (5, 3), (32, 9), (17, 16400)]])


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("lower,m,n", LEAF_EXACT)
def test_single_leaf_bit_exact(gpu, lower, dt, m, n):
    a, b = _system(lower, m, n, dt, 1)
    x_gpu = b.ravel().copy()
    x_orc = b.ravel().copy()
    _gpu(lower, m, n, 1.0, a.ravel(), m, x_gpu, n, False)
    oracle.trsm(lower, m, n, dt(1.0), a.ravel(), m, x_orc, n, False)
    assert x_gpu.tobytes() == x_orc.tobytes()


# Solves whose order differs from the reference's substitution: the blocked
# recursion (m > 64), and upper single-leaf solves with 33 <= m <= 64, which
# the reference splits at 32 rows and the device solves in one leaf.
# Diagonal blocks of 65 to 256 rows run as one kernel (16-row leaves, 16
# columns per workgroup): whole and ragged leaves and column groups.
BLOCKED = ([(lower, m, n) for lower in (True, False)
            for m, n in [(65, 4), (80, 7), (300, 50), (1000, 33), (517, 1), (97, 3), (129, 16),
                         (255, 40), (256, 1000), (2048, 700)]] +
           [(False, m, n) for m, n in [(33, 5), (40, 300), (64, 16500), (64, 1)]])


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("lower,m,n", BLOCKED)
def test_blocked_against_oracle(gpu, lower, dt, m, n):
    a, b = _system(lower, m, n, dt, 2)
    x_gpu = b.ravel().copy()
    x_orc = b.ravel().copy()
    _gpu(lower, m, n, 1.0, a.ravel(), m, x_gpu, n, False)
    oracle.trsm(lower, m, n, dt(1.0), a.ravel(), m, x_orc, n, False)
    tol = 1e-10 if dt == np.float64 else 2e-4
    X = x_gpu.reshape(m, n).astype(np.float64)
    assert np.linalg.norm(X - x_orc.reshape(m, n)) <= tol * np.linalg.norm(x_orc)
    u = 2.0 ** -53 if dt == np.float64 else 2.0 ** -24
    A = a.astype(np.float64)
    back = np.linalg.norm(A @ X - b.astype(np.float64)) / (np.linalg.norm(A) * np.linalg.norm(X))
    assert back <= 4 * m * u


@pytest.mark.parametrize("lower", [True, False])
def test_reference_test_matrices(gpu, lower):
    # trsm.rs's own cases: m = 80, n = 7, diag 2 + i/2, off-diagonal 0.1 (i + j + 1)
    m, n = 80, 7
    a = np.zeros((m, m))
    for i in range(m):
        for j in (range(i + 1) if lower else range(i, m)):
            a[i, j] = 2.0 + i * 0.5 if i == j else 0.1 * (i + j + 1)
    b = np.sin(np.arange(m * n) * 0.13) + 1.0
    x = b.copy()
    _gpu(lower, m, n, 1.0, a.ravel(), m, x, n, False)
    assert np.max(np.abs(a @ x.reshape(m, n) - b.reshape(m, n))) < 1e-9


def test_alpha_unit_diag_and_strides(gpu):
    m, n, lda, ldb = 200, 13, 230, 20
    for lower in (True, False):
        a, b = _system(lower, m, n, np.float64, 3)
        np.fill_diagonal(a, 1.0)
        ap = np.full((m, lda), 7.0)       # padding that must never be read as data
        ap[:, :m] = a
        bp = np.full((m, ldb), -3.0)
        bp[:, :n] = b
        x_unit = bp.ravel().copy()
        _gpu(lower, m, n, 2.5, ap.ravel(), lda, x_unit, ldb, True)
        x_expl = b.ravel().copy()
        _gpu(lower, m, n, 2.5, a.ravel(), m, x_expl, n, False)
        xu = x_unit.reshape(m, ldb)
        assert np.all(xu[:, n:] == -3.0)
        assert np.linalg.norm(xu[:, :n] - x_expl.reshape(m, n)) <= 1e-12 * np.linalg.norm(x_expl)
        assert np.linalg.norm(a @ xu[:, :n] - 2.5 * b) <= 1e-11 * np.linalg.norm(2.5 * b)


def test_invalid_strides_rejected(gpu):
    a = np.eye(4).ravel()
    b = np.ones(8)
    assert not fa.lib().ferray_trsm_left_lower_f64(4, 2, 1.0, a.ctypes.data, 3, b.ctypes.data, 2,
                                                   False)
    with pytest.raises(ValueError):
        fa.trsm_left_lower_f64(4, 2, 1.0, a, 3, b, 2, False)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("lower,m,n", [(True, 1000, 4160), (False, 700, 4096), (True, 1300, 4224)])
def test_columns_of_wide_solves_equal_narrow_solves(gpu, lower, dt, m, n):
    """The columns of B are independent systems: every column of a wide
    solve equals the same column solved in a narrower call, bit for bit (the
    kernels fix each element's operation order whatever the column count),
    and the whole meets the oracle's bound."""
    a, b = _system(lower, m, n, dt, 7)
    x = b.ravel().copy()
    _gpu(lower, m, n, 1.0, a.ravel(), m, x, n, False)
    X = x.reshape(m, n)
    # 2000-column slices covering both ends and the middle
    for c0 in (0, n // 2 - 1000, n - 2000):
        part = np.ascontiguousarray(b[:, c0:c0 + 2000]).ravel()
        _gpu(lower, m, 2000, 1.0, a.ravel(), m, part, 2000, False)
        assert np.array_equal(part.reshape(m, 2000).view(np.uint8),
                              np.ascontiguousarray(X[:, c0:c0 + 2000]).view(np.uint8)), c0
    x_orc = b.ravel().copy()
    oracle.trsm(lower, m, n, dt(1.0), a.ravel(), m, x_orc, n, False)
    tol = 1e-10 if dt == np.float64 else 2e-4
    assert np.linalg.norm(X.astype(np.float64) - x_orc.reshape(m, n)) <= \
        tol * np.linalg.norm(x_orc)
