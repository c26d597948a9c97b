"""Host-pointer entry points with operands of 8 MiB and more: they move
through the per-thread ring of pinned 16 MiB slots (filled and drained by
host helper threads) instead of direct pageable copies.  The kernels are the
same, so every result must equal the device-pointer entry's bit for bit:
operands spanning more chunks than the ring has slots, a ragged last chunk,
C read back in (beta != 0), several dtypes, and several threads at once."""
from __future__ import annotations

import threading

import numpy as np
import pytest

import ferray_amd as fa
from ferray_amd._lib import lib
from tests.helpers import rng

pytestmark = pytest.mark.gpu


def _dev_f64(m, n, k, alpha, a, b, beta, c0):
    da, db, dc = fa.DeviceBuffer.from_array(a), fa.DeviceBuffer.from_array(b), \
        fa.DeviceBuffer.from_array(c0)
    assert lib().ferray_gemm_f64_dev(m, n, k, alpha, da.ptr, db.ptr, beta, dc.ptr, None) == 0
    return dc.download(np.float64, m * n)


@pytest.mark.parametrize("beta", [0.0, 0.5])
def test_f64_many_chunks_bit_identical(beta):
    # A 80 MiB (5 chunks, more than the 4 slots), B 32 MiB, C 20 MiB (ragged)
    m, n, k = 2560, 1024, 4096
    r = rng(101)
    a, b = r.standard_normal(m * k), r.standard_normal(k * n)
    c0 = r.standard_normal(m * n)
    c = c0.copy()
    assert fa.gemm_f64(m, n, k, 1.25, a, b, beta, c)
    assert np.array_equal(c, _dev_f64(m, n, k, 1.25, a, b, beta, c0))


def test_f64_ragged_sizes_bit_identical():
    m, n, k = 1531, 1209, 1777   # every operand a little over 8 MiB or under
    r = rng(102)
    a, b = r.standard_normal(m * k), r.standard_normal(k * n)
    c0 = r.standard_normal(m * n)
    c = c0.copy()
    assert fa.gemm_f64(m, n, k, -0.75, a, b, 2.0, c)
    assert np.array_equal(c, _dev_f64(m, n, k, -0.75, a, b, 2.0, c0))


def test_i8_large_output_bit_identical():
    m = n = k = 4096                 # A, B 16 MiB each, C 64 MiB of int32
    r = rng(103)
    a = r.integers(-128, 128, m * k).astype(np.int8)
    b = r.integers(-128, 128, k * n).astype(np.int8)
    c = np.zeros(m * n, np.int32)
    assert fa.gemm_i8_signed(m, n, k, a, b, c, False)
    da, db = fa.DeviceBuffer.from_array(a), fa.DeviceBuffer.from_array(b)
    dc = fa.DeviceBuffer(m * n * 4)
    assert lib().ferray_gemm_i8_signed_dev(m, n, k, da.ptr, db.ptr, dc.ptr, False, None) == 0
    assert np.array_equal(c, dc.download(np.int32, m * n))


def test_threads_share_the_copy_pool():
    m, n, k = 2048, 1536, 2048
    cases = []
    for seed in range(4):
        r = rng(200 + seed)
        cases.append((r.standard_normal(m * k), r.standard_normal(k * n)))
    alone = []
    for a, b in cases:
        c = np.zeros(m * n)
        assert fa.gemm_f64(m, n, k, 1.0, a, b, 0.0, c)
        alone.append(c)
    got = [None] * len(cases)
    errs = []

    def work(i):
        try:
            for _ in range(2):
                c = np.zeros(m * n)
                assert fa.gemm_f64(m, n, k, 1.0, cases[i][0], cases[i][1], 0.0, c)
                got[i] = c
        except Exception as e:  # pragma: no cover - reported below
            errs.append(e)

    ts = [threading.Thread(target=work, args=(i,)) for i in range(len(cases))]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errs
    for i in range(len(cases)):
        assert np.array_equal(got[i], alone[i])


# Products large enough for the row-panel pipeline (C of 8 Mi elements and
# more): B up first, then A (and C) in panels whose products start as soon as
# their rows land, C panels back while later ones compute.  The panels run
# the same kernels, so results still equal one device call bit for bit.

@pytest.mark.parametrize("beta", [0.0, 0.5])
def test_f64_row_panels_bit_identical(beta):
    m, n, k = 4000, 2048, 512       # two panels (2048 + 1952 rows)
    r = rng(110)
    a, b = r.standard_normal(m * k), r.standard_normal(k * n)
    c0 = r.standard_normal(m * n)
    c = c0.copy()
    assert fa.gemm_f64(m, n, k, 0.75, a, b, beta, c)
    assert np.array_equal(c, _dev_f64(m, n, k, 0.75, a, b, beta, c0))


def test_f32_and_c64_row_panels_bit_identical():
    r = rng(111)
    m, n, k = 4096, 4096, 256       # four panels
    a = r.standard_normal(m * k).astype(np.float32)
    b = r.standard_normal(k * n).astype(np.float32)
    c0 = r.standard_normal(m * n).astype(np.float32)
    c = c0.copy()
    assert fa.gemm_f32(m, n, k, np.float32(1.5), a, b, np.float32(-0.5), c)
    da, db, dc = (fa.DeviceBuffer.from_array(x) for x in (a, b, c0))
    assert lib().ferray_gemm_f32_dev(m, n, k, 1.5, da.ptr, db.ptr, -0.5, dc.ptr, None) == 0
    assert np.array_equal(c, dc.download(np.float32, m * n))

    m, n, k = 2048, 4096, 128       # two panels of complex rows
    a = (r.standard_normal(2 * m * k)).astype(np.float64)
    b = (r.standard_normal(2 * k * n)).astype(np.float64)
    c0 = r.standard_normal(2 * m * n)
    c = c0.copy()
    assert fa.gemm_c64(m, n, k, 1.0, 0.5, a, b, 0.25, -1.0, c)
    da, db, dc = (fa.DeviceBuffer.from_array(x) for x in (a, b, c0))
    assert lib().ferray_gemm_c64_dev(m, n, k, 1.0, 0.5, da.ptr, db.ptr, 0.25, -1.0, dc.ptr,
                                     None) == 0
    assert np.array_equal(c, dc.download(np.float64, 2 * m * n))


def test_integer_row_panels_bit_identical():
    m, n, k = 4096, 4096, 512       # four panels
    r = rng(112)
    a = r.integers(0, 256, m * k).astype(np.uint8)
    b = r.integers(-128, 128, k * n).astype(np.int8)
    c0 = r.integers(-2 ** 31, 2 ** 31, m * n).astype(np.int32)
    c = c0.copy()
    assert fa.gemm_i8(m, n, k, a, b, c, True)
    da, db, dc = (fa.DeviceBuffer.from_array(x) for x in (a, b, c0))
    assert lib().ferray_gemm_i8_dev(m, n, k, da.ptr, db.ptr, dc.ptr, True, None) == 0
    assert np.array_equal(c, dc.download(np.int32, m * n))

    q0 = r.standard_normal(m * n).astype(np.float32)
    q = q0.copy()
    assert fa.quantized_matmul(m, n, k, a, 3, np.float32(0.02), b, -5, np.float32(0.5), q, True)
    dq = fa.DeviceBuffer.from_array(q0)
    assert lib().ferray_quantized_matmul_dev(m, n, k, da.ptr, 3, 0.02, db.ptr, -5, 0.5, dq.ptr,
                                             True, None) == 0
    assert np.array_equal(q, dq.download(np.float32, m * n))


def test_threads_row_panels():
    m, n, k = 4096, 2048, 256
    cases = []
    for seed in range(3):
        r = rng(300 + seed)
        cases.append((r.standard_normal(m * k), r.standard_normal(k * n)))
    want = [_dev_f64(m, n, k, 1.0, a, b, 0.0, np.zeros(m * n)) for a, b in cases]
    got = [None] * len(cases)
    errs = []

    def work(i):
        try:
            for _ in range(2):
                c = np.zeros(m * n)
                assert fa.gemm_f64(m, n, k, 1.0, cases[i][0], cases[i][1], 0.0, c)
                got[i] = c
        except Exception as e:  # pragma: no cover - reported below
            errs.append(e)

    ts = [threading.Thread(target=work, args=(i,)) for i in range(len(cases))]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errs
    for i in range(len(cases)):
        assert np.array_equal(got[i], want[i])
