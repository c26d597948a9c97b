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
