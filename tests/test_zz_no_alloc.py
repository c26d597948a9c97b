"""The allocation-free i8 entry point (ferray_gemm_i8_signed_dev_noalloc).

Bars:
  * through the C ABI, two calls in a row (the second on a smaller, ragged
    problem that packs A into the same fixed scratch) equal
    ferray_gemm_i8_signed_dev bit for bit, written and accumulated;
  * the headline shape (8192^3, aligned) fits the fixed scratch and matches;
  * a problem whose packed planes need more than the fixed scratch is refused
    with FERRAY_ERR_SCRATCH before anything runs (C is left as it was), and a
    NULL stream is refused.
"""
from __future__ import annotations

import numpy as np
import pytest

import ferray_amd as fa
from ferray_amd._lib import lib

pytestmark = pytest.mark.gpu

ERR_INVALID, ERR_SCRATCH = 3, 6


def _operands(m, n, k, seed):
    da, db = fa.DeviceBuffer(max(m * k, 1)), fa.DeviceBuffer(max(k * n, 1))
    fa.fill(da, m * k, fa.FILL_I8, seed, -128, 127)
    fa.fill(db, k * n, fa.FILL_I8, seed + 1, -128, 127)
    return da, db


def _run(entry, m, n, k, da, db, c0, accumulate, stream):
    dc = fa.DeviceBuffer(m * n * 4)
    dc.upload(c0)
    st = entry(m, n, k, da.ptr, db.ptr, dc.ptr, accumulate, stream)
    assert st == 0, lib().ferray_status_string(st).decode()
    fa.sync(stream)
    out = dc.download(np.int32, m * n)
    dc.free()
    return out


@pytest.mark.parametrize("accumulate", [False, True])
def test_noalloc_matches_the_allocating_entry_twice_in_a_row(gpu, accumulate):
    L = lib()
    stream = L.ferray_default_stream()
    # large aligned (A read in place), then smaller and ragged (A packed)
    for i, (m, n, k) in enumerate([(4096, 3072, 2048), (1000, 700, 333)]):
        da, db = _operands(m, n, k, 40 + i)
        c0 = np.random.default_rng(i).integers(-2 ** 31, 2 ** 31, m * n).astype(np.int32)
        want = _run(L.ferray_gemm_i8_signed_dev, m, n, k, da, db, c0, accumulate, stream)
        got = _run(L.ferray_gemm_i8_signed_dev_noalloc, m, n, k, da, db, c0, accumulate, stream)
        da.free(); db.free()
        bad = np.flatnonzero(got != want)
        assert bad.size == 0, (m, n, k, bad.size, divmod(int(bad[0]), n))


def test_noalloc_headline_shape(gpu):
    L = lib()
    stream = L.ferray_default_stream()
    n = 8192
    da, db = _operands(n, n, n, 50)
    c0 = np.zeros(n * n, np.int32)
    want = _run(L.ferray_gemm_i8_signed_dev, n, n, n, da, db, c0, False, stream)
    got = _run(L.ferray_gemm_i8_signed_dev_noalloc, n, n, n, da, db, c0, False, stream)
    da.free(); db.free()
    assert np.array_equal(got, want)


def test_noalloc_refuses_what_does_not_fit(gpu):
    L = lib()
    stream = L.ferray_default_stream()
    # B^T planes alone: 16384 columns x 8256 bytes of k = 129 MiB > 128 MiB
    m, n, k = 64, 16384, 8256
    da, db = _operands(m, n, k, 60)
    dc = fa.DeviceBuffer(m * n * 4)
    c0 = np.arange(m * n, dtype=np.int32)
    dc.upload(c0)
    st = L.ferray_gemm_i8_signed_dev_noalloc(m, n, k, da.ptr, db.ptr, dc.ptr, False, stream)
    assert st == ERR_SCRATCH
    fa.sync(stream)
    assert np.array_equal(dc.download(np.int32, m * n), c0)
    assert L.ferray_gemm_i8_signed_dev_noalloc(m, n, k, da.ptr, db.ptr, dc.ptr, False,
                                               None) == ERR_INVALID
    for d in (da, db, dc):
        d.free()
