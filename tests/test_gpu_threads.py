"""Concurrent drop-in calls: the C ABI is reentrant (each calling thread gets
its own stream and staging buffers), as ferray's callers may call from rayon
workers (products/mod.rs:625-637).  Every result must equal the same call made
alone, bit for bit."""
from __future__ import annotations

import threading

import numpy as np

import pytest

import ferray_amd as fa
from tests.helpers import rng

pytestmark = pytest.mark.gpu


def _case(seed):
    r = rng(seed)
    m, n, k = (int(v) for v in r.integers(1, 300, 3))
    kind = seed % 3
    if kind == 0:
        a, b = r.standard_normal(m * k), r.standard_normal(k * n)
        c = np.zeros(m * n)
    elif kind == 1:
        a = r.standard_normal(m * k).astype(np.float32)
        b = r.standard_normal(k * n).astype(np.float32)
        c = np.zeros(m * n, np.float32)
    else:
        a = r.integers(-128, 128, m * k).astype(np.int8)
        b = r.integers(-128, 128, k * n).astype(np.int8)
        c = np.zeros(m * n, np.int32)
    return kind, m, n, k, a, b, c


def _run(case):
    kind, m, n, k, a, b, c = case
    c = c.copy()
    if kind == 0:
        ok = fa.gemm_f64(m, n, k, 1.0, a, b, 0.0, c)
    elif kind == 1:
        ok = fa.gemm_f32(m, n, k, np.float32(1), a, b, np.float32(0), c)
    else:
        ok = fa.gemm_i8_signed(m, n, k, a, b, c, False)
    assert ok, fa.last_status()
    return c


def test_concurrent_calls_match_serial(gpu):
    cases = [_case(s) for s in range(48)]
    serial = [_run(cs) for cs in cases]
    results = [None] * len(cases)
    errors = []

    def worker(t):
        try:
            for i in range(t, len(cases), 6):
                results[i] = _run(cases[i])
        except Exception as e:  # pragma: no cover - reported below
            errors.append(repr(e))

    threads = [threading.Thread(target=worker, args=(t,)) for t in range(6)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors
    for i, (x, y) in enumerate(zip(results, serial)):
        assert x.tobytes() == y.tobytes(), i
