"""Grouped persistent GEMM: one launch over many problems of mixed dtype.

Bars:
  * every problem's result is bit-identical to the same problem through its
    single-problem entry point, whatever else is in the batch, in any order
    and with either tile shape (each tile runs the entry point's k-sequence);
  * the single-problem results are themselves checked (integers exactly,
    f64 within 2(k+2)u(|alpha||A||B| + |beta||C0|)) so the comparison is not
    between two wrong answers;
  * a plan's captured graph replays give the eager result every time, with
    other device work and host synchronisation between the replays.
"""
from __future__ import annotations

import numpy as np
import pytest

import ferray_amd as fa
from ferray_amd import grouped_gemm as gg
from tests.helpers import exact_int_matmul, rng, wrap_i32

pytestmark = pytest.mark.gpu

F64, F32, C64, C32, I8 = gg.DTYPE_F64, gg.DTYPE_F32, gg.DTYPE_C64, gg.DTYPE_C32, gg.DTYPE_I8
REAL = {F64: np.float64, F32: np.float32, C64: np.float64, C32: np.float32}


class Case:
    """Host operands, device buffers and the single-entry result of one problem."""

    def __init__(self, dtype, m, n, k, seed, alpha=1.0, beta=0.0, accumulate=False):
        self.dtype, self.m, self.n, self.k = dtype, m, n, k
        self.alpha, self.beta, self.accumulate = complex(alpha), complex(beta), accumulate
        r = rng(seed)
        if dtype == I8:
            self.a = r.integers(-128, 128, m * k).astype(np.int8)
            self.b = r.integers(-128, 128, k * n).astype(np.int8)
            self.c0 = r.integers(-2 ** 31, 2 ** 31, m * n).astype(np.int32)
        else:
            f = 2 if dtype in (C64, C32) else 1
            t = REAL[dtype]
            self.a = r.standard_normal(m * k * f).astype(t)
            self.b = r.standard_normal(k * n * f).astype(t)
            self.c0 = r.standard_normal(m * n * f).astype(t)
        self.da = fa.DeviceBuffer.from_array(self.a)
        self.db = fa.DeviceBuffer.from_array(self.b)
        self.dc = fa.DeviceBuffer.from_array(self.c0)
        self.single = self._single()

    def _single(self):
        m, n, k, c = self.m, self.n, self.k, self.c0.copy()
        al, be = self.alpha, self.beta
        if self.dtype == F64:
            ok = fa.gemm_f64(m, n, k, al.real, self.a, self.b, be.real, c)
        elif self.dtype == F32:
            ok = fa.gemm_f32(m, n, k, np.float32(al.real), self.a, self.b, np.float32(be.real), c)
        elif self.dtype == C64:
            ok = fa.gemm_c64(m, n, k, al.real, al.imag, self.a, self.b, be.real, be.imag, c)
        elif self.dtype == C32:
            ok = fa.gemm_c32(m, n, k, np.float32(al.real), np.float32(al.imag), self.a, self.b,
                             np.float32(be.real), np.float32(be.imag), c)
        else:
            ok = fa.gemm_i8_signed(m, n, k, self.a, self.b, c, self.accumulate)
        assert ok, fa.last_status()
        return c

    def reset(self):
        self.dc.upload(self.c0)

    def problem(self):
        return gg.problem(self.dtype, self.m, self.n, self.k, self.da.ptr, self.db.ptr,
                          self.dc.ptr, self.alpha, self.beta, self.accumulate)

    def result(self):
        return self.dc.download(self.c0.dtype, self.c0.size)


def _mix(seed=0):
    cases = [
        Case(F64, 300, 200, 100, seed + 1),
        Case(F64, 64, 64, 64, seed + 2, alpha=-0.5, beta=2.0),
        Case(F64, 129, 130, 17, seed + 3, beta=1.0),
        Case(F64, 1, 1, 1, seed + 4),
        Case(F64, 513, 257, 129, seed + 5, alpha=0.75, beta=-1.0),
        Case(F32, 256, 256, 256, seed + 6),
        Case(F32, 100, 37, 53, seed + 7, alpha=2.0, beta=0.5),
        Case(C64, 70, 50, 33, seed + 8, alpha=0.5 - 1j, beta=1 + 2j),
        Case(C64, 128, 64, 128, seed + 9),
        Case(C32, 65, 66, 67, seed + 10, alpha=1j, beta=-1.0),
        Case(I8, 300, 200, 128, seed + 11),
        Case(I8, 77, 99, 1000, seed + 12, accumulate=True),
        Case(I8, 128, 128, 64, seed + 13),
        Case(I8, 5, 3, 7, seed + 14, accumulate=True),
        Case(F64, 10, 10, 0, seed + 15, beta=2.0),     # k == 0: C = beta C
        Case(I8, 4, 4, 0, seed + 16),                  # k == 0: C = 0
        Case(F32, 33, 17, 0, seed + 17, beta=1.0),     # k == 0, beta 1: untouched
    ]
    return cases


def _run(cases, order=None):
    order = list(range(len(cases))) if order is None else order
    for c in cases:
        c.reset()
    gg.grouped([cases[i].problem() for i in order])
    fa.sync()
    return [c.result() for c in cases]


def _same(x, y):
    return x.tobytes() == y.tobytes()


@pytest.fixture(scope="module")
def mix(gpu):
    return _mix()


def test_single_results_are_right(mix):
    for c in mix:
        if c.dtype == I8 and c.k:
            ref = exact_int_matmul(c.a.reshape(c.m, c.k), c.b.reshape(c.k, c.n)).ravel()
            if c.accumulate:
                ref = ref + c.c0
            assert np.array_equal(c.single, wrap_i32(ref))
        elif c.dtype == F64 and c.k:
            a, b, c0 = c.a.reshape(c.m, c.k), c.b.reshape(c.k, c.n), c.c0.reshape(c.m, c.n)
            ref = c.alpha.real * (a @ b) + c.beta.real * c0
            mag = abs(c.alpha) * (np.abs(a) @ np.abs(b)) + abs(c.beta) * np.abs(c0)
            assert np.all(np.abs(c.single.reshape(c.m, c.n) - ref) <= 2 * (c.k + 2) * 2.0 ** -53 * mag)


def test_mix_bit_identical_to_single(mix):
    for c, got in zip(mix, _run(mix)):
        assert _same(got, c.single), (c.dtype, c.m, c.n, c.k)


@pytest.mark.parametrize("how", ["reversed", "shuffled"])
def test_order_does_not_change_results(mix, how):
    order = list(range(len(mix)))[::-1] if how == "reversed" else \
        list(rng(5).permutation(len(mix)))
    for c, got in zip(mix, _run(mix, order)):
        assert _same(got, c.single)


def test_alone_equals_in_mix(mix):
    for c in mix:
        assert _same(_run([c])[0], c.single)


@pytest.mark.parametrize("cfg", [0, 1])
def test_tile_shape_does_not_change_results(mix, cfg):
    saved = gg.tile_table()
    try:
        gg.set_tile_table([[cfg] * gg.N_BUCKETS for _ in range(gg.N_DTYPES)])
        assert set(v for v in gg._flatten(gg.tile_table())) == {cfg}
        for c, got in zip(mix, _run(mix)):
            assert _same(got, c.single)
    finally:
        gg.set_tile_table(saved)


def test_many_small_problems(gpu):
    r = rng(42)
    cases = []
    for i in range(120):
        dt = int(r.integers(0, 5))
        m, n, k = (int(v) for v in r.integers(1, 160, 3))
        cases.append(Case(dt, m, n, k, 1000 + i, alpha=1.0, beta=0.5 if dt != I8 else 0.0,
                          accumulate=bool(i % 2)))
    for c, got in zip(cases, _run(cases)):
        assert _same(got, c.single), (c.dtype, c.m, c.n, c.k)


def test_plan_replays_match_eager(mix):
    plan = gg.Plan(64, 8192)
    try:
        probs = [c.problem() for c in mix]
        other_a = rng(9).standard_normal(96 * 96)
        for rep in range(4):
            for c in mix:
                c.reset()
            # Eager device work and a host round trip between replays.
            scratch = np.zeros(96 * 96)
            assert fa.gemm_f64(96, 96, 96, 1.0, other_a, other_a, 0.0, scratch)
            plan.launch(probs)
            plan.sync()
            for c in mix:
                assert _same(c.result(), c.single), (rep, c.dtype, c.m, c.n, c.k)
    finally:
        plan.close()


def test_two_plans_in_flight(mix):
    first, second = mix[:8], mix[8:]
    p1, p2 = gg.Plan(16, 4096), gg.Plan(16, 4096)
    try:
        for c in mix:
            c.reset()
        p1.launch([c.problem() for c in first])
        p2.launch([c.problem() for c in second])
        p1.sync()
        p2.sync()
        for c in mix:
            assert _same(c.result(), c.single)
    finally:
        p1.close()
        p2.close()


def test_errors(gpu):
    big = gg.problem(I8, 4, 4, 65537, 1, 1, 1)
    with pytest.raises(fa.FerrayError) as e:
        gg.grouped([big])
    assert e.value.status == 5
    with pytest.raises(fa.FerrayError) as e:
        gg.grouped([gg.problem(7, 4, 4, 4, 1, 1, 1)])
    assert e.value.status == 3
    with pytest.raises(fa.FerrayError) as e:
        gg.grouped([gg.problem(F64, 4, 4, 4, 0, 0, 1)])
    assert e.value.status == 3
    plan = gg.Plan(1, 16)
    try:
        p = gg.problem(F64, 8, 8, 8, 1, 1, 1)
        with pytest.raises(fa.FerrayError):
            plan.launch([p, p])
    finally:
        plan.close()
