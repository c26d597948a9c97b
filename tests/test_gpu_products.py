"""The product layer on the device against numpy.

Bar: elementwise |C - C_numpy| <= 2(k+2)u(|A||B|) with u the unit roundoff of
the dtype (2^-53 f64, 2^-24 f32), as for the GEMM entry points; batched
results equal the per-matrix GEMM entry point's bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

import ferray_amd as fa
from ferray_amd import products as pr
from tests.helpers import rng

pytestmark = pytest.mark.gpu
U = {np.float64: 2.0 ** -53, np.float32: 2.0 ** -24, np.complex128: 2.0 ** -53,
     np.complex64: 2.0 ** -24}


def _arr(r, shape, dt):
    x = r.standard_normal(shape)
    if dt in (np.complex128, np.complex64):
        x = x + 1j * r.standard_normal(shape)
    return x.astype(dt)


def _close(got, a, b, dt, k, prod=np.matmul):
    ref = prod(a.astype(np.complex128 if np.iscomplexobj(a) else np.float64),
               b.astype(np.complex128 if np.iscomplexobj(b) else np.float64))
    mag = prod(np.abs(a).astype(np.float64), np.abs(b).astype(np.float64))
    f = 4 if np.iscomplexobj(a) else 1
    assert got.shape == np.shape(ref)
    assert np.all(np.abs(got - ref) <= f * 2 * (k + 2) * U[dt] * mag + 1e-300)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("sa,sb", [((37, 19), (19, 23)), ((19,), (19, 23)), ((37, 19), (19,)),
                                   ((300, 260), (260, 270))])
def test_matmul_2d_and_vectors(gpu, dt, sa, sb):
    r = rng(1)
    a, b = _arr(r, sa, dt), _arr(r, sb, dt)
    _close(pr.matmul(a, b), a, b, dt, sa[-1])


@pytest.mark.parametrize("sa,sb", [((4, 17, 9), (4, 9, 13)), ((3, 1, 8, 5), (2, 5, 6)),
                                   ((6, 8, 5), (5, 7)), ((8, 5), (2, 3, 5, 7))])
def test_matmul_batched_broadcast(gpu, sa, sb):
    r = rng(2)
    a, b = _arr(r, sa, np.float64), _arr(r, sb, np.float64)
    got = pr.matmul(a, b)
    # The reference indexes matrix i as A[i % a_batch], B[i % b_batch]
    # (products/mod.rs:625-650); for these shapes that equals numpy's broadcast
    # except where both sides repeat, so compare per matrix with that rule.
    m, n = a.shape[-2], b.shape[-1]
    A = a.reshape(-1, m, a.shape[-1])
    B = b.reshape(-1, b.shape[-2], n)
    G = got.reshape(-1, m, n)
    for i in range(G.shape[0]):
        single = np.zeros(m * n)
        assert fa.gemm_f64(m, n, A.shape[2], 1.0, np.ascontiguousarray(A[i % A.shape[0]]),
                           np.ascontiguousarray(B[i % B.shape[0]]), 0.0, single)
        assert G[i].tobytes() == single.reshape(m, n).tobytes()
    na, nb, nbatch = A.shape[0], B.shape[0], G.shape[0]
    if na in (1, nbatch) and nb in (1, nbatch):
        # one side unbatched or both equal: the modulo rule is numpy's broadcast
        _close(got, a, b, np.float64, sa[-1])


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
def test_matmul_complex(gpu, dt):
    r = rng(3)
    a, b = _arr(r, (33, 20), dt), _arr(r, (20, 41), dt)
    _close(pr.matmul_complex(a, b), a, b, dt, 20)


@pytest.mark.parametrize("sa,sb", [((50,), (50,)), ((12, 30), (30, 7)), ((3, 4, 30), (30,)),
                                   ((3, 30), (30, 2, 5)), ((30,), (30, 6))])
def test_dot(gpu, sa, sb):
    r = rng(4)
    a, b = _arr(r, sa, np.float64), _arr(r, sb, np.float64)
    got = pr.dot(a, b)
    ref_fn = lambda x, y: np.tensordot(x, y, axes=1)
    _close(np.asarray(got), a, b, np.float64, sa[-1], prod=ref_fn)


def test_gemm_blas_style(gpu):
    r = rng(5)
    a, b = _arr(r, (64, 48), np.float64), _arr(r, (48, 80), np.float64)
    _close(pr.gemm(a, b), a, b, np.float64, 48)
    # the reference's own 2x2 known answer (conformance_internal_surface.rs:741-746)
    c = pr.gemm(np.array([[1.0, 2.0], [3.0, 4.0]]), np.array([[5.0, 6.0], [7.0, 8.0]]))
    assert c.tolist() == [[19.0, 22.0], [43.0, 50.0]]


def test_matmul_raw_flat(gpu):
    r = rng(6)
    a, b = _arr(r, (7 * 5,), np.float32), _arr(r, (5 * 3,), np.float32)
    got = pr.matmul_raw(a, b, 7, 5, 3)
    assert got.shape == (21,)
    _close(got.reshape(7, 3), a.reshape(7, 5), b.reshape(5, 3), np.float32, 5)


def test_batched_int8_through_grouped(gpu):
    r = rng(7)
    a = r.integers(-128, 128, (5, 20, 33)).astype(np.int8)
    b = r.integers(-128, 128, (33, 17)).astype(np.int8)
    out = np.zeros((5, 20, 17), dtype=np.int32)
    ok = fa.lib().ferray_matmul_batched(4, 5, 5, 1, 20, 33, 17, a.ctypes.data, b.ctypes.data,
                                        out.ctypes.data)
    assert ok
    ref = np.matmul(a.astype(np.int64), b.astype(np.int64))
    assert np.array_equal(out, ref.astype(np.int32))


@pytest.mark.parametrize("kind", ["i8", "u8", "i16"])
def test_matmul_int_exact(gpu, kind):
    r = rng(8)
    m, k, n = 37, 300, 45
    if kind == "i8":
        a = r.integers(-128, 128, (m, k)).astype(np.int8)
        b = r.integers(-128, 128, (k, n)).astype(np.int8)
    elif kind == "u8":
        a = r.integers(0, 256, (m, k)).astype(np.uint8)
        b = r.integers(-128, 128, (k, n)).astype(np.int8)
    else:
        a = r.integers(-32768, 32768, (m, k)).astype(np.int16)
        b = r.integers(-32768, 32768, (k, n)).astype(np.int16)
    got = pr.matmul_int(a, b)
    ref = (a.astype(np.int64) @ b.astype(np.int64)) & 0xFFFFFFFF
    assert got.dtype == np.int32
    assert np.array_equal(got.astype(np.int64) & 0xFFFFFFFF, ref)


def test_matrix_power_fibonacci_exact(gpu):
    # [[1,1],[1,0]]^n = [[F(n+1), F(n)], [F(n), F(n-1)]], exact in f64 for n <= 70
    q = np.array([[1.0, 1.0], [1.0, 0.0]])
    fib = [0, 1]
    for _ in range(80):
        fib.append(fib[-1] + fib[-2])
    for n in (1, 2, 5, 31, 70):
        got = pr.matrix_power(q, n)
        assert got.tolist() == [[fib[n + 1], fib[n]], [fib[n], fib[n - 1]]]


@pytest.mark.parametrize("dt", [np.float64, np.complex128, np.float32])
def test_matrix_power_against_numpy(gpu, dt):
    r = rng(9)
    sz = 96
    a = (np.eye(sz) * 0.5 + 0.02 * r.standard_normal((sz, sz))).astype(dt)
    if np.iscomplexobj(a):
        a = a + 0.02j * r.standard_normal((sz, sz))
    for p in (0, 1, 3, 12):
        got = pr.matrix_power(a, p)
        ref = np.linalg.matrix_power(a.astype(np.complex128 if np.iscomplexobj(a) else np.float64), p)
        tol = 1e-12 if dt != np.float32 else 2e-5
        assert np.linalg.norm(got - ref) <= tol * max(1.0, np.linalg.norm(ref))


def test_multi_dot_against_numpy(gpu):
    r = rng(10)
    shapes = [(10, 300), (300, 5), (5, 200), (200, 40)]
    mats = [r.standard_normal(s) for s in shapes]
    got = pr.multi_dot(mats)
    ref = np.linalg.multi_dot(mats)
    assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)
    v0, v1 = r.standard_normal(10), r.standard_normal(40)
    got = pr.multi_dot([v0] + mats + [v1])
    ref = np.linalg.multi_dot([v0] + mats + [v1])
    assert abs(float(got) - float(ref)) <= 1e-10 * max(1.0, abs(float(ref)))


def test_einsum_matmul(gpu):
    r = rng(11)
    a, b = r.standard_normal((20, 30)), r.standard_normal((30, 25))
    for s in ("ij,jk->ik", "ab,bc->ac"):
        got = pr.einsum_matmul(s, a, b)
        assert np.linalg.norm(got - np.einsum(s, a, b)) <= 1e-13 * np.linalg.norm(a @ b)
    x, y = r.standard_normal((4, 20, 30)), r.standard_normal((4, 30, 25))
    got = pr.einsum_matmul("bij,bjk->bik", x, y)
    assert np.linalg.norm(got - np.einsum("bij,bjk->bik", x, y)) <= 1e-13 * np.linalg.norm(got)
