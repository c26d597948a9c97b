"""Serving loop: admission, continuous batching and result identity.

Bars:
  * every request's C is bit-identical whether it ran alone, in a mixed
    batch, through the one-launch-per-request mode, with host or
    device-resident operands, and through its single-problem entry point;
  * integers are exact against numpy; f64 is within
    2(k+2)u|alpha||A||B| of numpy's product (u = 2^-53);
  * admission never holds more HBM than the arena, and a request larger
    than the arena is refused at submission.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import ferray_amd as fa
from ferray_amd import serve as sv
from ferray_amd._lib import DTYPE_C32, DTYPE_C64, DTYPE_F32, DTYPE_F64, DTYPE_I8, lib
from tests.helpers import exact_int_matmul, rng, wrap_i32

pytestmark = pytest.mark.gpu
ELEM = {DTYPE_F64: 8, DTYPE_F32: 4, DTYPE_C64: 16, DTYPE_C32: 8, DTYPE_I8: 1}
HOST_NP = {DTYPE_F64: np.float64, DTYPE_F32: np.float32, DTYPE_C64: np.complex128,
           DTYPE_C32: np.complex64, DTYPE_I8: np.int8}


@pytest.fixture(scope="module")
def pools(gpu):
    p = sv.OperandPools(fa, elems=1 << 22, seed=11)
    yield p
    p.free()


def _requests(pools, seed=0, count=40, hi=300):
    r = rng(seed)
    out = []
    shapes = [(1, 1, 1), (129, 65, 33), (64, 64, 64), (7, 300, 5)]
    for i in range(count):
        dt = [DTYPE_F64, DTYPE_F32, DTYPE_C64, DTYPE_C32, DTYPE_I8][i % 5]
        m, n, k = shapes[i] if i < len(shapes) else (int(v) for v in r.integers(1, hi, 3))
        alpha = [1.0, -0.5, 0.25 + 1j, 2.0][i % 4] if dt != DTYPE_I8 else 1.0
        if dt in (DTYPE_F64, DTYPE_F32):
            alpha = complex(alpha).real
        e = ELEM[dt]
        a = pools.slice(dt, m * k * e, r)
        b = pools.slice(dt, k * n * e, r)
        out.append(sv.request(dt, m, n, k, a, b, alpha))
    out.append(sv.request(DTYPE_F64, 5, 6, 0, None, None, 1.0))  # k == 0: C = 0
    out.append(sv.request(DTYPE_I8, 3, 4, 0, None, None))
    return out


def _single(q) -> np.ndarray:
    """The request through its single-problem entry point."""
    L = lib()
    out = fa.DeviceBuffer(q.m * q.n * sv.c_elem_bytes(q.dtype))
    if q.dtype == DTYPE_F64:
        st = L.ferray_gemm_f64_dev(q.m, q.n, q.k, q.alpha_re, q.a, q.b, 0.0, out.ptr, None)
    elif q.dtype == DTYPE_F32:
        st = L.ferray_gemm_f32_dev(q.m, q.n, q.k, q.alpha_re, q.a, q.b, 0.0, out.ptr, None)
    elif q.dtype == DTYPE_C64:
        st = L.ferray_gemm_c64_dev(q.m, q.n, q.k, q.alpha_re, q.alpha_im, q.a, q.b, 0.0, 0.0,
                                   out.ptr, None)
    elif q.dtype == DTYPE_C32:
        st = L.ferray_gemm_c32_dev(q.m, q.n, q.k, q.alpha_re, q.alpha_im, q.a, q.b, 0.0, 0.0,
                                   out.ptr, None)
    else:
        st = L.ferray_gemm_i8_signed_dev(q.m, q.n, q.k, q.a, q.b, out.ptr, False, None)
    assert st == 0
    fa.sync()
    res = out.download(sv.c_dtype(q.dtype), q.m * q.n)
    out.free()
    return res


def _host_operands(q):
    t = HOST_NP[q.dtype]
    nb_a = q.m * q.k * ELEM[q.dtype]
    nb_b = q.k * q.n * ELEM[q.dtype]
    a = np.empty(q.m * q.k, dtype=t)
    b = np.empty(q.k * q.n, dtype=t)
    if nb_a:
        assert lib().ferray_dev_download(a.ctypes.data, q.a, nb_a, None) == 0
    if nb_b:
        assert lib().ferray_dev_download(b.ctypes.data, q.b, nb_b, None) == 0
    return a, b


def _serve(reqs, mode=sv.SERVE_GROUPED, lanes=3, max_batch=16, host=False):
    cfg = sv.config(hbm_bytes=256 << 20, max_resident=1024, max_batch=max_batch, lanes=lanes,
                    mode=mode, keep_results=True)
    with sv.Server(cfg) as srv:
        tickets = []
        for q in reqs:
            if host:
                a, b = _host_operands(q)
                hq = sv.request(q.dtype, q.m, q.n, q.k, a.ctypes.data if a.size else None,
                                b.ctypes.data if b.size else None,
                                complex(q.alpha_re, q.alpha_im), on_host=True)
                tickets.append(srv.submit(hq, keep_alive=(a, b)))
            else:
                tickets.append(srv.submit(q))
        out = [srv.result(t, q.dtype, q.m, q.n) for t, q in zip(tickets, reqs)]
        st = srv.stats()
    return out, st


@pytest.fixture(scope="module")
def mixed(pools):
    reqs = _requests(pools)
    res, st = _serve(reqs)
    return reqs, res, st


def _same(x, y):
    return x.tobytes() == y.tobytes()


def test_results_are_right(mixed):
    reqs, res, _ = mixed
    for q, c in zip(reqs, res):
        if q.k == 0:
            assert not np.any(c)
            continue
        a, b = _host_operands(q)
        if q.dtype == DTYPE_I8:
            ref = wrap_i32(exact_int_matmul(a.reshape(q.m, q.k), b.reshape(q.k, q.n)).ravel())
            assert np.array_equal(c, ref)
        elif q.dtype == DTYPE_F64:
            A, B = a.reshape(q.m, q.k), b.reshape(q.k, q.n)
            ref = q.alpha_re * (A @ B)
            bound = 2 * (q.k + 2) * 2.0 ** -53 * abs(q.alpha_re) * (np.abs(A) @ np.abs(B))
            assert np.all(np.abs(c.reshape(q.m, q.n) - ref) <= bound)


def test_mix_equals_single_entry(mixed):
    reqs, res, _ = mixed
    for q, c in zip(reqs, res):
        assert _same(c, _single(q)), (q.dtype, q.m, q.n, q.k)


def test_mix_equals_alone(mixed, pools):
    reqs, res, _ = mixed
    for q, c in list(zip(reqs, res))[::3]:
        alone, _ = _serve([q], lanes=1)
        assert _same(alone[0], c)


def test_per_request_mode_identical(mixed):
    reqs, res, _ = mixed
    out, st = _serve(reqs, mode=sv.SERVE_PER_REQUEST)
    for a, b in zip(out, res):
        assert _same(a, b)
    assert st.launches == len(reqs)


def test_host_operands_identical(mixed):
    reqs, res, _ = mixed
    out, _ = _serve(reqs[:15], host=True)
    for a, b in zip(out, res[:15]):
        assert _same(a, b)


def test_every_request_launched_once(mixed):
    reqs, _, st = mixed
    assert st.completed == len(reqs) and st.launched_requests == len(reqs)
    assert 1 <= st.launches <= len(reqs)


def test_admission_bounded_by_arena(pools):
    r = rng(3)
    cap = 1 << 20
    cfg = sv.config(hbm_bytes=cap, max_resident=64, max_batch=8, lanes=2)
    with sv.Server(cfg) as srv:
        tickets = []
        for _ in range(60):
            m, n, k = (int(v) for v in r.integers(16, 200, 3))
            a = pools.slice(DTYPE_F64, m * k * 8, r)
            b = pools.slice(DTYPE_F64, k * n * 8, r)
            tickets.append(srv.submit(sv.request(DTYPE_F64, m, n, k, a, b)))
        lat = [srv.wait(t) for t in tickets]
        st = srv.stats()
        assert st.completed == 60 and all(x >= 0 for x in lat)
        assert st.arena_bytes == cap and 0 < st.arena_peak <= cap
        big = sv.request(DTYPE_F64, 1024, 1024, 4, pools.f64.ptr, pools.f64.ptr)
        with pytest.raises(fa.FerrayError):
            srv.submit(big)  # 8 MiB of C cannot fit a 1 MiB arena


def test_max_resident_one_with_kept_results(mixed):
    reqs, res, _ = mixed
    cfg = sv.config(hbm_bytes=64 << 20, max_resident=1, max_batch=4, lanes=2, keep_results=True)
    with sv.Server(cfg) as srv:
        tickets = [srv.submit(q) for q in reqs[:6]]
        assert srv.poll(tickets[-1]) in (0, 1, 2, 3)
        out = [srv.result(t, q.dtype, q.m, q.n) for t, q in zip(tickets, reqs[:6])]
    for a, b in zip(out, res[:6]):
        assert _same(a, b)


def test_close_drains_unfetched(pools):
    reqs = _requests(pools, seed=5, count=20)
    cfg = sv.config(hbm_bytes=8 << 20, max_resident=8, max_batch=4, lanes=2, keep_results=True)
    srv = sv.Server(cfg)
    for q in reqs:
        srv.submit(q)
    srv.close()  # kept results are dropped; queued requests still run


@pytest.mark.parametrize("mode", [sv.SERVE_GROUPED, sv.SERVE_PER_REQUEST])
def test_trace_replay(pools, mode):
    reqs, arrival, flops = sv.make_trace(pools, 400, rate_per_s=2e5, seed=1, lo=16, hi=256)
    cfg = sv.config(hbm_bytes=256 << 20, max_resident=512, max_batch=64, lanes=3, mode=mode)
    lat, wall, st = sv.run_trace(cfg, reqs, arrival)
    assert st.completed == 400 and np.all(lat > 0) and wall > 0
    if mode == sv.SERVE_PER_REQUEST:
        assert st.launches == 400
    else:
        assert st.launches < 400
