"""Every benchmark workload checked at its stated size.

The workloads (one MI355X, synthetic seeded data):
  * plain DGEMM / SGEMM, square M = N = K in {1024, 4096, 8192}, entries
    normal(0, 1); ZGEMM / CGEMM at 4096 with independent normal real and
    imaginary parts;
  * i8 x i8 -> i32 at 8192^3, entries uniform in [-128, 127], bit-exact;
    i16 x i16 -> i32 at 4096^3 over the full range, bit-exact mod 2^32;
  * split-precision (Ozaki) DGEMM at 4096^3 and 8192^3 with 12 to 16 moduli,
    on normal(0, 1) and on u * 2^e (u uniform [-1, 1), e uniform in
    [-20, 20]) inputs;
  * a serving trace of 2000 requests, M, N, K log-uniform in [16, 2048],
    50 % f64 / 30 % f32 / 10 % c64 / 10 % i8, every result kept.

Bars:
  * real floats: |C - C_oracle| <= 2(k + 2) u |A||B| elementwise against the
    oracle's restatement of ferray's Goto DGEMM / SGEMM on every element, and
    ||C - C_dd|| / ||C_dd|| <= 8 sqrt(k) u over 32 full rows against the
    compensated (double-double) reference;
  * complex: the same bound, doubled, against numpy's complex product on every
    element, and against the double-double complex reference on 16 full rows;
  * integers: equal to the exact product mod 2^32 on every element (float64
    BLAS products, exact because every partial sum stays below 2^53), and
    to the oracle's closed form on 16 full rows; written and accumulated;
  * Ozaki: normwise ||C - C_dd||_F / (||A||_F ||B||_F), estimated from 64
    full rows, <= 1e-15 at 13 moduli and under the per-count limits below;
    |C - C_dd| <= 2^(6 - g) ||a_i|| ||b_j|| + 4 u |C_dd| elementwise (g the
    scaling exponent of the modulus count); errors fall as moduli are added;
  * serving: every result equals its single-problem entry point's bit for bit,
    batching happened, and 60 sampled results meet the bounds above against
    numpy.

The device generator that makes these inputs is itself checked against the
oracle's generators first.
"""
from __future__ import annotations

import math
import time

import numpy as np
import pytest

import ferray_amd as fa
import oracle
from ferray_amd import serve as sv
from ferray_amd._lib import DTYPE_C32, DTYPE_C64, DTYPE_F32, DTYPE_F64, DTYPE_I8, lib
from tests.helpers import rng, wrap_i32

pytestmark = pytest.mark.gpu

U64, U32 = 2.0 ** -53, 2.0 ** -24


def _dev(nbytes):
    return fa.DeviceBuffer(nbytes)


def _filled(count, kind, seed, dtype, lo=0, hi=0):
    """A device buffer of `count` generated elements and its host copy."""
    buf = _dev(count * np.dtype(dtype).itemsize)
    fa.fill(buf, count, kind, seed, lo, hi)
    fa.sync()
    return buf, buf.download(dtype, count)


def _ok(status, what):
    assert status == 0, f"{what}: {lib().ferray_status_string(status).decode()}"


# ------------------------------------------------------------ the input generator

def test_device_generator_matches_the_oracle_generators(gpu):
    n = 1 << 16
    buf, u = _filled(n, fa.FILL_F64_UNIFORM, 5, np.float64)
    buf.free()
    assert u.tobytes() == oracle.uniform_f64(n, 5).tobytes()
    for kind, dt, lo, hi in ((fa.FILL_I8, np.int8, -128, 127), (fa.FILL_I16, np.int16, -32768, 32767),
                             (fa.FILL_U8, np.uint8, 0, 255)):
        buf, x = _filled(n, kind, 6, dt, lo, hi)
        buf.free()
        assert np.array_equal(x.astype(np.int64), oracle.uniform_int(n, 6, lo, hi)), dt
    # normal(0,1): Box-Muller on draws 2i and 2i+1 of the same stream
    buf, g = _filled(4096, fa.FILL_F64_NORMAL, 8, np.float64)
    buf.free()
    mix = np.array([oracle.lib().orc_splitmix64(8, i) for i in range(8192)], np.uint64)
    u1 = ((mix[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * 2.0 ** -53
    u2 = (mix[1::2] >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    want = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    assert np.all(np.abs(g - want) <= 1e-13 * np.maximum(1.0, np.abs(want)))
    # u * 2^e: exact
    buf, w = _filled(4096, fa.FILL_F64_WIDE, 9, np.float64, -20, 20)
    buf.free()
    mix = np.array([oracle.lib().orc_splitmix64(9, i) for i in range(8192)], np.uint64)
    uu = (mix[0::2] >> np.uint64(11)).astype(np.float64) * 2.0 ** -52 - 1.0
    ee = -20 + (((mix[1::2] >> np.uint64(32)) * np.uint64(41)) >> np.uint64(32)).astype(np.int64)
    assert w.tobytes() == np.ldexp(uu, ee).tobytes()


# ------------------------------------------------------------ real float sweep

@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [1024, 4096, 8192])
def test_real_sweep_at_size(gpu, dtype, n):
    f64 = dtype == np.float64
    kind = fa.FILL_F64_NORMAL if f64 else fa.FILL_F32_NORMAL
    dA, a = _filled(n * n, kind, 1, dtype)
    dB, b = _filled(n * n, kind, 2, dtype)
    dC = _dev(n * n * a.itemsize)
    run = lib().ferray_gemm_f64_dev if f64 else lib().ferray_gemm_f32_dev
    _ok(run(n, n, n, 1.0, dA.ptr, dB.ptr, 0.0, dC.ptr, None), "gemm")
    fa.sync()
    c = dC.download(dtype, n * n)
    for d in (dA, dB, dC):
        d.free()
    ref = np.zeros(n * n, dtype)
    (oracle.gemm_f64_goto if f64 else oracle.gemm_f32_goto)(n, n, n, dtype(1.0), a, b,
                                                             dtype(0.0), ref)
    A = a.reshape(n, n).astype(np.float64)
    B = b.reshape(n, n).astype(np.float64)
    u = U64 if f64 else U32
    bound = 2 * (n + 2) * u * (np.abs(A) @ np.abs(B))
    err = np.abs(c.reshape(n, n).astype(np.float64) - ref.reshape(n, n))
    bad = np.flatnonzero(err > bound)
    assert bad.size == 0, f"{bad.size} elements over the bound, first {divmod(int(bad[0]), n)}"
    rows = rng(n).choice(n, 32, replace=False)
    hi, lo = oracle.dd_gemm_f64_sampled(n, n, n, A.ravel(), B.ravel(), rows, np.arange(n))
    d = (c.reshape(n, n)[rows].astype(np.float64) - hi) - lo
    assert np.linalg.norm(d) <= 8 * math.sqrt(n) * u * np.linalg.norm(hi)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_complex_at_4096(gpu, dtype):
    n = 4096
    c64 = dtype == np.complex128
    real = np.float64 if c64 else np.float32
    kind = fa.FILL_F64_NORMAL if c64 else fa.FILL_F32_NORMAL
    dA, a = _filled(2 * n * n, kind, 3, real)
    dB, b = _filled(2 * n * n, kind, 4, real)
    dC = _dev(2 * n * n * a.itemsize)
    if c64:
        st = lib().ferray_gemm_c64_dev(n, n, n, 1.0, 0.0, dA.ptr, dB.ptr, 0.0, 0.0, dC.ptr, None)
    else:
        st = lib().ferray_gemm_c32_dev(n, n, n, 1.0, 0.0, dA.ptr, dB.ptr, 0.0, 0.0, dC.ptr, None)
    _ok(st, "complex gemm")
    fa.sync()
    c = dC.download(dtype, n * n).reshape(n, n).astype(np.complex128)
    for d in (dA, dB, dC):
        d.free()
    A = a.astype(np.float64).view(np.complex128).reshape(n, n)
    B = b.astype(np.float64).view(np.complex128).reshape(n, n)
    u = U64 if c64 else U32
    mag = (np.abs(A.real) + np.abs(A.imag)) @ (np.abs(B.real) + np.abs(B.imag))
    ref = A @ B
    bound = 2 * 2 * (n + 2) * u * mag + 4 * (n + 2) * U64 * mag   # device + numpy's own error
    err = np.maximum(np.abs(c.real - ref.real), np.abs(c.imag - ref.imag))
    bad = np.flatnonzero(err > bound)
    assert bad.size == 0, f"{bad.size} elements over the bound, first {divmod(int(bad[0]), n)}"
    rows = rng(7).choice(n, 16, replace=False)
    dd = oracle.dd_gemm_c64_sampled(n, n, n, A.view(np.float64).ravel(),
                                    B.view(np.float64).ravel(), rows, np.arange(n))
    assert np.all(np.abs(c[rows] - dd) <= 2 * (n + 2) * u * mag[rows])
    assert np.linalg.norm(c[rows] - dd) <= 8 * math.sqrt(2 * n) * u * np.linalg.norm(dd)


# ------------------------------------------------------------ integers

def _exact_f64(a, b, m, n, k):
    """Exact integer product through float64 BLAS: every partial sum is an
    integer below 2^53 in magnitude, so no step rounds."""
    return np.rint(a.reshape(m, k).astype(np.float64) @
                   b.reshape(k, n).astype(np.float64)).astype(np.int64)


def _same(c, want, n):
    bad = np.flatnonzero(c != want)
    assert bad.size == 0, (f"{bad.size} wrong, first at {divmod(int(bad[0]), n)}: "
                           f"got {c[bad[0]]}, want {want[bad[0]]}")


@pytest.mark.parametrize("kind", ["i8", "i16"])
def test_integer_bit_exact_at_size(gpu, kind):
    if kind == "i8":
        n, fill, dt, lo, hi = 8192, fa.FILL_I8, np.int8, -128, 127
        run = lib().ferray_gemm_i8_signed_dev
    else:
        n, fill, dt, lo, hi = 4096, fa.FILL_I16, np.int16, -32768, 32767
        run = lib().ferray_gemm_i16_dev
    dA, a = _filled(n * n, fill, 11, dt, lo, hi)
    dB, b = _filled(n * n, fill, 12, dt, lo, hi)
    dC = _dev(n * n * 4)
    _ok(run(n, n, n, dA.ptr, dB.ptr, dC.ptr, False, None), "written")
    fa.sync()
    c = dC.download(np.int32, n * n)
    exact = _exact_f64(a, b, n, n, n)
    want = wrap_i32(exact).ravel()
    _same(c, want, n)
    # the oracle's closed form on 16 full rows
    rows = rng(13).choice(n, 16, replace=False)
    sub = np.ascontiguousarray(a.reshape(n, n)[rows]).ravel()
    ro = np.zeros(16 * n, np.int32)
    if kind == "i8":
        oracle.gemm_i8_signed_threaded(16, n, n, sub, b, ro)
    else:
        oracle.gemm_i16(16, n, n, sub, b, ro, False)
    assert np.array_equal(ro.reshape(16, n), c.reshape(n, n)[rows])
    # accumulate onto arbitrary i32 values
    c0 = rng(14).integers(-2 ** 31, 2 ** 31, n * n).astype(np.int32)
    dC.upload(c0)
    _ok(run(n, n, n, dA.ptr, dB.ptr, dC.ptr, True, None), "accumulated")
    fa.sync()
    c = dC.download(np.int32, n * n)
    _same(c, wrap_i32(exact.ravel() + c0), n)
    for d in (dA, dB, dC):
        d.free()


# ------------------------------------------------------------ split precision

OZ_MODULI = [256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197, 193]
# normwise error limits per modulus count (each modulus adds about 4 bits)
OZ_LIMIT = {12: 3e-14, 13: 1e-15, 14: 1e-16, 15: 1.5e-17, 16: 2e-18}


def _oz_gamma(nmod):
    M = math.prod(OZ_MODULI[:nmod])
    return (M.bit_length() - 1 - 2) // 2


@pytest.mark.parametrize("dist", ["normal", "wide"])
@pytest.mark.parametrize("n", [4096, 8192])
def test_ozaki_at_size(gpu, n, dist):
    kind, lo, hi = (fa.FILL_F64_NORMAL, 0, 0) if dist == "normal" else (fa.FILL_F64_WIDE, -20, 20)
    dA, a = _filled(n * n, kind, 21, np.float64, lo, hi)
    dB, b = _filled(n * n, kind, 22, np.float64, lo, hi)
    dC = _dev(n * n * 8)
    rows = np.sort(rng(n + 1).choice(n, 64, replace=False))
    ref_hi, ref_lo = oracle.dd_gemm_f64_sampled(n, n, n, a, b, rows, np.arange(n))
    nab = np.linalg.norm(a) * np.linalg.norm(b)
    ra = np.linalg.norm(a.reshape(n, n)[rows], axis=1)[:, None]
    cb = np.linalg.norm(b.reshape(n, n), axis=0)[None, :]
    del a, b

    def rows_of_c():
        return np.stack([dC.download(np.float64, n, int(r) * n * 8) for r in rows])

    errs = {}
    for nm in (12, 13, 14, 15, 16):
        _ok(lib().ferray_dgemm_ozaki_dev(n, n, n, 1.0, dA.ptr, dB.ptr, 0.0, dC.ptr, nm, None),
            f"ozaki {nm}")
        fa.sync()
        d = (rows_of_c() - ref_hi) - ref_lo
        errs[nm] = math.sqrt(n / len(rows)) * np.linalg.norm(d) / nab
        lim = 2.0 ** (6 - _oz_gamma(nm)) * ra * cb + 4 * U64 * np.abs(ref_hi)
        assert np.all(np.abs(d) <= lim), (nm, float(np.max(np.abs(d) / lim)))
        assert errs[nm] <= OZ_LIMIT[nm], (nm, errs)
    assert errs[13] <= 1e-15, errs
    assert all(errs[q + 1] < errs[q] for q in range(12, 16)), errs
    for d in (dA, dB, dC):
        d.free()


# ------------------------------------------------------------ serving trace

ELEM = {DTYPE_F64: 8, DTYPE_F32: 4, DTYPE_C64: 16, DTYPE_C32: 8, DTYPE_I8: 1}
OP_NP = {DTYPE_F64: np.float64, DTYPE_F32: np.float32, DTYPE_C64: np.complex128,
         DTYPE_C32: np.complex64, DTYPE_I8: np.int8}


def _single(q, out):
    L = lib()
    if q.dtype == DTYPE_F64:
        st = L.ferray_gemm_f64_dev(q.m, q.n, q.k, q.alpha_re, q.a, q.b, 0.0, out.ptr, None)
    elif q.dtype == DTYPE_F32:
        st = L.ferray_gemm_f32_dev(q.m, q.n, q.k, q.alpha_re, q.a, q.b, 0.0, out.ptr, None)
    elif q.dtype == DTYPE_C64:
        st = L.ferray_gemm_c64_dev(q.m, q.n, q.k, q.alpha_re, q.alpha_im, q.a, q.b, 0.0, 0.0,
                                   out.ptr, None)
    else:
        st = L.ferray_gemm_i8_signed_dev(q.m, q.n, q.k, q.a, q.b, out.ptr, False, None)
    _ok(st, "single entry")
    fa.sync()
    return out.download(sv.c_dtype(q.dtype), q.m * q.n)


def _operand(pools, q, ptr, count):
    buf = pools.buffer(q.dtype)
    return buf.download(OP_NP[q.dtype], count, ptr - buf.ptr)


def _numpy_check(pools, q, c):
    A = _operand(pools, q, q.a, q.m * q.k).reshape(q.m, q.k)
    B = _operand(pools, q, q.b, q.k * q.n).reshape(q.k, q.n)
    c = c.reshape(q.m, q.n)
    if q.dtype == DTYPE_I8:
        assert np.array_equal(c, wrap_i32(_exact_f64(A, B, q.m, q.n, q.k)))
        return
    u = U64 if q.dtype in (DTYPE_F64, DTYPE_C64) else U32
    if q.dtype == DTYPE_C64:
        mag = (np.abs(A.real) + np.abs(A.imag)) @ (np.abs(B.real) + np.abs(B.imag))
        ref = complex(q.alpha_re, q.alpha_im) * (A @ B)
        bound = 4 * (q.k + 2) * u * abs(complex(q.alpha_re, q.alpha_im)) * mag
        assert np.all(np.abs(c - ref) <= bound)
    else:
        A64, B64 = A.astype(np.float64), B.astype(np.float64)
        ref = q.alpha_re * (A64 @ B64)
        bound = 2 * (q.k + 2) * u * abs(q.alpha_re) * (np.abs(A64) @ np.abs(B64))
        assert np.all(np.abs(c.astype(np.float64) - ref) <= bound)


def test_serving_trace_2000_requests_every_result_checked(gpu):
    pools = sv.OperandPools(fa, elems=1 << 24, seed=31)
    try:
        reqs, arrival, _ = sv.make_trace(pools, 2000, rate_per_s=2e4, seed=5, lo=16, hi=2048)
        cfg = sv.config(hbm_bytes=4 << 30, max_resident=4096, max_batch=256, lanes=4,
                        keep_results=True)
        with sv.Server(cfg) as srv:
            t0 = time.perf_counter()
            tickets = []
            for q, at in zip(reqs, arrival):
                while time.perf_counter() - t0 < at:
                    pass
                tickets.append(srv.submit(q))
            got = [srv.result(t, q.dtype, q.m, q.n) for t, q in zip(tickets, reqs)]
            st = srv.stats()
        assert st.completed == 2000
        assert st.launches < 2000 and st.launched_requests == 2000, (st.launches,
                                                                    st.launched_requests)
        biggest = max(q.m * q.n * 16 for q in reqs)
        out = fa.DeviceBuffer(biggest)
        for i, (q, c) in enumerate(zip(reqs, got)):
            alone = _single(q, out)
            assert c.tobytes() == alone.tobytes(), (i, q.dtype, q.m, q.n, q.k)
        out.free()
        for i in rng(6).choice(len(reqs), 60, replace=False):
            _numpy_check(pools, reqs[i], got[i])
    finally:
        pools.free()
