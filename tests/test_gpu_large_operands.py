"""Operands and results larger than 4 GiB.

An MI355X holds 288 GB, so one operand of a product can exceed 2^32 bytes.
Every kernel forms its addresses in 64-bit arithmetic (32-bit offsets are
used only inside one block tile, where the launchers check they fit); these
tests run each product family with one matrix past 4 GiB and check rows (or
columns) of C whose data lie on both sides of the 2^31- and 2^32-byte
offsets, and the last rows and columns.

Bars:
  * integers: equal to the exact product (float64 BLAS products, exact since
    every partial sum stays below 2^53) on every checked element, reduced
    mod 2^32 for i16;
  * f64 / f32: |C - C_oracle| <= 2(k + 2) u |A||B| against the oracle's
    restatement of ferray's block formula on every checked element;
  * Ozaki (14 moduli): |C - C_dd| <= 2^(6 - g) ||a_i|| ||b_j|| + 4 u |C_dd|
    against the compensated reference (g the scaling exponent of the modulus
    count), as at the benchmark sizes.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import ferray_amd as fa
import oracle
from ferray_amd._lib import lib
from tests.helpers import rng, wrap_i32

pytestmark = pytest.mark.gpu

U64, U32 = 2.0 ** -53, 2.0 ** -24
GiB4 = 1 << 32


def _ok(status, what):
    assert status == 0, f"{what}: {lib().ferray_status_string(status).decode()}"


def _filled(count, kind, seed, itemsize, lo=0, hi=0):
    buf = fa.DeviceBuffer(count * itemsize)
    fa.fill(buf, count, kind, seed, lo, hi)
    fa.sync()
    return buf


def _lines(count, pitch_bytes, seed, extra=24):
    """Row (or column) indices around the 2^31- and 2^32-byte offsets of a
    matrix whose lines are pitch_bytes apart, its first and last lines and a
    few random ones."""
    picks = {0, 1, 127, 128, count - 2, count - 1}
    for edge in (1 << 31, 1 << 32):
        r = edge // pitch_bytes
        picks.update(x for x in (r - 1, r, r + 1) if 0 <= x < count)
    picks.update(int(x) for x in rng(seed).choice(count, extra, replace=False))
    return np.array(sorted(picks), dtype=np.int64)


def _rows(buf, dtype, rows, width):
    item = np.dtype(dtype).itemsize
    return np.stack([buf.download(dtype, width, int(r) * width * item) for r in rows])


def _exact(a, b):
    return np.rint(a.astype(np.float64) @ b.astype(np.float64)).astype(np.int64)


# ------------------------------------------------------------------ integers

@pytest.mark.parametrize("accumulate", [False, True])
def test_i8_a_past_4gib(gpu, accumulate):
    m, n, k = 69632, 256, 65536          # A: 4.56 GB
    assert m * k > GiB4
    dA = _filled(m * k, fa.FILL_I8, 71, 1, -128, 127)
    dB = _filled(k * n, fa.FILL_I8, 72, 1, -128, 127)
    rows = _lines(m, k, 1)
    c0 = rng(2).integers(-2 ** 31, 2 ** 31, m * n).astype(np.int32)
    dC = fa.DeviceBuffer.from_array(c0)
    _ok(lib().ferray_gemm_i8_signed_dev(m, n, k, dA.ptr, dB.ptr, dC.ptr, accumulate, None), "i8")
    fa.sync()
    a = _rows(dA, np.int8, rows, k)
    b = dB.download(np.int8, k * n).reshape(k, n)
    c = _rows(dC, np.int32, rows, n)
    for d in (dA, dB, dC):
        d.free()
    want = _exact(a, b) + (c0.reshape(m, n)[rows].astype(np.int64) if accumulate else 0)
    bad = np.argwhere(c != wrap_i32(want))
    assert bad.size == 0, (len(bad), rows[bad[0][0]], bad[0][1])


def test_i8_b_past_4gib(gpu):
    m, n, k = 256, 69632, 65536          # B: 4.56 GB
    assert k * n > GiB4
    dA = _filled(m * k, fa.FILL_I8, 73, 1, -128, 127)
    dB = _filled(k * n, fa.FILL_I8, 74, 1, -128, 127)
    dC = fa.DeviceBuffer(m * n * 4)
    _ok(lib().ferray_gemm_i8_signed_dev(m, n, k, dA.ptr, dB.ptr, dC.ptr, False, None), "i8")
    fa.sync()
    cols = _lines(n, 1, 3)
    a = dA.download(np.int8, m * k).reshape(m, k)
    bcols = dB.download(np.int8, k * n).reshape(k, n)[:, cols]
    c = dC.download(np.int32, m * n).reshape(m, n)[:, cols]
    for d in (dA, dB, dC):
        d.free()
    bad = np.argwhere(c != wrap_i32(_exact(a, bcols)))
    assert bad.size == 0, (len(bad), bad[0][0], cols[bad[0][1]])


def test_i8_c_past_4gib(gpu):
    m = n = 36864
    k = 128                              # C: 5.4 GB of int32
    assert m * n * 4 > GiB4
    dA = _filled(m * k, fa.FILL_I8, 75, 1, -128, 127)
    dB = _filled(k * n, fa.FILL_I8, 76, 1, -128, 127)
    dC = fa.DeviceBuffer(m * n * 4)
    _ok(lib().ferray_gemm_i8_signed_dev(m, n, k, dA.ptr, dB.ptr, dC.ptr, False, None), "i8")
    fa.sync()
    rows = _lines(m, n * 4, 4)
    a = dA.download(np.int8, m * k).reshape(m, k)[rows]
    b = dB.download(np.int8, k * n).reshape(k, n)
    c = _rows(dC, np.int32, rows, n)
    for d in (dA, dB, dC):
        d.free()
    bad = np.argwhere(c != wrap_i32(_exact(a, b)))
    assert bad.size == 0, (len(bad), rows[bad[0][0]], bad[0][1])


def test_i16_a_past_4gib(gpu):
    m, n, k = 36864, 256, 65536          # A: 4.83 GB of int16
    assert m * k * 2 > GiB4
    dA = _filled(m * k, fa.FILL_I16, 77, 2, -32768, 32767)
    dB = _filled(k * n, fa.FILL_I16, 78, 2, -32768, 32767)
    dC = fa.DeviceBuffer(m * n * 4)
    _ok(lib().ferray_gemm_i16_dev(m, n, k, dA.ptr, dB.ptr, dC.ptr, False, None), "i16")
    fa.sync()
    rows = _lines(m, k * 2, 5)
    a = _rows(dA, np.int16, rows, k)
    b = dB.download(np.int16, k * n).reshape(k, n)
    c = _rows(dC, np.int32, rows, n)
    for d in (dA, dB, dC):
        d.free()
    bad = np.argwhere(c != wrap_i32(_exact(a, b)))
    assert bad.size == 0, (len(bad), rows[bad[0][0]], bad[0][1])


# ------------------------------------------------------------------ floats

def _bound(a, b, k, u):
    return 2 * (k + 2) * u * (np.abs(a.astype(np.float64)) @ np.abs(b.astype(np.float64)))


def _oracle_rows(a, b, k, dtype):
    r, n = a.shape[0], b.shape[1]
    ref = np.zeros(r * n, dtype)
    (oracle.gemm_f64 if dtype == np.float64 else oracle.gemm_f32)(
        r, n, k, dtype(1.0), np.ascontiguousarray(a).ravel(), np.ascontiguousarray(b).ravel(),
        dtype(0.0), ref)
    return ref.reshape(r, n)


@pytest.mark.parametrize("dtype,m,n,k", [
    (np.float64, 36864, 640, 16384),     # A: 4.83 GB, 128 x 128 tiles
    (np.float32, 65536, 256, 17408),     # A: 4.56 GB, 256 x 256 tiles
])
def test_real_a_past_4gib(gpu, dtype, m, n, k):
    f64 = dtype == np.float64
    item = np.dtype(dtype).itemsize
    assert m * k * item > GiB4
    kind = fa.FILL_F64_NORMAL if f64 else fa.FILL_F32_NORMAL
    dA = _filled(m * k, kind, 81, item)
    dB = _filled(k * n, kind, 82, item)
    dC = fa.DeviceBuffer(m * n * item)
    run = lib().ferray_gemm_f64_dev if f64 else lib().ferray_gemm_f32_dev
    _ok(run(m, n, k, 1.0, dA.ptr, dB.ptr, 0.0, dC.ptr, None), "gemm")
    fa.sync()
    rows = _lines(m, k * item, 6)
    a = _rows(dA, dtype, rows, k)
    b = dB.download(dtype, k * n).reshape(k, n)
    c = _rows(dC, dtype, rows, n).astype(np.float64)
    for d in (dA, dB, dC):
        d.free()
    err = np.abs(c - _oracle_rows(a, b, k, dtype))
    bound = _bound(a, b, k, U64 if f64 else U32)
    assert np.all(err <= bound), float(np.max(err / bound))


def test_f64_b_past_4gib(gpu):
    m, n, k = 640, 36864, 16384          # B: 4.83 GB
    assert k * n * 8 > GiB4
    dA = _filled(m * k, fa.FILL_F64_NORMAL, 83, 8)
    dB = _filled(k * n, fa.FILL_F64_NORMAL, 84, 8)
    dC = fa.DeviceBuffer(m * n * 8)
    _ok(lib().ferray_gemm_f64_dev(m, n, k, 1.0, dA.ptr, dB.ptr, 0.0, dC.ptr, None), "gemm")
    fa.sync()
    cols = _lines(n, 8, 7)
    rows = np.sort(rng(8).choice(m, 32, replace=False))
    a = dA.download(np.float64, m * k).reshape(m, k)[rows]
    bcols = np.ascontiguousarray(dB.download(np.float64, k * n).reshape(k, n)[:, cols])
    c = dC.download(np.float64, m * n).reshape(m, n)[np.ix_(rows, cols)]
    for d in (dA, dB, dC):
        d.free()
    err = np.abs(c - _oracle_rows(a, bcols, k, np.float64))
    bound = _bound(a, bcols, k, U64)
    assert np.all(err <= bound), float(np.max(err / bound))


def test_f64_c_past_4gib(gpu):
    m = n = 24576
    k = 64                               # C: 4.83 GB
    assert m * n * 8 > GiB4
    dA = _filled(m * k, fa.FILL_F64_NORMAL, 85, 8)
    dB = _filled(k * n, fa.FILL_F64_NORMAL, 86, 8)
    c0 = _filled(m * n, fa.FILL_F64_NORMAL, 87, 8)
    _ok(lib().ferray_gemm_f64_dev(m, n, k, 1.0, dA.ptr, dB.ptr, 0.5, c0.ptr, None), "gemm")
    fa.sync()
    rows = _lines(m, n * 8, 9)
    a = dA.download(np.float64, m * k).reshape(m, k)[rows]
    b = dB.download(np.float64, k * n).reshape(k, n)
    c = _rows(c0, np.float64, rows, n)
    for d in (dA, dB, c0):
        d.free()
    # C0 regenerated from the same seeded stream: the oracle's generator is
    # bit-identical to the device's (checked in test_gpu_configs.py)
    cin = _filled(m * n, fa.FILL_F64_NORMAL, 87, 8)
    cin_rows = _rows(cin, np.float64, rows, n)
    cin.free()
    ref = np.ascontiguousarray(cin_rows).ravel().copy()
    oracle.gemm_f64(len(rows), n, k, 1.0, np.ascontiguousarray(a).ravel(), b.ravel(), 0.5, ref)
    err = np.abs(c - ref.reshape(len(rows), n))
    bound = _bound(a, b, k, U64) + 4 * U64 * np.abs(cin_rows)
    assert np.all(err <= bound), float(np.max(err / bound))


# ------------------------------------------------------------------ Ozaki

OZ_MODULI = [256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199]


def test_ozaki_a_past_4gib(gpu):
    m, n, k = 36864, 256, 16384          # A: 4.83 GB; 14 residue planes of A: 8.5 GB
    assert m * k * 8 > GiB4
    nm = 14
    dA = _filled(m * k, fa.FILL_F64_NORMAL, 91, 8)
    dB = _filled(k * n, fa.FILL_F64_NORMAL, 92, 8)
    dC = fa.DeviceBuffer(m * n * 8)
    _ok(lib().ferray_dgemm_ozaki_dev(m, n, k, 1.0, dA.ptr, dB.ptr, 0.0, dC.ptr, nm, None),
        "ozaki")
    fa.sync()
    rows = _lines(m, k * 8, 10)
    a = np.ascontiguousarray(_rows(dA, np.float64, rows, k))
    b = dB.download(np.float64, k * n)
    c = _rows(dC, np.float64, rows, n)
    for d in (dA, dB, dC):
        d.free()
    hi, lo = oracle.dd_gemm_f64_sampled(len(rows), n, k, a.ravel(), b, np.arange(len(rows)),
                                        np.arange(n))
    d = (c - hi) - lo
    M = math.prod(OZ_MODULI[:nm])
    g = (M.bit_length() - 1 - 2) // 2
    ra = np.linalg.norm(a, axis=1)[:, None]
    cb = np.linalg.norm(b.reshape(k, n), axis=0)[None, :]
    lim = 2.0 ** (6 - g) * ra * cb + 4 * U64 * np.abs(hi)
    assert np.all(np.abs(d) <= lim), float(np.max(np.abs(d) / lim))


# ------------------------------------------------------------------ long k
#
# A grid's y extent is at most 65535 blocks; the passes whose grid y walks k
# (64- or 128-deep tiles) are launched in slices past that.  k just above
# 2^22 (and 2^23 for Ozaki's 128-deep tiles), ragged, on thin products.

KLONG = (1 << 22) + 1000


def _upload(arr):
    return fa.DeviceBuffer.from_array(np.ascontiguousarray(arr))


def _ones(shape, seed):
    """Entries in {-1, 0, 1}: every product and partial sum below 2^24 is
    exact in f32, so float results must equal the integer product."""
    return rng(seed).integers(-1, 2, shape).astype(np.int64)


@pytest.mark.parametrize("kind", ["i8", "i16"])
def test_integer_long_k(gpu, kind):
    m, n, k = 8, 24, KLONG
    dt, lo, hi, fill = ((np.int8, -128, 127, fa.FILL_I8) if kind == "i8"
                        else (np.int16, -32768, 32767, fa.FILL_I16))
    item = np.dtype(dt).itemsize
    dA = _filled(m * k, fill, 101, item, lo, hi)
    dB = _filled(k * n, fill, 102, item, lo, hi)
    dC = fa.DeviceBuffer(m * n * 4)
    run = lib().ferray_gemm_i8_signed_dev if kind == "i8" else lib().ferray_gemm_i16_dev
    _ok(run(m, n, k, dA.ptr, dB.ptr, dC.ptr, False, None), kind)
    fa.sync()
    a = dA.download(dt, m * k).reshape(m, k)
    b = dB.download(dt, k * n).reshape(k, n)
    c = dC.download(np.int32, m * n).reshape(m, n)
    for d in (dA, dB, dC):
        d.free()
    # |sum| <= k 2^30 < 2^53: the float64 product is exact
    assert np.array_equal(c, wrap_i32(_exact(a, b)))


@pytest.mark.parametrize("kind", ["bf16", "e4m3", "bf16x3"])
def test_float_long_k(gpu, kind):
    m, n, k = 8, 24, KLONG
    a, b = _ones((m, k), 103), _ones((k, n), 104)
    want = (a.astype(np.float64) @ b.astype(np.float64)).astype(np.float32)
    if kind == "bf16":   # -1, 0, 1 as bf16 bit patterns
        code = {-1: 0xBF80, 0: 0, 1: 0x3F80}
        enc = lambda x: np.vectorize(code.get, otypes=[np.uint16])(x)
        run = lambda da, db, dc: lib().ferray_gemm_bf16_f32_dev(m, n, k, da, db, dc, False, None)
    elif kind == "e4m3":  # OCP E4M3: 1.0 = 0x38
        code = {-1: 0xB8, 0: 0, 1: 0x38}
        enc = lambda x: np.vectorize(code.get, otypes=[np.uint8])(x)
        run = lambda da, db, dc: lib().ferray_gemm_e4m3_f32_dev(m, n, k, da, db, dc, False, None)
    else:
        enc = lambda x: x.astype(np.float32)
        run = lambda da, db, dc: lib().ferray_sgemm_bf16x3_dev(m, n, k, 1.0, da, db, 0.0, dc, 3,
                                                               None)
    dA, dB = _upload(enc(a)), _upload(enc(b))
    dC = fa.DeviceBuffer(m * n * 4)
    _ok(run(dA.ptr, dB.ptr, dC.ptr), kind)
    fa.sync()
    c = dC.download(np.float32, m * n).reshape(m, n)
    for d in (dA, dB, dC):
        d.free()
    assert np.array_equal(c, want)


@pytest.mark.parametrize("m,n,k", [
    (8, 16, (1 << 23) + 1000),           # B's residue pass: Kp / 128 > 65535 tiles
    (512, 512, 70000),                   # two k chunks on the 256 x 256 kernel
])
def test_ozaki_long_k(gpu, m, n, k):
    """k past one chunk of 2^16: later chunks add their products to the
    stored residues mod m."""
    nm = 14
    dA = _filled(m * k, fa.FILL_F64_NORMAL, 105, 8)
    dB = _filled(k * n, fa.FILL_F64_NORMAL, 106, 8)
    dC = fa.DeviceBuffer(m * n * 8)
    _ok(lib().ferray_dgemm_ozaki_dev(m, n, k, 1.0, dA.ptr, dB.ptr, 0.0, dC.ptr, nm, None),
        "ozaki")
    fa.sync()
    rows = np.arange(m) if m <= 16 else np.sort(rng(11).choice(m, 16, replace=False))
    a = np.ascontiguousarray(_rows(dA, np.float64, rows, k))
    b = dB.download(np.float64, k * n)
    c = _rows(dC, np.float64, rows, n)
    for d in (dA, dB, dC):
        d.free()
    hi, lo = oracle.dd_gemm_f64_sampled(len(rows), n, k, a.ravel(), b, np.arange(len(rows)),
                                        np.arange(n))
    d = (c - hi) - lo
    M = math.prod(OZ_MODULI[:nm])
    g = (M.bit_length() - 1 - 2) // 2
    ra = np.linalg.norm(a, axis=1)[:, None]
    cb = np.linalg.norm(b.reshape(k, n), axis=0)[None, :]
    lim = 2.0 ** (6 - g) * ra * cb + 4 * U64 * np.abs(hi)
    assert np.all(np.abs(d) <= lim), float(np.max(np.abs(d) / lim))
