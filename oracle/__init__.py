"""ferray GEMM oracle — TEST INFRASTRUCTURE ONLY.

ctypes face of liboracle (ferray_oracle.c), the CPU restatement of
ferray-linalg's GEMM arithmetic.  Only tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg may import this package, and only as the checker
or the timed CPU reference.  The product (ferray_amd) never imports it.

Every function below names the reference code it restates; how each one is
pinned (golden vectors, known-answer tests, closed forms) is in
tests/test_oracle.py.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(ORACLE_DIR, "libferray_oracle.so")

_sz, _p, _d, _f, _i = C.c_size_t, C.c_void_p, C.c_double, C.c_float, C.c_int
_SIGS = {
    "orc_lcg_matrix": (None, [_sz, C.c_uint64, _p]),
    "orc_splitmix64": (C.c_uint64, [C.c_uint64, C.c_uint64]),
    "orc_uniform_f64": (None, [_sz, C.c_uint64, _p]),
    "orc_uniform_int": (None, [_sz, C.c_uint64, C.c_int64, C.c_int64, _p]),
    "orc_gemm_f64": (None, [_sz, _sz, _sz, _d, _p, _p, _d, _p]),
    "orc_gemm_f32": (None, [_sz, _sz, _sz, _f, _p, _p, _f, _p]),
    "orc_gemm_c64": (None, [_sz, _sz, _sz, _d, _d, _p, _p, _d, _d, _p]),
    "orc_gemm_c32": (None, [_sz, _sz, _sz, _f, _f, _p, _p, _f, _f, _p]),
    "orc_gemm_i8": (None, [_sz, _sz, _sz, _p, _p, _p, _i]),
    "orc_gemm_i8_signed": (None, [_sz, _sz, _sz, _p, _p, _p, _i]),
    "orc_gemm_i16": (None, [_sz, _sz, _sz, _p, _p, _p, _i]),
    "orc_gemm_i8_avx2_saturating": (None, [_sz, _sz, _sz, _p, _p, _p, _i]),
    "orc_quantized_matmul": (None, [_sz, _sz, _sz, _p, C.c_uint8, _f, _p, C.c_int8, _f, _p, _i]),
    "orc_bf16_to_f32": (_f, [C.c_uint16]),
    "orc_f16_to_f32": (_f, [C.c_uint16]),
    "orc_gemm_bf16_f32": (None, [_sz, _sz, _sz, _p, _p, _p, _i]),
    "orc_gemm_f16_f32": (None, [_sz, _sz, _sz, _p, _p, _p, _i]),
    "orc_e4m3_to_f64": (_d, [C.c_uint8]),
    "orc_e5m2_to_f64": (_d, [C.c_uint8]),
    "orc_gemm_fp8_f64": (None, [_sz, _sz, _sz, _p, _p, _i, _p, _i]),
    "orc_dd_gemm_f64_sampled": (None, [_sz, _sz, _sz, _p, _p, _sz, _p, _sz, _p, _p, _p]),
    "orc_dd_gemm_c64_sampled": (None, [_sz, _sz, _sz, _p, _p, _sz, _p, _sz, _p, _p]),
    "orc_gemm_f64_goto": (_i, [_sz, _sz, _sz, _d, _p, _p, _d, _p, _i]),
    "orc_gemm_f32_goto": (_i, [_sz, _sz, _sz, _f, _p, _p, _f, _p, _i]),
    "orc_gemm_i8_signed_threaded": (_i, [_sz, _sz, _sz, _p, _p, _p, _i]),
    "orc_cpu_has_avx2_fma": (_i, []),
    "orc_trsm_left_lower_f64": (None, [_sz, _sz, _d, _p, _sz, _p, _sz, _i]),
    "orc_trsm_left_upper_f64": (None, [_sz, _sz, _d, _p, _sz, _p, _sz, _i]),
    "orc_trsm_left_lower_f32": (None, [_sz, _sz, _f, _p, _sz, _p, _sz, _i]),
    "orc_trsm_left_upper_f32": (None, [_sz, _sz, _f, _p, _sz, _p, _sz, _i]),
}
_lib = None


def build() -> str:
    """Compiles the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def _p_(a: np.ndarray) -> int:
    assert a.flags.c_contiguous
    return a.ctypes.data


# ----------------------------------------------------------------- generators

def lcg_matrix(n: int, seed: int) -> np.ndarray:
    """ferray-linalg/examples/bench_matmul.rs:21-34 make_matrix(n, seed)."""
    out = np.empty(n * n, dtype=np.float64)
    lib().orc_lcg_matrix(out.size, seed, _p_(out))
    return out


def uniform_f64(count: int, seed: int) -> np.ndarray:
    out = np.empty(count, dtype=np.float64)
    lib().orc_uniform_f64(count, seed, _p_(out))
    return out


def uniform_int(count: int, seed: int, lo: int, hi: int) -> np.ndarray:
    out = np.empty(count, dtype=np.int64)
    lib().orc_uniform_int(count, seed, lo, hi, _p_(out))
    return out


# ------------------------------------------------------------ GEMM restatements

def gemm_f64(m, n, k, alpha, a, b, beta, c) -> None:
    """ferray KC-block formula, f64 (outer.rs:198-488, kernel_avx2.rs:53-291)."""
    lib().orc_gemm_f64(m, n, k, alpha, _p_(a), _p_(b), beta, _p_(c))


def gemm_f32(m, n, k, alpha, a, b, beta, c) -> None:
    """ferray KC-block formula, f32 (outer.rs:499-687, kernel_avx2_f32.rs)."""
    lib().orc_gemm_f32(m, n, k, alpha, _p_(a), _p_(b), beta, _p_(c))


def gemm_c64(m, n, k, ar, ai, a, b, br, bi, c) -> None:
    """Interior-tile ZGEMM formula (kernel_avx2_complex_f64.rs:4-230)."""
    lib().orc_gemm_c64(m, n, k, ar, ai, _p_(a), _p_(b), br, bi, _p_(c))


def gemm_c32(m, n, k, ar, ai, a, b, br, bi, c) -> None:
    """Interior-tile CGEMM formula (kernel_avx2_complex_f32.rs)."""
    lib().orc_gemm_c32(m, n, k, ar, ai, _p_(a), _p_(b), br, bi, _p_(c))


def gemm_i8(m, n, k, a, b, c, accumulate) -> None:
    """u8 x i8 -> i32 wrapping (faer_fallback.rs:278-310 closed form)."""
    lib().orc_gemm_i8(m, n, k, _p_(a), _p_(b), _p_(c), int(accumulate))


def gemm_i8_signed(m, n, k, a, b, c, accumulate) -> None:
    """i8 x i8 -> i32 wrapping (integer_tests.rs:27-37 closed form)."""
    lib().orc_gemm_i8_signed(m, n, k, _p_(a), _p_(b), _p_(c), int(accumulate))


def gemm_i16(m, n, k, a, b, c, accumulate) -> None:
    """i16 x i16 -> i32 wrapping (integer_tests.rs:15-25 closed form)."""
    lib().orc_gemm_i16(m, n, k, _p_(a), _p_(b), _p_(c), int(accumulate))


def gemm_i8_avx2_saturating(m, n, k, a, b, c, accumulate) -> None:
    """Default-AVX2 u8xi8 interior tile with i16-saturating pair sums
    (kernel_avx2_i8.rs:1-36) — documents the divergence only."""
    lib().orc_gemm_i8_avx2_saturating(m, n, k, _p_(a), _p_(b), _p_(c), int(accumulate))


def quantized_matmul(m, n, k, a, a_zp, sa, b, b_zp, sb, c, accumulate) -> None:
    """mod.rs:484-575."""
    lib().orc_quantized_matmul(m, n, k, _p_(a), int(a_zp), sa, _p_(b), int(b_zp), sb, _p_(c),
                               int(accumulate))


def gemm_bf16_f32(m, n, k, a, b, c, accumulate) -> None:
    """mod.rs:851-882: widen, then the f32 formula with beta in {0, 1}."""
    lib().orc_gemm_bf16_f32(m, n, k, _p_(a), _p_(b), _p_(c), int(accumulate))


def gemm_f16_f32(m, n, k, a, b, c, accumulate) -> None:
    """mod.rs:906-933."""
    lib().orc_gemm_f16_f32(m, n, k, _p_(a), _p_(b), _p_(c), int(accumulate))


def bf16_to_f32(h: int) -> float:
    return lib().orc_bf16_to_f32(h)


def f16_to_f32(h: int) -> float:
    return lib().orc_f16_to_f32(h)


def e4m3_to_f64(x: int) -> float:
    """OCP E4M3 byte -> value (no reference counterpart; OCP FP8 spec v1.0)."""
    return lib().orc_e4m3_to_f64(x)


def e5m2_to_f64(x: int) -> float:
    """OCP E5M2 byte -> value (OCP FP8 spec v1.0)."""
    return lib().orc_e5m2_to_f64(x)


def gemm_fp8_f64(m, n, k, a, b, fmt: str, c, accumulate) -> None:
    """C (+)= A B over fp8 bytes (fmt 'e4m3' or 'e5m2') summed in double; the
    fp8 member of the family has no reference entry point, so this is the
    exact product (to double rounding) the device result is measured against."""
    assert c.dtype == np.float64
    lib().orc_gemm_fp8_f64(m, n, k, _p_(a), _p_(b), 0 if fmt == "e4m3" else 1, _p_(c),
                           int(accumulate))


# ------------------------------------------------------------ error reference

def dd_gemm_f64_sampled(m, n, k, a, b, rows, cols):
    """Compensated (Dot2) C[rows][:, cols] for f64 A, B: (hi, lo)."""
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    hi = np.empty((rows.size, cols.size), dtype=np.float64)
    lo = np.empty_like(hi)
    lib().orc_dd_gemm_f64_sampled(m, n, k, _p_(a), _p_(b), rows.size, _p_(rows), cols.size,
                                  _p_(cols), _p_(hi), _p_(lo))
    return hi, lo


def dd_gemm_c64_sampled(m, n, k, a, b, rows, cols):
    """Compensated complex C[rows][:, cols] (interleaved f64 in, complex128 out)."""
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    out = np.empty((rows.size, cols.size), dtype=np.complex128)
    lib().orc_dd_gemm_c64_sampled(m, n, k, _p_(a), _p_(b), rows.size, _p_(rows), cols.size,
                                  _p_(cols), _p_(out.view(np.float64)))
    return out


# ------------------------------------------------------------ timed CPU baseline

def gemm_f64_goto(m, n, k, alpha, a, b, beta, c, workers: int = 0) -> int:
    """Goto 5-loop restatement with AVX2 4x8 micro-kernel and pick_grid
    (outer.rs:66-488); returns the worker count used."""
    return lib().orc_gemm_f64_goto(m, n, k, alpha, _p_(a), _p_(b), beta, _p_(c), workers)


def gemm_f32_goto(m, n, k, alpha, a, b, beta, c, workers: int = 0) -> int:
    return lib().orc_gemm_f32_goto(m, n, k, alpha, _p_(a), _p_(b), beta, _p_(c), workers)


def gemm_i8_signed_threaded(m, n, k, a, b, c, workers: int = 0) -> int:
    return lib().orc_gemm_i8_signed_threaded(m, n, k, _p_(a), _p_(b), _p_(c), workers)


def cpu_has_avx2_fma() -> bool:
    return bool(lib().orc_cpu_has_avx2_fma())


# ------------------------------------------------------------ triangular solve

def trsm(lower: bool, m, n, alpha, a, lda, b, ldb, unit_diag) -> None:
    """trsm.rs restated (NB = 32 block recursion, scalar updates); in place."""
    f32 = a.dtype == np.float32
    name = f"orc_trsm_left_{'lower' if lower else 'upper'}_{'f32' if f32 else 'f64'}"
    getattr(lib(), name)(m, n, alpha, _p_(a), lda, _p_(b), ldb, int(bool(unit_diag)))
