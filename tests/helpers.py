"""Shared helpers for the test suite (inputs, error norms, hex floats)."""
from __future__ import annotations

import numpy as np


def hexf(s: str) -> float:
    return float.fromhex(s)


def ulp_diff(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """|ulp distance| between two f64 arrays of the same sign pattern."""
    xi = np.asarray(x, dtype=np.float64).view(np.int64).astype(object)
    yi = np.asarray(y, dtype=np.float64).view(np.int64).astype(object)

    def ordered(v):
        return v if v >= 0 else -(v & 0x7FFFFFFFFFFFFFFF)
    return np.array([abs(ordered(a) - ordered(b)) for a, b in zip(xi.ravel(), yi.ravel())])


def normwise_error(c_hat: np.ndarray, c_ref: np.ndarray, a: np.ndarray, b: np.ndarray):
    """(||C^ - C||_F / (||A||_F ||B||_F), ||C^ - C||_F / ||C||_F)."""
    d = np.linalg.norm((np.asarray(c_hat, np.float64) - np.asarray(c_ref, np.float64)).ravel())
    na = np.linalg.norm(np.asarray(a, np.float64).ravel())
    nb = np.linalg.norm(np.asarray(b, np.float64).ravel())
    nc = np.linalg.norm(np.asarray(c_ref, np.float64).ravel())
    return d / (na * nb), (d / nc if nc else d)


def rng(seed: int) -> np.random.Generator:
    return np.random.default_rng(seed)


def wrap_i32(x: np.ndarray) -> np.ndarray:
    """Reduce int64 / object integers mod 2^32 into int32 (two's complement)."""
    return (np.asarray(x).astype(np.int64) & 0xFFFFFFFF).astype(np.uint32).view(np.int32)


def exact_int_matmul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Exact integer product in int64 (all test sizes keep |sum| < 2^63)."""
    return a.astype(np.int64) @ b.astype(np.int64)
