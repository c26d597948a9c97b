"""Grouped (batched, mixed-dtype) GEMM on device-resident operands.

``grouped(problems)`` enqueues many independent problems in one persistent
kernel launch; ``Plan`` keeps reusable launch state (descriptor buffers and a
captured hipGraph) for callers with several batches in flight.  The tile
shape of each problem comes from a table the autotuner writes
(``ferray_amd/tile_table.json``), loaded when this module is imported.
"""
from __future__ import annotations

import ctypes as C
import json
import os

from ._lib import (DTYPE_C32, DTYPE_C64, DTYPE_F32, DTYPE_F64, DTYPE_I8, GemmProblem, check,
                   lib)

__all__ = ["GemmProblem", "Plan", "grouped", "problem", "load_tile_table", "tile_table",
           "set_tile_table", "DTYPE_F64", "DTYPE_F32", "DTYPE_C64", "DTYPE_C32", "DTYPE_I8",
           "TABLE_PATH"]

TABLE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tile_table.json")
N_DTYPES, N_BUCKETS, N_SMALL, N_KCLS = 5, 8, 3, 3


def problem(dtype: int, m: int, n: int, k: int, a: int, b: int, c: int, alpha=1.0, beta=0.0,
            accumulate: bool = False) -> GemmProblem:
    """Builds one problem descriptor; a, b, c are device addresses."""
    alpha = complex(alpha)
    beta = complex(beta)
    return GemmProblem(dtype, int(bool(accumulate)), m, n, k, a, b, c, alpha.real, alpha.imag,
                       beta.real, beta.imag)


def _array(problems):
    if isinstance(problems, C.Array):
        return problems
    arr = (GemmProblem * len(problems))()
    for i, p in enumerate(problems):
        arr[i] = p
    return arr


def grouped(problems, stream=None) -> None:
    arr = _array(problems)
    check(lib().ferray_gemm_grouped_dev(C.cast(arr, C.c_void_p), len(arr), stream), "grouped")


class Plan:
    """A reusable grouped launch (own descriptor buffers + captured graph)."""

    def __init__(self, max_problems: int, max_tiles: int):
        p = C.c_void_p()
        check(lib().ferray_grouped_plan_create(max_problems, max_tiles, C.byref(p)), "plan_create")
        self.ptr = p.value

    def launch(self, problems, stream=None) -> None:
        arr = _array(problems)
        check(lib().ferray_grouped_plan_launch(self.ptr, C.cast(arr, C.c_void_p), len(arr),
                                               stream), "plan_launch")

    def sync(self) -> None:
        check(lib().ferray_grouped_plan_sync(self.ptr), "plan_sync")

    def close(self) -> None:
        if self.ptr:
            lib().ferray_grouped_plan_destroy(self.ptr)
            self.ptr = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


def tile_table() -> list:
    """The table as nested lists [dtype][size bucket][small-side class][k class]
    (0 = 64x64 tiles, 1 = 128x128; see include/ferray_gemm.h)."""
    n = N_DTYPES * N_BUCKETS * N_SMALL * N_KCLS
    buf = (C.c_int * n)()
    check(lib().ferray_get_tile_table(C.cast(buf, C.c_void_p), n), "get_tile_table")
    flat = list(buf)
    it = iter(flat)
    return [[[[next(it) for _ in range(N_KCLS)] for _ in range(N_SMALL)]
             for _ in range(N_BUCKETS)] for _ in range(N_DTYPES)]


def _flatten(x):
    if isinstance(x, (list, tuple)):
        for y in x:
            yield from _flatten(y)
    else:
        yield int(x)


def set_tile_table(table) -> None:
    """Installs a table: the full nesting of tile_table(), or [dtype][bucket]
    (each entry then covers every small-side and k class of its bucket)."""
    flat = list(_flatten(table))
    buf = (C.c_int * len(flat))(*flat)
    check(lib().ferray_set_tile_table(C.cast(buf, C.c_void_p), len(flat)), "set_tile_table")


def load_tile_table(path: str = TABLE_PATH) -> bool:
    """Loads the autotuner's table if present; returns whether it was applied."""
    if not os.path.exists(path):
        return False
    with open(path) as f:
        set_tile_table(json.load(f)["table"])
    return True


try:  # the autotuner's table, when one has been recorded for this device
    load_tile_table()
except (OSError, ValueError, KeyError):
    pass
