"""Host-side pieces of the grouped GEMM API that need no device."""
from __future__ import annotations

import ctypes as C
import re

import pytest

from ferray_amd import grouped_gemm as gg
from tests.test_abi import ROOT

HEADER = ROOT + "/include/ferray_gemm.h"


def test_problem_struct_matches_header():
    text = open(HEADER).read()
    body = re.search(r"typedef struct ferray_gemm_problem \{(.*?)\}", text, re.S).group(1)
    fields = re.findall(r"(\w+)\s*;", body.replace(",", ";"))
    names = [f[0] for f in gg.GemmProblem._fields_]
    assert fields == names
    # int, int, 3 x size_t, 3 pointers, 4 doubles
    assert C.sizeof(gg.GemmProblem) == 8 + 3 * 8 + 3 * 8 + 4 * 8


def test_problem_builder():
    p = gg.problem(gg.DTYPE_C64, 3, 4, 5, 16, 32, 48, alpha=1 - 2j, beta=0.5j, accumulate=True)
    assert (p.dtype, p.m, p.n, p.k, p.a, p.b, p.c) == (2, 3, 4, 5, 16, 32, 48)
    assert (p.alpha_re, p.alpha_im, p.beta_re, p.beta_im, p.accumulate) == (1, -2, 0, 0.5, 1)


def test_tile_table_roundtrip_and_validation():
    saved = gg.tile_table()
    try:
        full = [[[[(d + b + s + q) % 2 for q in range(gg.N_KCLS)] for s in range(gg.N_SMALL)]
                 for b in range(gg.N_BUCKETS)] for d in range(gg.N_DTYPES)]
        gg.set_tile_table(full)
        assert gg.tile_table() == full
        # the per-bucket form covers every small-side and k class of its bucket
        t = [[(d + b) % 2 for b in range(gg.N_BUCKETS)] for d in range(gg.N_DTYPES)]
        gg.set_tile_table(t)
        assert gg.tile_table() == [[[[t[d][b]] * gg.N_KCLS] * gg.N_SMALL
                                    for b in range(gg.N_BUCKETS)] for d in range(gg.N_DTYPES)]
        with pytest.raises(Exception):
            gg.set_tile_table([[2] * gg.N_BUCKETS for _ in range(gg.N_DTYPES)])
        with pytest.raises(Exception):
            gg.set_tile_table([[0] * 3])
    finally:
        gg.set_tile_table(saved)
