"""Autotuner for the grouped kernel's tile-shape table.

For every dtype and every cell the grouped kernel distinguishes -- size
bucket of the longer side, class of the shorter side (< 64, < 128, >= 128)
and depth class (k < 128, < 512, >= 512, real elements) -- it times a batch
of problems shaped like the cell (shorter side m, longer side n, depth k)
with 64x64 and with 128x128 tiles (HIP events around grouped launches on one
stream) and keeps the faster shape.  Cells no problem can fall in (a shorter
side longer than the bucket) take the entry of the nearest valid class.  The
table and the measurements are written to ``ferray_amd/tile_table.json``,
which ``ferray_amd.grouped`` loads on import.

    python -m ferray_amd.autotune [--out PATH] [--reps N]
"""
from __future__ import annotations

import argparse
import json
import os
import sys

import numpy as np

from . import gemm as fa
from . import grouped as gg

NAMES = {gg.DTYPE_F64: "f64", gg.DTYPE_F32: "f32", gg.DTYPE_C64: "c64", gg.DTYPE_C32: "c32",
         gg.DTYPE_I8: "i8"}
ELEM = {gg.DTYPE_F64: 8, gg.DTYPE_F32: 4, gg.DTYPE_C64: 16, gg.DTYPE_C32: 8, gg.DTYPE_I8: 1}
C_ELEM = {gg.DTYPE_F64: 8, gg.DTYPE_F32: 4, gg.DTYPE_C64: 16, gg.DTYPE_C32: 8, gg.DTYPE_I8: 4}
FILL = {gg.DTYPE_F64: fa.FILL_F64_NORMAL, gg.DTYPE_F32: fa.FILL_F32_NORMAL,
        gg.DTYPE_C64: fa.FILL_F64_NORMAL, gg.DTYPE_C32: fa.FILL_F32_NORMAL,
        gg.DTYPE_I8: fa.FILL_I8}


SMALL_REP = (32, 96, None)     # shorter side per class (None: the bucket width)
K_REP = (64, 256, None)        # depth per class (None: max(512, bucket width))


def cell_shape(dtype: int, bucket: int, small: int, kcls: int):
    """(m, n, k) in the dtype's own elements for a table cell, or None if no
    problem falls in it."""
    width = 1 << (bucket + 4)
    short = SMALL_REP[small] if SMALL_REP[small] is not None else width
    if short > width or (small == 2 and width < 128) or (small == 1 and width < 64):
        return None
    k_real = K_REP[kcls] if K_REP[kcls] is not None else max(512, width)
    cplx = dtype in (gg.DTYPE_C64, gg.DTYPE_C32)
    n = width // 2 if cplx else width
    k = k_real // 2 if cplx else k_real
    return short, max(n, 1), max(k, 1)


def time_batch(dtype: int, m: int, n: int, k: int, count: int, cfg: int, reps: int) -> float:
    """ms per grouped launch of `count` m x n x k problems with tile shape cfg."""
    cplx = dtype in (gg.DTYPE_C64, gg.DTYPE_C32)
    a = fa.DeviceBuffer(m * k * ELEM[dtype])
    b = fa.DeviceBuffer(k * n * ELEM[dtype])
    c = fa.DeviceBuffer(count * m * n * C_ELEM[dtype])
    fa.fill(a, m * k * (2 if cplx else 1), FILL[dtype], 1, -128, 127)
    fa.fill(b, k * n * (2 if cplx else 1), FILL[dtype], 2, -128, 127)
    probs = [gg.problem(dtype, m, n, k, a.ptr, b.ptr, c.ptr + i * m * n * C_ELEM[dtype])
             for i in range(count)]
    saved = gg.tile_table()
    try:
        gg.set_tile_table([[cfg] * gg.N_BUCKETS for _ in range(gg.N_DTYPES)])
        plan = gg.Plan(count, 1 << 30)
        for _ in range(2):
            plan.launch(probs)
        plan.sync()
        e0, e1 = fa.Event(), fa.Event()
        e0.record()
        for _ in range(reps):
            plan.launch(probs)
        e1.record()
        plan.sync()
        fa.device_sync()
        ms = e0.elapsed_ms(e1) / reps
        plan.close()
    finally:
        gg.set_tile_table(saved)
        for buf in (a, b, c):
            buf.free()
    return ms


def default_cfg(bucket: int) -> int:
    """The library's built-in choice: 128x128 from a longer side of 256 up."""
    return 1 if bucket >= 4 else 0


def tune(reps: int = 5, log=print) -> dict:
    table = [[[[default_cfg(b)] * gg.N_KCLS for _ in range(gg.N_SMALL)]
              for b in range(gg.N_BUCKETS)] for _ in range(gg.N_DTYPES)]
    measured = {}
    for dtype in NAMES:
        cplx = dtype in (gg.DTYPE_C64, gg.DTYPE_C32)
        for bucket in range(gg.N_BUCKETS):
            for small in range(gg.N_SMALL):
                for kcls in range(gg.N_KCLS):
                    shape = cell_shape(dtype, bucket, small, kcls)
                    if shape is None:   # no problem falls here: nearest valid class
                        table[dtype][bucket][small][kcls] = table[dtype][bucket][small - 1][kcls]
                        continue
                    m, n, k = shape
                    n_real = 2 * n if cplx else n
                    tiles128 = -(-m // 128) * -(-n_real // 128)
                    count = int(min(256, max(4, 2048 // tiles128)))
                    t = [time_batch(dtype, m, n, k, count, cfg, reps) for cfg in (0, 1)]
                    best = int(np.argmin(t))
                    table[dtype][bucket][small][kcls] = best
                    key = f"{NAMES[dtype]}_{m}x{n}x{k}"
                    measured[key] = {"count": count, "ms_64x64": round(t[0], 4),
                                     "ms_128x128": round(t[1], 4), "pick": best,
                                     "default": default_cfg(bucket)}
                    log(f"{key} x{count}: 64x64 {t[0]:.3f} ms, 128x128 {t[1]:.3f} ms -> "
                        f"{'128x128' if best else '64x64'}"
                        f"{'' if best == default_cfg(bucket) else ' (differs from default)'}")
    return {"table": table, "measured": measured,
            "note": "entry [dtype][b][s][q]: 0 = 64x64 tiles, 1 = 128x128, for problems with "
                    "clamp(ceil(log2(max(m, n_real))), 4, 11) = b + 4, min(m, n_real) in "
                    "[0,64)/[64,128)/[128,..) for s = 0/1/2 and k_real in [0,128)/[128,512)/"
                    "[512,..) for q = 0/1/2"}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=gg.TABLE_PATH)
    ap.add_argument("--reps", type=int, default=5)
    args = ap.parse_args(argv)
    if not fa.gpu_supports():
        print("no usable gfx950 device", file=sys.stderr)
        return 1
    result = tune(args.reps, log=lambda s: print(s, flush=True))
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(result, f, indent=1)
    print(f"wrote {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
