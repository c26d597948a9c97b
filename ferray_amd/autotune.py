"""Autotuner for the grouped kernel's tile-shape table.

For every dtype and size bucket the grouped kernel distinguishes, it times a
batch of square problems of that bucket's largest size with 64x64 and with
128x128 tiles (HIP events around grouped launches on one stream) and keeps
the faster shape.  The table and the measurements are written to
``ferray_amd/tile_table.json``, which ``ferray_amd.grouped`` loads on import.

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


def bucket_problem(dtype: int, bucket: int) -> int:
    """Side n of the square problem for a bucket (its real width is 2^(bucket+4))."""
    width = 1 << (bucket + 4)
    return width // 2 if dtype in (gg.DTYPE_C64, gg.DTYPE_C32) else width


def time_batch(dtype: int, n: int, count: int, cfg: int, reps: int) -> float:
    """ms per grouped launch of `count` n x n x n problems with tile shape cfg."""
    a = fa.DeviceBuffer(n * n * ELEM[dtype])
    b = fa.DeviceBuffer(n * n * ELEM[dtype])
    c = fa.DeviceBuffer(count * n * n * C_ELEM[dtype])
    scalars = n * n * (2 if dtype in (gg.DTYPE_C64, gg.DTYPE_C32) else 1)
    fa.fill(a, scalars, FILL[dtype], 1, -128, 127)
    fa.fill(b, scalars, FILL[dtype], 2, -128, 127)
    probs = [gg.problem(dtype, n, n, n, a.ptr, b.ptr, c.ptr + i * n * n * C_ELEM[dtype])
             for i in range(count)]
    saved = gg.tile_table()
    try:
        table = [row[:] for row in saved]
        table[dtype] = [cfg] * gg.N_BUCKETS
        gg.set_tile_table(table)
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


def tune(reps: int = 5, log=print) -> dict:
    table = [[0] * gg.N_BUCKETS for _ in range(gg.N_DTYPES)]
    measured = {}
    for dtype in NAMES:
        for bucket in range(gg.N_BUCKETS):
            n = bucket_problem(dtype, bucket)
            # enough problems for a few waves of 128x128 tiles on 256 CUs
            tiles128 = max(1, (n * (2 if dtype in (gg.DTYPE_C64, gg.DTYPE_C32) else 1) // 128)) * \
                max(1, n // 128)
            count = int(min(256, max(4, 2048 // tiles128)))
            t = [time_batch(dtype, n, count, cfg, reps) for cfg in (0, 1)]
            best = int(np.argmin(t))
            table[dtype][bucket] = best
            measured[f"{NAMES[dtype]}_{n}"] = {"count": count, "ms_64x64": round(t[0], 4),
                                               "ms_128x128": round(t[1], 4), "pick": best}
            log(f"{NAMES[dtype]} n={n} x{count}: 64x64 {t[0]:.3f} ms, 128x128 {t[1]:.3f} ms "
                f"-> {'128x128' if best else '64x64'}")
    return {"table": table, "measured": measured,
            "note": "entry [dtype][b]: 0 = 64x64 tiles, 1 = 128x128, for problems with "
                    "clamp(ceil(log2(max(m, n_real))), 4, 11) = b + 4"}


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
