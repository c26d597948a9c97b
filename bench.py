"""Benchmark: native FP64 MFMA DGEMM at M=N=K=8192 on one MI355X.

A step is one C = A.B call (f64, row-major, entries normal(0,1) from the
seeded device generator) with A, B and C resident in HBM.  W untimed warm-up
steps, then K steps timed with HIP events on the stream the kernels run on,
bracketed by device synchronisation.  One JSON line is printed.

Besides the headline the line carries:
  roofline      - the DGEMM kernel's achieved FLOP/s against the FP64 matrix
                  peak (78.6 TFLOP/s), plus HBM traffic per launch when the
                  rocprofv3 counter summary for it is present under profiles/;
  cpu_baseline  - the oracle's restatement of ferray's CPU GEMM (Goto 5-loop,
                  AVX2 4x8 kernel, min(nproc, 8) workers) timed on this host
                  at 8192^3 f64 (median of 3 after one warm-up), plus the
                  512^3 config timed the reference's way (median of 30 after
                  5 warm-ups) beside ferray's published figure;
  error         - normwise error of the 8192 result against a compensated
                  (double-double) dot product on sampled rows and columns;
  split_precision - the Ozaki-scheme DGEMM on INT8 MFMA at the same size, per
                  modulus count: time, speed-up over the native DGEMM above and
                  sampled normwise error, for normal(0,1) and u*2^e inputs;
  split_sgemm   - SGEMM at the same size on BF16 matrix cores by operand
                  splitting (3 and 6 piece products) against native FP32
                  MFMA: time, speed-up and sampled normwise error;
  sweep         - the other dtypes of the GEMM family at their config sizes
                  (f32 8192, c64/c32 4096, i8 8192, i16 4096, bf16/f16
                  8192, OCP E4M3 fp8 8192, quantized_matmul u8 x i8 8192),
                  one timed number each (--no-sweep skips them);
  consumers     - the product layer's GEMM consumers on the device: TRSM at
                  8192 x 8192, 256 batched 256^3 products in one grouped
                  launch against 256 launches, matrix_power (--no-consumers
                  skips them);
  serving       - 20,000-request Poisson trace (M, N, K log-uniform in
                  [16, 2048], f64/f32/c64/i8 mix, operands resident in HBM)
                  through the serving loop with continuous batching and
                  through one launch per request: served rate, TFLOP/s, p50
                  and p99 latency per offered rate, and the best rate of each
                  at p99 <= 2 ms (--no-serve skips it).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("GEMM TFLOP/s at 8192 (f64 native vs split-precision, % peak); "
          "serving p99 ms")
PEAK = {  # MI355X published dense peaks, TFLOP/s (TOP/s for int8)
    "f64": 78.6, "f32": 157.3, "bf16": 2500.0, "f16": 2500.0, "i8": 5000.0, "fp8": 5000.0,
}
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def time_steps(fa, launch, steps: int, warmup: int) -> float:
    """ms per step: events on the library's stream around `steps` calls."""
    for _ in range(warmup):
        launch()
    fa.device_sync()
    e0, e1 = fa.Event(), fa.Event()
    e0.record()
    for _ in range(steps):
        launch()
    e1.record()
    ms = e0.elapsed_ms(e1) / steps
    fa.device_sync()
    return ms


def load_traffic(name: str):
    """HBM bytes per launch from profiles/<name>_traffic.json, if recorded."""
    path = os.path.join(ROOT, "profiles", f"{name}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def dgemm_error(fa, dA, dB, dC, n: int, samples: int = 16):
    import oracle
    a = dA.download(np.float64, n * n)
    b = dB.download(np.float64, n * n)
    rng = np.random.default_rng(123)
    rows = np.sort(rng.choice(n, samples, replace=False))
    cols = np.sort(rng.choice(n, samples, replace=False))
    hi, lo = oracle.dd_gemm_f64_sampled(n, n, n, a, b, rows, cols)
    c_rows = np.stack([dC.download(np.float64, n, offset_bytes=int(r) * n * 8) for r in rows])
    got = c_rows[:, cols]
    d = (got - hi) - lo
    na, nb = np.linalg.norm(a), np.linalg.norm(b)
    # scale the sampled error to the full matrix (mean square per element)
    frac = (n * n) / (samples * samples)
    err_f = np.sqrt(np.sum(d * d) * frac)
    c_f = np.sqrt(np.sum(hi * hi) * frac)
    return {"normwise_ab": float(err_f / (na * nb)), "rel_c": float(err_f / c_f),
            "max_abs": float(np.max(np.abs(d))),
            "sample": f"{samples} rows x {samples} cols vs compensated dot products"}


def split_precision(fa, dA, dB, n: int, steps: int, native_ms: float, dist: str,
                    moduli=(12, 13, 14, 15, 16), samples: int = 16):
    """Ozaki-scheme DGEMM on INT8 MFMA: time and sampled error per modulus count."""
    import oracle
    a = dA.download(np.float64, n * n)
    b = dB.download(np.float64, n * n)
    rng = np.random.default_rng(321)
    rows = np.sort(rng.choice(n, samples, replace=False))
    cols = np.sort(rng.choice(n, samples, replace=False))
    hi, lo = oracle.dd_gemm_f64_sampled(n, n, n, a, b, rows, cols)
    na, nb = np.linalg.norm(a), np.linalg.norm(b)
    frac = (n * n) / (samples * samples)
    c_f = np.sqrt(np.sum(hi * hi) * frac)
    dC = fa.DeviceBuffer(n * n * 8)
    L = fa.lib()
    out = {}
    for nm in moduli:
        def step():
            st = L.ferray_dgemm_ozaki_dev(n, n, n, 1.0, dA.ptr, dB.ptr, 0.0, dC.ptr, nm, None)
            if st != 0:
                raise RuntimeError(f"ferray_dgemm_ozaki_dev status {st}")
        ms = time_steps(fa, step, max(3, steps // 2), 1)
        got = np.stack([dC.download(np.float64, n, offset_bytes=int(r) * n * 8) for r in rows])[:, cols]
        d = (got - hi) - lo
        err_f = np.sqrt(np.sum(d * d) * frac)
        out[str(nm)] = {"ms": round(ms, 4), "tflops_equiv": round(2.0 * n ** 3 / (ms * 1e-3) / 1e12, 2),
                        "speedup_vs_native": round(native_ms / ms, 3),
                        "normwise_ab": float(err_f / (na * nb)), "rel_c": float(err_f / c_f)}
        log(f"ozaki[{dist}] moduli={nm}: {ms:.3f} ms ({native_ms / ms:.2f}x native), "
            f"normwise {err_f / (na * nb):.2e}")
    dC.free()
    return out


def split_sgemm(fa, n: int, steps: int, warmup: int, samples: int = 16):
    """SGEMM at n^3 (normal(0,1) f32): native FP32 MFMA against the bf16
    operand-split forms (terms 3 and 6); time, speed-up and normwise error
    of sampled rows against an f64 product of the same f32 operands."""
    dA, dB, dC = fa.DeviceBuffer(n * n * 4), fa.DeviceBuffer(n * n * 4), fa.DeviceBuffer(n * n * 4)
    fa.fill(dA, n * n, fa.FILL_F32_NORMAL, 21)
    fa.fill(dB, n * n, fa.FILL_F32_NORMAL, 22)
    L = fa.lib()
    a = dA.download(np.float32, n * n).reshape(n, n)
    b = dB.download(np.float32, n * n).reshape(n, n).astype(np.float64)
    rows = np.sort(np.random.default_rng(5).choice(n, samples, replace=False))
    ref = a[rows].astype(np.float64) @ b
    scale = np.linalg.norm(a[rows].astype(np.float64)) * np.linalg.norm(b)
    runs = {"native_f32": lambda: L.ferray_gemm_f32_dev(n, n, n, 1.0, dA.ptr, dB.ptr, 0.0, dC.ptr, None),
            "bf16x3": lambda: L.ferray_sgemm_bf16x3_dev(n, n, n, 1.0, dA.ptr, dB.ptr, 0.0, dC.ptr, 3, None),
            "bf16x6": lambda: L.ferray_sgemm_bf16x3_dev(n, n, n, 1.0, dA.ptr, dB.ptr, 0.0, dC.ptr, 6, None)}
    out = {}
    for name, fn in runs.items():
        if fn() != 0:
            out[name] = {"error": "launch failed"}
            continue
        ms = time_steps(fa, fn, max(3, steps // 2), max(1, warmup))
        got = np.stack([dC.download(np.float32, n, offset_bytes=int(r) * n * 4) for r in rows])
        err = float(np.linalg.norm(got.astype(np.float64) - ref) / scale)
        out[name] = {"ms": round(ms, 4), "tflops_equiv": round(2.0 * n ** 3 / (ms * 1e-3) / 1e12, 2),
                     "normwise_ab_sampled_rows": err}
        log(f"sgemm {name} {n}: {ms:.3f} ms, normwise {err:.2e}")
    for name in ("bf16x3", "bf16x6"):
        if "ms" in out.get(name, {}) and "ms" in out.get("native_f32", {}):
            out[name]["speedup_vs_native"] = round(out["native_f32"]["ms"] / out[name]["ms"], 3)
    for buf in (dA, dB, dC):
        buf.free()
    return out


def consumers(fa, steps: int, warmup: int):
    """GEMM consumers of the product layer, device resident:
      trsm      - L X = B, L lower 8192 x 8192, B 8192 x 8192 (f64): m^2 n flop
      batched   - 256 independent 256^3 f64 products: one grouped launch
                  (ferray_matmul_batched_dev) against 256 single launches
      matrix_power - A^13 of a 4096 x 4096 f64 matrix (5 squarings/products)"""
    L = fa.lib()
    out = {}
    m = n = 8192
    dA, dB = fa.DeviceBuffer(m * m * 8), fa.DeviceBuffer(m * n * 8)
    fa.fill(dA, m * m, fa.FILL_F64_UNIFORM, 31)
    fa.fill(dB, m * n, fa.FILL_F64_NORMAL, 32)
    # A well-conditioned lower triangle: the kernel only reads the lower part;
    # a large diagonal keeps the solve stable for the timing.
    a_h = dA.download(np.float64, m * m).reshape(m, m) / np.sqrt(m)
    np.fill_diagonal(a_h, 4.0)
    dA.upload(np.ascontiguousarray(a_h))
    del a_h
    b0 = fa.DeviceBuffer(m * n * 8)
    L.ferray_dev_copy(b0.ptr, dB.ptr, m * n * 8, None)

    def trsm_once():
        """One solve from a fresh B, timed alone (the reset copy is outside
        the events)."""
        L.ferray_dev_copy(dB.ptr, b0.ptr, m * n * 8, None)
        e0, e1 = fa.Event(), fa.Event()
        e0.record()
        st = L.ferray_trsm_dev(0, 0, m, n, 1.0, dA.ptr, m, dB.ptr, n, False, None)
        e1.record()
        if st != 0:
            raise RuntimeError(f"ferray_trsm_dev status {st}")
        fa.device_sync()
        return e0.elapsed_ms(e1)

    for _ in range(3):
        trsm_once()
    runs = [trsm_once() for _ in range(7)]
    t = float(np.median(runs))
    out["trsm_lower_f64_8192x8192"] = {"ms": round(t, 3), "tflops": round(m * m * n / (t * 1e-3) / 1e12, 2),
                                       "min_ms": round(min(runs), 3), "max_ms": round(max(runs), 3),
                                       "note": "m^2 n flop; median of 7 solves after 3 warm-ups, each "
                                               "timed alone with events on the library stream (the B "
                                               "reset copy runs before the first event)"}
    log(f"trsm f64 8192x8192: {t:.2f} ms, {m * m * n / (t * 1e-3) / 1e12:.1f} TFLOP/s")
    for buf in (dA, dB, b0):
        buf.free()

    batch, sz = 256, 256
    e = sz * sz * 8
    dA, dB, dC = fa.DeviceBuffer(batch * e), fa.DeviceBuffer(batch * e), fa.DeviceBuffer(batch * e)
    fa.fill(dA, batch * sz * sz, fa.FILL_F64_NORMAL, 33)
    fa.fill(dB, batch * sz * sz, fa.FILL_F64_NORMAL, 34)

    def grouped_step():
        st = L.ferray_matmul_batched_dev(0, batch, batch, batch, sz, sz, sz, dA.ptr, dB.ptr,
                                         dC.ptr, None)
        if st != 0:
            raise RuntimeError(f"ferray_matmul_batched_dev status {st}")

    def loop_step():
        for i in range(batch):
            L.ferray_gemm_f64_dev(sz, sz, sz, 1.0, dA.ptr + i * e, dB.ptr + i * e, 0.0,
                                  dC.ptr + i * e, None)

    flops = 2.0 * batch * sz ** 3
    ms_g = time_steps(fa, grouped_step, max(3, steps), warmup)
    ms_l = time_steps(fa, loop_step, max(3, steps), warmup)
    out["batched_256x256cubed_f64"] = {
        "grouped_ms": round(ms_g, 3), "grouped_tflops": round(flops / (ms_g * 1e-3) / 1e12, 2),
        "per_matrix_launches_ms": round(ms_l, 3),
        "per_matrix_tflops": round(flops / (ms_l * 1e-3) / 1e12, 2),
        "speedup": round(ms_l / ms_g, 2),
        "note": "the 256 single launches are issued from Python (ctypes), so they include "
                "host call overhead, as a caller looping over matrices would pay"}
    log(f"batched 256 x 256^3 f64: grouped {ms_g:.3f} ms, 256 launches {ms_l:.3f} ms")
    for buf in (dA, dB, dC):
        buf.free()

    from ferray_amd import products
    sz = 4096
    a = np.eye(sz) * 0.5 + np.random.default_rng(35).standard_normal((sz, sz)) * (0.1 / np.sqrt(sz))
    t0 = time.perf_counter()
    products.matrix_power(a, 13)
    w = time.perf_counter() - t0
    out["matrix_power_4096_p13_f64"] = {"wall_ms": round(w * 1e3, 2), "gemms": 6,
                                        "note": "includes the 128 MiB upload and download"}
    log(f"matrix_power 4096^13: {w * 1e3:.1f} ms wall")
    return out


def _median_time(fn, reps: int, warm: int) -> float:
    for _ in range(warm):
        fn()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    return float(np.median(times))


def cpu_baseline(n: int = 8192, reps: int = 3):
    """The oracle's restatement of ferray's CPU DGEMM on this host at the
    headline size (median of `reps` after one warm-up), plus config 1
    (512^3, the reference's LCG inputs) timed the reference's way: median of
    30 after 5 warm-ups, beside ferray's published 1768 us on 8 threads
    (outer.rs:136-141)."""
    import oracle
    a = oracle.uniform_f64(n * n, 1)
    b = oracle.uniform_f64(n * n, 2)
    c = np.zeros(n * n)
    workers = oracle.gemm_f64_goto(n, n, n, 1.0, a, b, 0.0, c)
    t = _median_time(lambda: oracle.gemm_f64_goto(n, n, n, 1.0, a, b, 0.0, c), reps, 0)
    del a, b, c
    m = 512
    a1, b1 = oracle.lcg_matrix(m, 1), oracle.lcg_matrix(m, 2)
    c1, ref = np.zeros(m * m), np.zeros(m * m)
    t1 = _median_time(lambda: oracle.gemm_f64_goto(m, m, m, 1.0, a1, b1, 0.0, c1), 30, 5)
    oracle.gemm_f64(m, m, m, 1.0, a1, b1, 0.0, ref)
    return {"value": round(2.0 * n ** 3 / t / 1e12, 4), "unit": "TFLOP/s", "cores": int(workers),
            "kind": "port",
            "sample": (f"ferray CPU DGEMM restated in C (Goto 5-loop MC/KC/NC=64/256/512, AVX2 4x8, "
                       f"pick_grid, min(nproc,8) workers); f64 {n}^3 uniform[-1,1), median of "
                       f"{reps} after 1 warm-up, {t * 1e3:.0f} ms per call; avx2="
                       f"{oracle.cpu_has_avx2_fma()}"),
            "config1_512": {
                "us": round(t1 * 1e6, 1), "gflops": round(2.0 * m ** 3 / t1 / 1e9, 1),
                "method": "LCG seeds 1, 2 (examples/bench_matmul.rs:21-34); median of 30 after 5 "
                          "warm-ups",
                "bit_identical_to_block_formula": bool(np.array_equal(c1, ref)),
                "ferray_published_us_8_threads": 1768.0}}


def offload_crossover(fa):
    """Host-pointer GPU calls (upload, product, download, synchronise) against
    the CPU restatement of ferray's GEMM on this host, square products of
    size n: the smallest n from which the GPU call is faster at that and
    every larger size, per dtype.  The CPU side follows ferray's dispatch:
    one thread below max-dim 256 (faer Par::Seq, products/mod.rs:499-576),
    min(nproc, 8) workers from 256 (the hand-tuned pool, outer.rs:142-155).
    Complex products are timed on the CPU as the real Goto GEMM of the same
    flop count (m x 2n x 2k): the restatement's complex kernel is the plain
    block formula, not ferray's AVX2 one."""
    import oracle
    sizes = [8, 16, 32, 48, 64, 96, 128, 192, 256, 384, 512]
    out = {}
    for dt in ("f64", "f32", "c64", "c32"):
        real = np.float64 if dt in ("f64", "c64") else np.float32
        cplx = dt[0] == "c"
        rows, cross = [], None
        for n in sizes:
            r = np.random.default_rng(n)
            if cplx:
                a = r.standard_normal(2 * n * n).astype(real)
                b = r.standard_normal(2 * n * n).astype(real)
                c = np.zeros(2 * n * n, real)
                gpu = (fa.gemm_c64 if dt == "c64" else fa.gemm_c32)
                g = lambda: gpu(n, n, n, 1.0, 0.0, a, b, 0.0, 0.0, c)  # noqa: E731
                ca = r.standard_normal(4 * n * n).astype(real)
                cc = np.zeros(4 * n * n, real)
                cpu_f = oracle.gemm_f64_goto if real == np.float64 else oracle.gemm_f32_goto
                h = lambda: cpu_f(n, 2 * n, 2 * n, real(1.0), ca, ca, real(0.0), cc,  # noqa: E731
                                  1 if n < 256 else 0)
            else:
                a = r.standard_normal(n * n).astype(real)
                b = r.standard_normal(n * n).astype(real)
                c = np.zeros(n * n, real)
                gpu = fa.gemm_f64 if dt == "f64" else fa.gemm_f32
                g = lambda: gpu(n, n, n, real(1.0), a, b, real(0.0), c)  # noqa: E731
                cpu_f = oracle.gemm_f64_goto if dt == "f64" else oracle.gemm_f32_goto
                h = lambda: cpu_f(n, n, n, real(1.0), a, b, real(0.0), c,  # noqa: E731
                                  1 if n < 256 else 0)
            tg = _median_time(g, 15, 3)
            tc = _median_time(h, 15, 3)
            rows.append({"n": n, "gpu_us": round(tg * 1e6, 1), "cpu_us": round(tc * 1e6, 1)})
        for i in range(len(rows)):
            if all(r["gpu_us"] < r["cpu_us"] for r in rows[i:]):
                cross = rows[i]["n"]
                break
        out[dt] = {"crossover_n": cross, "points": rows}
        log(f"offload crossover {dt}: n = {cross}")
    return out


def run_sweep(fa, steps: int, warmup: int):
    out = {}
    cases = [("sgemm_f32_8192", "f32", 8192), ("zgemm_c64_4096", "c64", 4096),
             ("cgemm_c32_4096", "c32", 4096), ("igemm_i8_8192", "i8", 8192),
             ("igemm_i16_4096", "i16", 4096), ("hgemm_bf16_8192", "bf16", 8192),
             ("hgemm_f16_8192", "f16", 8192), ("fp8gemm_e4m3_8192", "fp8", 8192),
             ("quantized_matmul_8192", "quant", 8192)]
    for name, dt, n in cases:
        esz = {"f32": 4, "c64": 16, "c32": 8, "i8": 1, "i16": 2, "bf16": 2, "f16": 2, "fp8": 1,
               "quant": 1}[dt]
        csz = {"f32": 4, "c64": 16, "c32": 8, "i8": 4, "i16": 4, "bf16": 4, "f16": 4, "fp8": 4,
               "quant": 4}[dt]
        dA, dB, dC = fa.DeviceBuffer(n * n * esz), fa.DeviceBuffer(n * n * esz), \
            fa.DeviceBuffer(n * n * csz)
        if dt in ("f32", "c32"):
            cnt = n * n * (2 if dt == "c32" else 1)
            fa.fill(dA, cnt, fa.FILL_F32_NORMAL, 11); fa.fill(dB, cnt, fa.FILL_F32_NORMAL, 12)
        elif dt == "c64":
            fa.fill(dA, 2 * n * n, fa.FILL_F64_NORMAL, 11); fa.fill(dB, 2 * n * n, fa.FILL_F64_NORMAL, 12)
        elif dt in ("i8", "quant"):
            fa.fill(dA, n * n, fa.FILL_I8, 11, -128, 127); fa.fill(dB, n * n, fa.FILL_I8, 12, -128, 127)
        elif dt == "fp8":  # finite positive E4M3 codes (exponent field <= 14)
            fa.fill(dA, n * n, fa.FILL_I8, 11, 0, 0x77); fa.fill(dB, n * n, fa.FILL_I8, 12, 0, 0x77)
        elif dt == "i16":
            fa.fill(dA, n * n, fa.FILL_I16, 11, -32768, 32767)
            fa.fill(dB, n * n, fa.FILL_I16, 12, -32768, 32767)
        else:  # bf16 / f16: random bit patterns with a bounded exponent
            fa.fill(dA, n * n, fa.FILL_I16, 11, 0x3c00, 0x3fff)
            fa.fill(dB, n * n, fa.FILL_I16, 12, 0xbc00, 0xbfff)
        L = fa.lib()
        fns = {
            "f32": lambda: L.ferray_gemm_f32_dev(n, n, n, 1.0, dA.ptr, dB.ptr, 0.0, dC.ptr, None),
            "c64": lambda: L.ferray_gemm_c64_dev(n, n, n, 1.0, 0.0, dA.ptr, dB.ptr, 0.0, 0.0, dC.ptr, None),
            "c32": lambda: L.ferray_gemm_c32_dev(n, n, n, 1.0, 0.0, dA.ptr, dB.ptr, 0.0, 0.0, dC.ptr, None),
            "i8": lambda: L.ferray_gemm_i8_signed_dev(n, n, n, dA.ptr, dB.ptr, dC.ptr, False, None),
            "i16": lambda: L.ferray_gemm_i16_dev(n, n, n, dA.ptr, dB.ptr, dC.ptr, False, None),
            "bf16": lambda: L.ferray_gemm_bf16_f32_dev(n, n, n, dA.ptr, dB.ptr, dC.ptr, False, None),
            "f16": lambda: L.ferray_gemm_f16_f32_dev(n, n, n, dA.ptr, dB.ptr, dC.ptr, False, None),
            "fp8": lambda: L.ferray_gemm_e4m3_f32_dev(n, n, n, dA.ptr, dB.ptr, dC.ptr, False, None),
            "quant": lambda: L.ferray_quantized_matmul_dev(n, n, n, dA.ptr, 128, 0.02, dB.ptr, -3,
                                                           0.05, dC.ptr, False, None),
        }
        fn = fns[dt]
        st = fn()
        if st != 0:
            out[name] = {"error": f"status {st}"}
            continue
        # at least ~40 ms of timed work per product, so that sub-millisecond
        # calls are averaged over many launches
        est = time_steps(fa, fn, 2, max(1, warmup))
        reps = max(max(3, steps // 2), min(200, int(40.0 / max(est, 1e-3))))
        ms = time_steps(fa, fn, reps, 1)
        flops = (8.0 if dt in ("c64", "c32") else 2.0) * n ** 3
        rate = flops / (ms * 1e-3) / 1e12
        peak_key = {"c64": "f64", "c32": "f32", "i16": "i8", "quant": "i8"}.get(dt, dt)
        peak = PEAK[peak_key] / (4.0 if dt == "i16" else 1.0)  # i16 = 4 int8 products
        out[name] = {"ms": round(ms, 4), "rate": round(rate, 2),
                     "unit": "TOP/s" if dt in ("i8", "i16", "quant") else "TFLOP/s",
                     "frac_of_peak": round(rate / peak, 4), "calls_timed": reps}
        log(f"{name}: {ms:.3f} ms  {rate:.1f}  ({rate / peak * 100:.1f}% of peak)")
        for buf in (dA, dB, dC):
            buf.free()
    return out


SERVE_P99_TARGET_MS = 2.0


def serving_check(fa, sv, pools, count: int = 500):
    """Served results against their single-problem entry points, bit for bit,
    on `count` requests of the serving trace kept by a keep_results server."""
    L = fa.lib()
    reqs, arrival, _ = sv.make_trace(pools, count, rate_per_s=1e5, seed=11)
    cfg = sv.config(hbm_bytes=4 << 30, max_resident=4096, max_batch=256, lanes=4,
                    keep_results=True)
    with sv.Server(cfg) as srv:
        tickets = [srv.submit(q) for q in reqs]
        got = [srv.result(t, q.dtype, q.m, q.n) for t, q in zip(tickets, reqs)]
    out = fa.DeviceBuffer(max(q.m * q.n * 16 for q in reqs))
    same = 0
    for q, c in zip(reqs, got):
        if q.dtype == sv.DTYPE_F64:
            st = L.ferray_gemm_f64_dev(q.m, q.n, q.k, q.alpha_re, q.a, q.b, 0.0, out.ptr, None)
        elif q.dtype == sv.DTYPE_F32:
            st = L.ferray_gemm_f32_dev(q.m, q.n, q.k, q.alpha_re, q.a, q.b, 0.0, out.ptr, None)
        elif q.dtype == sv.DTYPE_C64:
            st = L.ferray_gemm_c64_dev(q.m, q.n, q.k, q.alpha_re, q.alpha_im, q.a, q.b, 0.0,
                                       0.0, out.ptr, None)
        else:
            st = L.ferray_gemm_i8_signed_dev(q.m, q.n, q.k, q.a, q.b, out.ptr, False, None)
        fa.sync()
        if st == 0 and out.download(sv.c_dtype(q.dtype), q.m * q.n).tobytes() == c.tobytes():
            same += 1
    out.free()
    log(f"serving result check: {same} / {count} bit-identical to the single-problem entries")
    return {"requests": count, "bit_identical_to_single_entry": same, "ok": same == count}


def serving(fa, count: int = 20000):
    """The serving workload: `count` requests, Poisson arrivals, M, N, K
    log-uniform in [16, 2048], 50% f64 / 30% f32 / 10% c64 / 10% i8->i32,
    operands resident in HBM (device pools of seeded random data).  Each
    arrival rate is replayed through a fresh server, grouped (continuous
    batching onto 4 launch lanes) and one launch per request."""
    from ferray_amd import serve as sv
    elems = 1 << 29
    pools = sv.OperandPools(fa, elems=elems, seed=7)
    out = {"requests": count, "dims": "log-uniform [16, 2048]",
           "mix": "f64 0.5, f32 0.3, c64 0.1, i8 0.1", "p99_target_ms": SERVE_P99_TARGET_MS,
           "lanes": 4, "max_batch": 256,
           "operand_pools_gib": round(elems * (8 + 4 + 1) / 2 ** 30, 2),
           "operands": "each request reads A and B at random 256-byte-aligned offsets of "
                       "device pools (f64 4 GiB, f32 2 GiB, i8 0.5 GiB) far larger than the "
                       "256 MB last-level cache, so operands stream from HBM"}
    out["result_check"] = serving_check(fa, sv, pools)
    rates = {"grouped": [2e4, 5e4, 7.5e4, 1e5, 1.25e5, 1.5e5, 2e5],
             "per_request": [1e4, 2e4, 3e4, 3.5e4, 4e4, 5e4, 6e4]}
    best = {}
    for name, mode in (("grouped", sv.SERVE_GROUPED), ("per_request", sv.SERVE_PER_REQUEST)):
        rows = []
        for rate in rates[name]:
            reqs, arrival, flops = sv.make_trace(pools, count, rate_per_s=rate, seed=3)
            cfg = sv.config(hbm_bytes=16 << 30, max_resident=4096, max_batch=256, lanes=4,
                            mode=mode)
            lat, wall, st = sv.run_trace(cfg, reqs, arrival)
            row = {"offered_per_s": rate, "served_per_s": round(count / wall, 1),
                   "tflops": round(float(flops.sum()) / wall / 1e12, 3),
                   "p50_ms": round(float(np.percentile(lat, 50)), 4),
                   "p99_ms": round(float(np.percentile(lat, 99)), 4),
                   "launches": int(st.launches)}
            rows.append(row)
            log(f"serving {name} @ {rate:.0f}/s: {row['served_per_s']:.0f}/s, "
                f"p50 {row['p50_ms']:.3f} ms, p99 {row['p99_ms']:.3f} ms, {row['launches']} launches")
        out[name] = rows
        ok = [r for r in rows if r["p99_ms"] <= SERVE_P99_TARGET_MS]
        best[name] = max(ok, key=lambda r: r["served_per_s"]) if ok else None
    pools.free()
    out["best_at_p99_target"] = best
    if best["grouped"] and best["per_request"]:
        out["throughput_ratio_at_p99_target"] = round(
            best["grouped"]["served_per_s"] / best["per_request"]["served_per_s"], 3)
    return out


def vendor_products(n: int = 8192):
    """The same n^3 products through the vendor libraries torch calls on ROCm
    (rocBLAS / hipBLASLt): a comparison figure only, nothing of the library
    runs through torch.  Measured in a child process, since torch's own HIP
    runtime does not find the device in a process where the library's
    runtime already runs."""
    import subprocess
    res = subprocess.run([sys.executable, os.path.abspath(__file__), "--vendor-child", str(n)],
                         capture_output=True, text=True, timeout=600)
    sys.stderr.write(res.stderr)
    if res.returncode != 0:
        return {"error": f"child exited with {res.returncode}"}
    return json.loads(res.stdout.strip().splitlines()[-1])


def vendor_child(n: int, steps: int = 10, warmup: int = 3):
    """Times the vendor products in this (fresh) process; see vendor_products."""
    import torch
    torch.backends.cuda.matmul.allow_tf32 = False
    dev = torch.device("cuda:0")

    def timed(fn):
        # the sweep's rule: at least ~40 ms of timed work per product
        def run(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                fn()
            e1.record()
            torch.cuda.synchronize()
            return e0.elapsed_time(e1) / reps
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        est = run(2)
        return run(max(steps // 2, min(200, int(40.0 / max(est, 1e-3)))))

    out = {"library": "torch " + torch.__version__ + " (rocBLAS / hipBLASLt)"}
    g = torch.Generator(device=dev).manual_seed(7)
    work = 2.0 * n ** 3
    for name, dt in (("f64", torch.float64), ("f32", torch.float32), ("bf16", torch.bfloat16),
                     ("f16", torch.float16)):
        a = torch.randn(n, n, device=dev, dtype=dt, generator=g)
        b = torch.randn(n, n, device=dev, dtype=dt, generator=g)
        ms = timed(lambda: torch.matmul(a, b))
        out[name] = {"ms": round(ms, 4), "rate": round(work / (ms * 1e-3) / 1e12, 2),
                     "unit": "TFLOP/s", "call": "torch.matmul, row-major A and B"}
        log(f"vendor {name}: {ms:.3f} ms")
        del a, b
    a = torch.randint(-128, 128, (n, n), device=dev, dtype=torch.int8, generator=g)
    b = torch.randint(-128, 128, (n, n), device=dev, dtype=torch.int8, generator=g)
    bt = b.t().contiguous().t()  # the layout the vendor int8 kernels prefer
    for key, bb, lay in (("i8", b, "row-major B"), ("i8_bcol", bt, "column-major B")):
        ms = timed(lambda: torch._int_mm(a, bb))
        out[key] = {"ms": round(ms, 4), "rate": round(work / (ms * 1e-3) / 1e12, 2),
                    "unit": "TOP/s", "call": f"torch._int_mm, {lay}"}
        log(f"vendor {key}: {ms:.3f} ms")
    del a, b, bt
    a = torch.randn(n, n, device=dev, generator=g).to(torch.float8_e4m3fn)
    b = torch.randn(n, n, device=dev, generator=g).to(torch.float8_e4m3fn).t().contiguous().t()
    one = torch.tensor(1.0, device=dev)
    ms = timed(lambda: torch._scaled_mm(a, b, scale_a=one, scale_b=one, out_dtype=torch.float32))
    out["e4m3"] = {"ms": round(ms, 4), "rate": round(work / (ms * 1e-3) / 1e12, 2),
                   "unit": "TFLOP/s", "call": "torch._scaled_mm, column-major B, f32 out"}
    log(f"vendor e4m3: {ms:.3f} ms")
    del a, b
    torch.cuda.empty_cache()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=8192)
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    ap.add_argument("--no-split", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-host", action="store_true")
    ap.add_argument("--no-consumers", action="store_true")
    ap.add_argument("--crossover", action="store_true",
                    help="also time host-pointer products against the CPU restatement to find "
                         "the size from which offloading pays, per dtype")
    ap.add_argument("--vendor", action="store_true",
                    help="also time the same 8192^3 products through the vendor libraries "
                         "(rocBLAS / hipBLASLt via torch), for comparison")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run one 1 GiB device-to-device copy, the reference point "
                         "that scales rocprofv3's FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--vendor-child", type=int, default=0, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.vendor_child:
        print(json.dumps(vendor_child(args.vendor_child)))
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the GEMM path runs on a single MI355X); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        return 2

    import ferray_amd as fa
    if not fa.gpu_supports():
        print("no usable gfx950 device", file=sys.stderr)
        return 1
    n = args.size
    L = fa.lib()
    if args.copy_calibration:
        gib = 1 << 30
        src, dst = fa.DeviceBuffer(gib), fa.DeviceBuffer(gib)
        fa.fill(src, gib // 8, fa.FILL_F64_UNIFORM, 5)
        fa.device_sync()
        if L.ferray_dev_copy(dst.ptr, src.ptr, gib, None) != 0:
            raise RuntimeError("ferray_dev_copy failed")
        fa.device_sync()
        src.free()
        dst.free()
        log("copy calibration: 1 GiB read + 1 GiB written by copy_kernel")
    log(f"DGEMM {n}^3 f64: allocating and filling operands")
    dA, dB, dC = fa.DeviceBuffer(n * n * 8), fa.DeviceBuffer(n * n * 8), fa.DeviceBuffer(n * n * 8)
    fa.fill(dA, n * n, fa.FILL_F64_NORMAL, 1)
    fa.fill(dB, n * n, fa.FILL_F64_NORMAL, 2)

    def step():
        st = L.ferray_gemm_f64_dev(n, n, n, 1.0, dA.ptr, dB.ptr, 0.0, dC.ptr, None)
        if st != 0:
            raise RuntimeError(f"ferray_gemm_f64_dev status {st}")

    ms = time_steps(fa, step, args.steps, args.warmup)
    flops = 2.0 * n ** 3
    tflops = flops / (ms * 1e-3) / 1e12
    log(f"DGEMM {n}^3: {ms:.3f} ms/step, {tflops:.2f} TFLOP/s "
        f"({tflops / PEAK['f64'] * 100:.1f}% of FP64 matrix peak)")

    traffic = load_traffic(f"dgemm_f64_{n}")
    roof = {"bound": "mfma", "achieved": round(tflops, 3), "peak": PEAK["f64"],
            "unit": "TFLOP/s", "frac": round(tflops / PEAK["f64"], 4), "traffic": traffic,
            "kernel": "gemm_real_kernel<double,false,false>",
            "algorithmic_work": f"2*{n}^3 = {flops:.4g} flop per launch"}

    result = {
        "metric": METRIC, "value": round(tflops, 3), "unit": "TFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded device generator, normal(0,1))",
        "config": {"workload": f"dgemm_native_mfma_{n}", "m": n, "n": n, "k": n,
                   "layout": "row-major", "seeds": [1, 2], "parallelism": "single GPU"},
        "roofline": roof,
    }
    if not args.no_host:
        # The drop-in entry point with host buffers: upload, DGEMM and
        # download in one synchronous call (PCIe-inclusive; not the headline).
        a_h = dA.download(np.float64, n * n)
        b_h = dB.download(np.float64, n * n)
        c_h = np.zeros(n * n)
        walls = []
        for _ in range(2):
            t0 = time.perf_counter()
            if not fa.gemm_f64(n, n, n, 1.0, a_h, b_h, 0.0, c_h):
                raise RuntimeError("ferray_gemm_f64 (host pointers) failed")
            walls.append(time.perf_counter() - t0)
        w = min(walls)
        result["drop_in_host_pointers"] = {
            "ms": round(w * 1e3, 3), "tflops": round(flops / w / 1e12, 3),
            "note": "ferray_gemm_f64 on pageable host buffers: H2D of A, B, DGEMM, D2H of C"}
        log(f"drop-in DGEMM with host buffers: {w * 1e3:.1f} ms ({flops / w / 1e12:.2f} TFLOP/s)")
        del a_h, b_h, c_h
    if not args.no_host:
        # Config 1: the reference's own 512^3 LCG case, on the device.
        import oracle
        m1 = 512
        a1, b1 = oracle.lcg_matrix(m1, 1), oracle.lcg_matrix(m1, 2)
        c1 = np.zeros(m1 * m1)
        d1a, d1b, d1c = (fa.DeviceBuffer(m1 * m1 * 8) for _ in range(3))
        d1a.upload(a1)
        d1b.upload(b1)
        dev_ms = time_steps(fa, lambda: L.ferray_gemm_f64_dev(m1, m1, m1, 1.0, d1a.ptr, d1b.ptr,
                                                              0.0, d1c.ptr, None), 30, 5)
        host_s = _median_time(lambda: fa.gemm_f64(m1, m1, m1, 1.0, a1, b1, 0.0, c1), 30, 5)
        for buf in (d1a, d1b, d1c):
            buf.free()
        result["config1_512"] = {
            "device_resident_us": round(dev_ms * 1e3, 1),
            "host_pointers_us": round(host_s * 1e6, 1),
            "note": "LCG inputs of examples/bench_matmul.rs; device time by HIP events over 30 "
                    "calls after 5 warm-ups; host-pointer call median of 30 after 5 (PCIe "
                    "included)"}
        log(f"config-1 512^3: {dev_ms * 1e3:.1f} us resident, {host_s * 1e6:.1f} us with host "
            f"buffers")
    if not args.no_error:
        log("sampled error against compensated dot products")
        result["error"] = dgemm_error(fa, dA, dB, dC, n)
    if not args.no_split:
        sp = {"method": "integer-modular Ozaki scheme on INT8 MFMA; keys = number of moduli",
              "normal": split_precision(fa, dA, dB, n, args.steps, ms, "normal")}
        fa.fill(dA, n * n, fa.FILL_F64_WIDE, 3, -20, 20)
        fa.fill(dB, n * n, fa.FILL_F64_WIDE, 4, -20, 20)
        sp["wide_u_2e_pm20"] = split_precision(fa, dA, dB, n, args.steps, ms, "wide")
        result["split_precision"] = sp
    for buf in (dA, dB, dC):
        buf.free()
    if not args.no_sweep:
        result["sweep"] = run_sweep(fa, args.steps, args.warmup)
    if not args.no_split:
        result["split_sgemm"] = split_sgemm(fa, n, args.steps, args.warmup)
    if not args.no_consumers:
        result["consumers"] = consumers(fa, args.steps, args.warmup)
    if not args.no_serve:
        result["serving"] = serving(fa)
        b = result["serving"]["best_at_p99_target"]["grouped"]
        if b:
            result["serving_p99_ms"] = b["p99_ms"]
    if args.crossover:
        result["offload_crossover"] = offload_crossover(fa)
    if args.vendor:
        result["vendor_products"] = vendor_products(n)
    if not args.no_cpu_baseline:
        log("CPU baseline (oracle restatement of ferray's Goto DGEMM)")
        result["cpu_baseline"] = cpu_baseline()
        log(f"CPU baseline: {result['cpu_baseline']['value'] * 1e3:.1f} GFLOP/s on "
            f"{result['cpu_baseline']['cores']} threads")
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
