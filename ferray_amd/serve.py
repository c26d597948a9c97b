"""Serving loop: continuous batching of matmul requests on one device.

``Server`` wraps the library's request loop (admission by HBM footprint,
grouped launches on several lanes, completion polling); ``run_trace``
replays a timed request trace through a fresh server in native code and
returns per-request latencies.  ``make_trace`` builds the synthetic trace
of the serving workload: Poisson arrivals, M, N, K log-uniform, a dtype mix,
operands taken from device-resident pools of seeded random data.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import (DTYPE_C32, DTYPE_C64, DTYPE_F32, DTYPE_F64, DTYPE_I8, SERVE_GROUPED,
                   SERVE_PER_REQUEST, Request, ServerConfig, ServerStats, check, lib)

__all__ = ["Server", "Request", "ServerConfig", "ServerStats", "config", "request", "run_trace",
           "OperandPools", "make_trace", "SERVE_GROUPED", "SERVE_PER_REQUEST", "c_dtype",
           "c_elem_bytes"]

_ELEM = {DTYPE_F64: 8, DTYPE_F32: 4, DTYPE_C64: 16, DTYPE_C32: 8, DTYPE_I8: 1}
_C_ELEM = {DTYPE_F64: 8, DTYPE_F32: 4, DTYPE_C64: 16, DTYPE_C32: 8, DTYPE_I8: 4}
_C_NP = {DTYPE_F64: np.float64, DTYPE_F32: np.float32, DTYPE_C64: np.complex128,
         DTYPE_C32: np.complex64, DTYPE_I8: np.int32}


def c_dtype(dtype: int):
    return _C_NP[dtype]


def c_elem_bytes(dtype: int) -> int:
    return _C_ELEM[dtype]


def config(hbm_bytes=8 << 30, max_resident=4096, max_batch=256, lanes=4, mode=SERVE_GROUPED,
           keep_results=False) -> ServerConfig:
    return ServerConfig(hbm_bytes, max_resident, max_batch, lanes, mode, int(bool(keep_results)))


def request(dtype, m, n, k, a, b, alpha=1.0, on_host=False) -> Request:
    alpha = complex(alpha)
    return Request(dtype, int(bool(on_host)), m, n, k, alpha.real, alpha.imag, a, b)


class Server:
    def __init__(self, cfg: ServerConfig):
        p = C.c_void_p()
        check(lib().ferray_server_create(C.byref(cfg), C.byref(p)), "server_create")
        self.ptr = p.value
        self._keep = []  # host operands must outlive their requests

    def submit(self, req: Request, keep_alive=None) -> int:
        t = C.c_uint64()
        check(lib().ferray_server_submit(self.ptr, C.byref(req), C.byref(t)), "server_submit")
        if keep_alive is not None:
            self._keep.append(keep_alive)
        return t.value

    def poll(self, ticket: int) -> int:
        st = C.c_int()
        check(lib().ferray_server_poll(self.ptr, ticket, C.byref(st)), "server_poll")
        return st.value

    def wait(self, ticket: int) -> float:
        ms = C.c_double()
        check(lib().ferray_server_wait(self.ptr, ticket, C.byref(ms)), "server_wait")
        return ms.value

    def result(self, ticket: int, dtype: int, m: int, n: int) -> np.ndarray:
        out = np.empty(m * n, dtype=_C_NP[dtype])
        check(lib().ferray_server_result(self.ptr, ticket, out.ctypes.data, out.nbytes),
              "server_result")
        return out

    def result_dev(self, ticket: int) -> int:
        p = C.c_void_p()
        check(lib().ferray_server_result_dev(self.ptr, ticket, C.byref(p)), "server_result_dev")
        return p.value

    def release(self, ticket: int) -> None:
        check(lib().ferray_server_release(self.ptr, ticket), "server_release")

    def stats(self) -> ServerStats:
        st = ServerStats()
        check(lib().ferray_server_stats_get(self.ptr, C.byref(st)), "server_stats")
        return st

    def close(self) -> None:
        if self.ptr:
            check(lib().ferray_server_destroy(self.ptr), "server_destroy")
            self.ptr = None
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


def run_trace(cfg: ServerConfig, reqs, arrival_s: np.ndarray):
    """Replays the trace natively; returns (latency_ms array, wall seconds, stats)."""
    n = len(reqs)
    arr = (Request * n)(*reqs)
    arrival = np.ascontiguousarray(arrival_s, dtype=np.float64)
    lat = np.empty(n, dtype=np.float64)
    wall = C.c_double()
    st = ServerStats()
    check(lib().ferray_serve_trace(C.byref(cfg), C.cast(arr, C.c_void_p), arrival.ctypes.data, n,
                                   lat.ctypes.data, C.byref(wall), C.byref(st)), "serve_trace")
    return lat, wall.value, st


class OperandPools:
    """Device-resident pools of seeded random data that requests slice.

    f64 (also read as interleaved complex<f64>), f32 (also complex<f32>) and
    i8 pools, each filled by the library's seeded generator: normal(0,1) for
    floats, uniform integers in [-128, 127] for i8.
    """

    def __init__(self, fa, elems: int = 1 << 24, seed: int = 7):
        self.elems = elems
        self.f64 = fa.DeviceBuffer(elems * 8)
        self.f32 = fa.DeviceBuffer(elems * 4)
        self.i8 = fa.DeviceBuffer(elems)
        fa.fill(self.f64, elems, fa.FILL_F64_NORMAL, seed)
        fa.fill(self.f32, elems, fa.FILL_F32_NORMAL, seed + 1)
        fa.fill(self.i8, elems, fa.FILL_I8, seed + 2, -128, 127)
        fa.device_sync()

    def buffer(self, dtype):
        return {DTYPE_F64: self.f64, DTYPE_C64: self.f64, DTYPE_F32: self.f32,
                DTYPE_C32: self.f32, DTYPE_I8: self.i8}[dtype]

    def slice(self, dtype: int, nbytes: int, rng: np.random.Generator) -> int:
        buf = self.buffer(dtype)
        limit = buf.nbytes - nbytes
        if limit < 0:
            raise ValueError("operand larger than its pool")
        return buf.ptr + int(rng.integers(0, limit // 256 + 1)) * 256

    def free(self):
        for b in (self.f64, self.f32, self.i8):
            b.free()


MIX = ((DTYPE_F64, 0.5), (DTYPE_F32, 0.3), (DTYPE_C64, 0.1), (DTYPE_I8, 0.1))


def make_trace(pools: OperandPools, count: int, rate_per_s: float, seed: int = 0,
               lo: int = 16, hi: int = 2048, mix=MIX):
    """(requests, arrival seconds, flop per request) for a Poisson stream."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_per_s, size=count)
    arrival = np.cumsum(gaps) - gaps[0]
    dims = np.exp(rng.uniform(np.log(lo), np.log(hi), size=(count, 3))).astype(np.int64)
    dims = np.clip(dims, lo, hi)
    kinds = np.array([d for d, _ in mix])
    probs = np.array([p for _, p in mix])
    dts = rng.choice(kinds, size=count, p=probs / probs.sum())
    reqs, flops = [], np.empty(count)
    for i in range(count):
        dt = int(dts[i])
        m, n, k = (int(v) for v in dims[i])
        e = _ELEM[dt]
        a = pools.slice(dt, m * k * e, rng)
        b = pools.slice(dt, k * n * e, rng)
        reqs.append(request(dt, m, n, k, a, b))
        flops[i] = (8.0 if dt in (DTYPE_C64, DTYPE_C32) else 2.0) * m * n * k
    return reqs, arrival, flops
