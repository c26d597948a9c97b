//! FFI to `libferray_gemm.so`, the MI355X (gfx950) implementation of
//! ferray-linalg's GEMM family, and safe wrappers that keep ferray's own
//! signatures so `ferray_linalg::gemm` can try the device first behind a
//! feature gate, exactly as it tries OpenBLAS (gemm/mod.rs:211-229), and fall
//! back to its CPU kernels when a call returns `false`.
//!
//! The `extern "C"` block below is the C ABI of `include/ferray_gemm.h`, one
//! declaration per function, in the header's order.
#![allow(non_camel_case_types)]

use std::os::raw::{c_char, c_int, c_void};

// ---------------------------------------------------------------- constants

pub const FERRAY_OK: c_int = 0;
pub const FERRAY_ERR_NO_DEVICE: c_int = 1;
pub const FERRAY_ERR_OOM: c_int = 2;
pub const FERRAY_ERR_INVALID: c_int = 3;
pub const FERRAY_ERR_HIP: c_int = 4;
pub const FERRAY_ERR_UNSUPPORTED: c_int = 5;
pub const FERRAY_ERR_SCRATCH: c_int = 6;

pub const FERRAY_DTYPE_F64: c_int = 0;
pub const FERRAY_DTYPE_F32: c_int = 1;
pub const FERRAY_DTYPE_C64: c_int = 2;
pub const FERRAY_DTYPE_C32: c_int = 3;
pub const FERRAY_DTYPE_I8: c_int = 4;

pub const FERRAY_SERVE_GROUPED: c_int = 0;
pub const FERRAY_SERVE_PER_REQUEST: c_int = 1;

// ---------------------------------------------------------------- C structs

/// One problem of a grouped launch (device pointers, tight row-major).
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct ferray_gemm_problem {
    pub dtype: c_int,
    pub accumulate: c_int,
    pub m: usize,
    pub n: usize,
    pub k: usize,
    pub a: *const c_void,
    pub b: *const c_void,
    pub c: *mut c_void,
    pub alpha_re: f64,
    pub alpha_im: f64,
    pub beta_re: f64,
    pub beta_im: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct ferray_server_config {
    pub hbm_bytes: usize,
    pub max_resident: usize,
    pub max_batch: usize,
    pub lanes: c_int,
    pub mode: c_int,
    pub keep_results: c_int,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct ferray_request {
    pub dtype: c_int,
    pub operands_on_host: c_int,
    pub m: usize,
    pub n: usize,
    pub k: usize,
    pub alpha_re: f64,
    pub alpha_im: f64,
    pub a: *const c_void,
    pub b: *const c_void,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct ferray_server_stats {
    pub submitted: u64,
    pub completed: u64,
    pub launches: u64,
    pub launched_requests: u64,
    pub arena_bytes: usize,
    pub arena_peak: usize,
}

/// Opaque handles.
#[repr(C)]
pub struct ferray_grouped_plan {
    _private: [u8; 0],
}
#[repr(C)]
pub struct ferray_server {
    _private: [u8; 0],
}

// ---------------------------------------------------------------- the C ABI

#[link(name = "ferray_gemm")]
extern "C" {
    pub fn ferray_gemm_f64(m: usize, n: usize, k: usize, alpha: f64, a: *const f64, b: *const f64, beta: f64, c: *mut f64) -> bool;
    pub fn ferray_gemm_f32(m: usize, n: usize, k: usize, alpha: f32, a: *const f32, b: *const f32, beta: f32, c: *mut f32) -> bool;
    pub fn ferray_gemm_c64(m: usize, n: usize, k: usize, alpha_re: f64, alpha_im: f64, a: *const f64, b: *const f64, beta_re: f64, beta_im: f64, c: *mut f64) -> bool;
    pub fn ferray_gemm_c32(m: usize, n: usize, k: usize, alpha_re: f32, alpha_im: f32, a: *const f32, b: *const f32, beta_re: f32, beta_im: f32, c: *mut f32) -> bool;
    pub fn ferray_gemm_i8(m: usize, n: usize, k: usize, a: *const u8, b: *const i8, c: *mut i32, accumulate: bool) -> bool;
    pub fn ferray_gemm_i8_signed(m: usize, n: usize, k: usize, a: *const i8, b: *const i8, c: *mut i32, accumulate: bool) -> bool;
    pub fn ferray_gemm_i16(m: usize, n: usize, k: usize, a: *const i16, b: *const i16, c: *mut i32, accumulate: bool) -> bool;
    pub fn ferray_quantized_matmul(m: usize, n: usize, k: usize, a: *const u8, a_zp: u8, scale_a: f32, b: *const i8, b_zp: i8, scale_b: f32, c: *mut f32, accumulate: bool) -> bool;
    pub fn ferray_gemm_bf16_f32(m: usize, n: usize, k: usize, a: *const u16, b: *const u16, c: *mut f32, accumulate: bool) -> bool;
    pub fn ferray_gemm_f16_f32(m: usize, n: usize, k: usize, a: *const u16, b: *const u16, c: *mut f32, accumulate: bool) -> bool;
    pub fn ferray_gemm_e4m3_f32(m: usize, n: usize, k: usize, a: *const u8, b: *const u8, c: *mut f32, accumulate: bool) -> bool;
    pub fn ferray_gemm_e5m2_f32(m: usize, n: usize, k: usize, a: *const u8, b: *const u8, c: *mut f32, accumulate: bool) -> bool;
    pub fn ferray_gemm_offload_min_dim(dtype: c_int) -> usize;
    pub fn ferray_gpu_available() -> bool;
    pub fn ferray_gemm_last_status() -> c_int;
    pub fn ferray_status_string(status: c_int) -> *const c_char;
    pub fn ferray_gemm_f64_dev(m: usize, n: usize, k: usize, alpha: f64, a: *const f64, b: *const f64, beta: f64, c: *mut f64, stream: *mut c_void) -> c_int;
    pub fn ferray_gemm_f32_dev(m: usize, n: usize, k: usize, alpha: f32, a: *const f32, b: *const f32, beta: f32, c: *mut f32, stream: *mut c_void) -> c_int;
    pub fn ferray_gemm_c64_dev(m: usize, n: usize, k: usize, alpha_re: f64, alpha_im: f64, a: *const f64, b: *const f64, beta_re: f64, beta_im: f64, c: *mut f64, stream: *mut c_void) -> c_int;
    pub fn ferray_gemm_c32_dev(m: usize, n: usize, k: usize, alpha_re: f32, alpha_im: f32, a: *const f32, b: *const f32, beta_re: f32, beta_im: f32, c: *mut f32, stream: *mut c_void) -> c_int;
    pub fn ferray_gemm_i8_dev(m: usize, n: usize, k: usize, a: *const u8, b: *const i8, c: *mut i32, accumulate: bool, stream: *mut c_void) -> c_int;
    pub fn ferray_gemm_i8_signed_dev(m: usize, n: usize, k: usize, a: *const i8, b: *const i8, c: *mut i32, accumulate: bool, stream: *mut c_void) -> c_int;
    pub fn ferray_gemm_i8_signed_dev_noalloc(m: usize, n: usize, k: usize, a: *const i8, b: *const i8, c: *mut i32, accumulate: bool, stream: *mut c_void) -> c_int;
    pub fn ferray_gemm_i16_dev(m: usize, n: usize, k: usize, a: *const i16, b: *const i16, c: *mut i32, accumulate: bool, stream: *mut c_void) -> c_int;
    pub fn ferray_quantized_matmul_dev(m: usize, n: usize, k: usize, a: *const u8, a_zp: u8, scale_a: f32, b: *const i8, b_zp: i8, scale_b: f32, c: *mut f32, accumulate: bool, stream: *mut c_void) -> c_int;
    pub fn ferray_gemm_bf16_f32_dev(m: usize, n: usize, k: usize, a: *const u16, b: *const u16, c: *mut f32, accumulate: bool, stream: *mut c_void) -> c_int;
    pub fn ferray_gemm_f16_f32_dev(m: usize, n: usize, k: usize, a: *const u16, b: *const u16, c: *mut f32, accumulate: bool, stream: *mut c_void) -> c_int;
    pub fn ferray_gemm_e4m3_f32_dev(m: usize, n: usize, k: usize, a: *const u8, b: *const u8, c: *mut f32, accumulate: bool, stream: *mut c_void) -> c_int;
    pub fn ferray_gemm_e5m2_f32_dev(m: usize, n: usize, k: usize, a: *const u8, b: *const u8, c: *mut f32, accumulate: bool, stream: *mut c_void) -> c_int;
    pub fn ferray_gemm_f64_strided_dev(m: usize, n: usize, k: usize, alpha: f64, a: *const f64, lda: usize, b: *const f64, ldb: usize, beta: f64, c: *mut f64, ldc: usize, stream: *mut c_void) -> c_int;
    pub fn ferray_gemm_f32_strided_dev(m: usize, n: usize, k: usize, alpha: f32, a: *const f32, lda: usize, b: *const f32, ldb: usize, beta: f32, c: *mut f32, ldc: usize, stream: *mut c_void) -> c_int;
    pub fn ferray_gemm_c64_strided_dev(m: usize, n: usize, k: usize, alpha_re: f64, alpha_im: f64, a: *const f64, lda: usize, b: *const f64, ldb: usize, beta_re: f64, beta_im: f64, c: *mut f64, ldc: usize, stream: *mut c_void) -> c_int;
    pub fn ferray_gemm_c32_strided_dev(m: usize, n: usize, k: usize, alpha_re: f32, alpha_im: f32, a: *const f32, lda: usize, b: *const f32, ldb: usize, beta_re: f32, beta_im: f32, c: *mut f32, ldc: usize, stream: *mut c_void) -> c_int;
    pub fn ferray_trsm_left_lower_f64(m: usize, n: usize, alpha: f64, a: *const f64, lda: usize, b: *mut f64, ldb: usize, unit_diag: bool) -> bool;
    pub fn ferray_trsm_left_upper_f64(m: usize, n: usize, alpha: f64, a: *const f64, lda: usize, b: *mut f64, ldb: usize, unit_diag: bool) -> bool;
    pub fn ferray_trsm_left_lower_f32(m: usize, n: usize, alpha: f32, a: *const f32, lda: usize, b: *mut f32, ldb: usize, unit_diag: bool) -> bool;
    pub fn ferray_trsm_left_upper_f32(m: usize, n: usize, alpha: f32, a: *const f32, lda: usize, b: *mut f32, ldb: usize, unit_diag: bool) -> bool;
    pub fn ferray_trsm_dev(dtype: c_int, upper: c_int, m: usize, n: usize, alpha: f64, a: *const c_void, lda: usize, b: *mut c_void, ldb: usize, unit_diag: bool, stream: *mut c_void) -> c_int;
    pub fn ferray_dgemm_ozaki(m: usize, n: usize, k: usize, alpha: f64, a: *const f64, b: *const f64, beta: f64, c: *mut f64, moduli: c_int) -> bool;
    pub fn ferray_dgemm_ozaki_dev(m: usize, n: usize, k: usize, alpha: f64, a: *const f64, b: *const f64, beta: f64, c: *mut f64, moduli: c_int, stream: *mut c_void) -> c_int;
    pub fn ferray_ozaki_max_moduli() -> c_int;
    pub fn ferray_sgemm_bf16x3(m: usize, n: usize, k: usize, alpha: f32, a: *const f32, b: *const f32, beta: f32, c: *mut f32, terms: c_int) -> bool;
    pub fn ferray_sgemm_bf16x3_dev(m: usize, n: usize, k: usize, alpha: f32, a: *const f32, b: *const f32, beta: f32, c: *mut f32, terms: c_int, stream: *mut c_void) -> c_int;
    pub fn ferray_gemm_grouped_dev(problems: *const ferray_gemm_problem, count: usize, stream: *mut c_void) -> c_int;
    pub fn ferray_grouped_plan_create(max_problems: usize, max_tiles: usize, out: *mut *mut ferray_grouped_plan) -> c_int;
    pub fn ferray_grouped_plan_launch(plan: *mut ferray_grouped_plan, problems: *const ferray_gemm_problem, count: usize, stream: *mut c_void) -> c_int;
    pub fn ferray_grouped_plan_sync(plan: *mut ferray_grouped_plan) -> c_int;
    pub fn ferray_grouped_plan_destroy(plan: *mut ferray_grouped_plan) -> c_int;
    pub fn ferray_matmul_batched(dtype: c_int, batch: usize, a_batch: usize, b_batch: usize, m: usize, k: usize, n: usize, a: *const c_void, b: *const c_void, c: *mut c_void) -> bool;
    pub fn ferray_matmul_batched_dev(dtype: c_int, batch: usize, a_batch: usize, b_batch: usize, m: usize, k: usize, n: usize, a: *const c_void, b: *const c_void, c: *mut c_void, stream: *mut c_void) -> c_int;
    pub fn ferray_set_tile_table(table: *const c_int, len: usize) -> c_int;
    pub fn ferray_get_tile_table(table: *mut c_int, len: usize) -> c_int;
    pub fn ferray_server_create(cfg: *const ferray_server_config, out: *mut *mut ferray_server) -> c_int;
    pub fn ferray_server_destroy(server: *mut ferray_server) -> c_int;
    pub fn ferray_server_submit(server: *mut ferray_server, req: *const ferray_request, ticket: *mut u64) -> c_int;
    pub fn ferray_server_poll(server: *mut ferray_server, ticket: u64, state: *mut c_int) -> c_int;
    pub fn ferray_server_wait(server: *mut ferray_server, ticket: u64, latency_ms: *mut f64) -> c_int;
    pub fn ferray_server_result(server: *mut ferray_server, ticket: u64, c_host: *mut c_void, bytes: usize) -> c_int;
    pub fn ferray_server_result_dev(server: *mut ferray_server, ticket: u64, c_dev: *mut *const c_void) -> c_int;
    pub fn ferray_server_release(server: *mut ferray_server, ticket: u64) -> c_int;
    pub fn ferray_server_stats_get(server: *mut ferray_server, out: *mut ferray_server_stats) -> c_int;
    pub fn ferray_serve_trace(cfg: *const ferray_server_config, reqs: *const ferray_request, arrival_s: *const f64, count: usize, latency_ms: *mut f64, wall_s: *mut f64, stats: *mut ferray_server_stats) -> c_int;
    pub fn ferray_dev_malloc(bytes: usize, out: *mut *mut c_void) -> c_int;
    pub fn ferray_dev_free(ptr: *mut c_void) -> c_int;
    pub fn ferray_dev_upload(dst: *mut c_void, src: *const c_void, bytes: usize, stream: *mut c_void) -> c_int;
    pub fn ferray_dev_download(dst: *mut c_void, src: *const c_void, bytes: usize, stream: *mut c_void) -> c_int;
    pub fn ferray_dev_memset(dst: *mut c_void, value: c_int, bytes: usize, stream: *mut c_void) -> c_int;
    pub fn ferray_dev_copy(dst: *mut c_void, src: *const c_void, bytes: usize, stream: *mut c_void) -> c_int;
    pub fn ferray_stream_sync(stream: *mut c_void) -> c_int;
    pub fn ferray_default_stream() -> *mut c_void;
    pub fn ferray_event_create(out: *mut *mut c_void) -> c_int;
    pub fn ferray_event_record(event: *mut c_void, stream: *mut c_void) -> c_int;
    pub fn ferray_event_elapsed_ms(start: *mut c_void, stop: *mut c_void, ms: *mut f32) -> c_int;
    pub fn ferray_event_destroy(event: *mut c_void) -> c_int;
    pub fn ferray_device_sync() -> c_int;
    pub fn ferray_dev_fill(dst: *mut c_void, count: usize, kind: c_int, seed: u64, lo: i64, hi: i64, stream: *mut c_void) -> c_int;
}

// ---------------------------------------------------------------- safe face

/// ferray-linalg's `gemm::*` entry points with their signatures unchanged
/// (ferray-linalg/tests/conformance/_surface.json), running on the device.
/// Each returns `false` when the device path cannot run, so callers keep
/// their CPU fallback.  `cpu_supports_*` stay ferray's own; the device probe
/// is `gpu_supports`.
pub mod gemm {
    use super::*;

    /// Device probe (the counterpart of `gemm::cpu_supports_avx2_fma`).
    pub fn gpu_supports() -> bool {
        unsafe { ferray_gpu_available() }
    }

    /// Smallest max(m, n, k) from which offloading a host-pointer product
    /// pays off against ferray's CPU kernels, per `FERRAY_DTYPE_*`.
    pub fn offload_min_dim(dtype: c_int) -> usize {
        unsafe { ferray_gemm_offload_min_dim(dtype) }
    }

    pub fn gemm_f64(m: usize, n: usize, k: usize, alpha: f64, a: &[f64], b: &[f64], beta: f64, c: &mut [f64]) -> bool {
        if a.len() < m * k || b.len() < k * n || c.len() < m * n {
            return false;
        }
        unsafe { ferray_gemm_f64(m, n, k, alpha, a.as_ptr(), b.as_ptr(), beta, c.as_mut_ptr()) }
    }

    pub fn gemm_f32(m: usize, n: usize, k: usize, alpha: f32, a: &[f32], b: &[f32], beta: f32, c: &mut [f32]) -> bool {
        if a.len() < m * k || b.len() < k * n || c.len() < m * n {
            return false;
        }
        unsafe { ferray_gemm_f32(m, n, k, alpha, a.as_ptr(), b.as_ptr(), beta, c.as_mut_ptr()) }
    }

    /// # Safety
    /// `a`, `b`, `c` hold m*k, k*n, m*n interleaved complex values.
    pub unsafe fn gemm_c64(m: usize, n: usize, k: usize, alpha_re: f64, alpha_im: f64, a: *const f64, b: *const f64, beta_re: f64, beta_im: f64, c: *mut f64) -> bool {
        unsafe { ferray_gemm_c64(m, n, k, alpha_re, alpha_im, a, b, beta_re, beta_im, c) }
    }

    /// # Safety
    /// As `gemm_c64`.
    pub unsafe fn gemm_c32(m: usize, n: usize, k: usize, alpha_re: f32, alpha_im: f32, a: *const f32, b: *const f32, beta_re: f32, beta_im: f32, c: *mut f32) -> bool {
        unsafe { ferray_gemm_c32(m, n, k, alpha_re, alpha_im, a, b, beta_re, beta_im, c) }
    }

    /// # Safety
    /// `a`, `b`, `c` hold m*k, k*n, m*n elements.
    pub unsafe fn gemm_i8(m: usize, n: usize, k: usize, a: *const u8, b: *const i8, c: *mut i32, accumulate: bool) -> bool {
        unsafe { ferray_gemm_i8(m, n, k, a, b, c, accumulate) }
    }

    /// # Safety
    /// As `gemm_i8`.
    pub unsafe fn gemm_i8_signed(m: usize, n: usize, k: usize, a: *const i8, b: *const i8, c: *mut i32, accumulate: bool) -> bool {
        unsafe { ferray_gemm_i8_signed(m, n, k, a, b, c, accumulate) }
    }

    /// # Safety
    /// As `gemm_i8`.
    pub unsafe fn gemm_i16(m: usize, n: usize, k: usize, a: *const i16, b: *const i16, c: *mut i32, accumulate: bool) -> bool {
        unsafe { ferray_gemm_i16(m, n, k, a, b, c, accumulate) }
    }

    /// # Safety
    /// As `gemm_i8`.
    pub unsafe fn quantized_matmul(m: usize, n: usize, k: usize, a: *const u8, a_zp: u8, scale_a: f32, b: *const i8, b_zp: i8, scale_b: f32, c: *mut f32, accumulate: bool) -> bool {
        unsafe { ferray_quantized_matmul(m, n, k, a, a_zp, scale_a, b, b_zp, scale_b, c, accumulate) }
    }

    /// # Safety
    /// As `gemm_i8`; bf16 values pass as their 16-bit patterns.
    pub unsafe fn gemm_bf16_f32(m: usize, n: usize, k: usize, a: *const half::bf16, b: *const half::bf16, c: *mut f32, accumulate: bool) -> bool {
        unsafe { ferray_gemm_bf16_f32(m, n, k, a.cast(), b.cast(), c, accumulate) }
    }

    /// # Safety
    /// As `gemm_bf16_f32`.
    pub unsafe fn gemm_f16_f32(m: usize, n: usize, k: usize, a: *const half::f16, b: *const half::f16, c: *mut f32, accumulate: bool) -> bool {
        unsafe { ferray_gemm_f16_f32(m, n, k, a.cast(), b.cast(), c, accumulate) }
    }
}
