// Drop-in entry points with ferray's host-pointer contract: operands are
// staged into per-thread device buffers, the device entry point runs on the
// thread's stream, and the call returns once C is back in host memory.
// Any failure returns false with the reason in ferray_gemm_last_status().
#include <algorithm>
#include <cstring>

#include "common.hpp"

extern "C" void ferray_set_last_status_internal(int s);

namespace fg {
namespace {

struct Staging {
  void* buf[3] = {nullptr, nullptr, nullptr};
  size_t cap[3] = {0, 0, 0};
  ~Staging() {
    for (int i = 0; i < 3; ++i)
      if (buf[i]) (void)hipFree(buf[i]);
  }
  int get(int slot, size_t bytes, void** out) {
    if (bytes == 0) bytes = 16;
    if (cap[slot] < bytes) {
      if (buf[slot]) {
        (void)hipFree(buf[slot]);
        buf[slot] = nullptr;
        cap[slot] = 0;
      }
      FG_CHECK(hipMalloc(&buf[slot], bytes));
      cap[slot] = bytes;
    }
    *out = buf[slot];
    return FERRAY_OK;
  }
};
thread_local Staging t_staging;

// Runs `launch(dA, dB, dC, stream)` with A (abytes), B (bbytes) uploaded and
// C (cbytes) uploaded first when `c_in` and downloaded afterwards.
template <typename F>
int staged(const void* a, size_t abytes, const void* b, size_t bbytes, void* c, size_t cbytes,
           bool c_in, F&& launch) {
  int s = ensure_device();
  if (s) return s;
  hipStream_t st = resolve_stream(nullptr);
  if (!st) return FERRAY_ERR_HIP;
  void *da = nullptr, *db = nullptr, *dc = nullptr;
  if ((s = t_staging.get(0, abytes, &da))) return s;
  if ((s = t_staging.get(1, bbytes, &db))) return s;
  if ((s = t_staging.get(2, cbytes, &dc))) return s;
  if (abytes) FG_CHECK(hipMemcpyAsync(da, a, abytes, hipMemcpyHostToDevice, st));
  if (bbytes) FG_CHECK(hipMemcpyAsync(db, b, bbytes, hipMemcpyHostToDevice, st));
  if (c_in && cbytes) FG_CHECK(hipMemcpyAsync(dc, c, cbytes, hipMemcpyHostToDevice, st));
  s = launch(da, db, dc, st);
  if (s) {
    (void)hipStreamSynchronize(st);
    return s;
  }
  if (cbytes) FG_CHECK(hipMemcpyAsync(c, dc, cbytes, hipMemcpyDeviceToHost, st));
  FG_CHECK(hipStreamSynchronize(st));
  return FERRAY_OK;
}

inline bool finish(int s) {
  ferray_set_last_status_internal(s);
  return s == FERRAY_OK;
}

}  // namespace
}  // namespace fg

using namespace fg;

extern "C" {

bool ferray_gemm_f64(size_t m, size_t n, size_t k, double alpha, const double* a,
                     const double* b, double beta, double* c) {
  if (m == 0 || n == 0) return finish(FERRAY_OK);
  return finish(staged(a, m * k * 8, b, k * n * 8, c, m * n * 8, beta != 0.0,
                       [&](void* da, void* db, void* dc, hipStream_t st) {
                         return ferray_gemm_f64_dev(m, n, k, alpha, (const double*)da,
                                                    (const double*)db, beta, (double*)dc, st);
                       }));
}

bool ferray_matmul_batched(int dtype, size_t batch, size_t a_batch, size_t b_batch, size_t m,
                           size_t k, size_t n, const void* a, const void* b, void* c) {
  if (dtype < FERRAY_DTYPE_F64 || dtype > FERRAY_DTYPE_I8 || a_batch == 0 || b_batch == 0)
    return finish(FERRAY_ERR_INVALID);
  if (batch == 0 || m == 0 || n == 0) return finish(FERRAY_OK);
  const size_t e = (dtype == FERRAY_DTYPE_F64 || dtype == FERRAY_DTYPE_C32) ? 8
                   : dtype == FERRAY_DTYPE_F32                              ? 4
                   : dtype == FERRAY_DTYPE_C64                              ? 16
                                                                            : 1;
  const size_t ce = dtype == FERRAY_DTYPE_I8 ? 4 : e;
  const size_t na = std::min(a_batch, batch), nb = std::min(b_batch, batch);
  return finish(staged(a, na * m * k * e, b, nb * k * n * e, c, batch * m * n * ce, false,
                       [&](void* da, void* db, void* dc, hipStream_t st) {
                         return ferray_matmul_batched_dev(dtype, batch, na, nb, m, k, n, da, db,
                                                          dc, st);
                       }));
}

bool ferray_trsm_left_lower_f64(size_t m, size_t n, double alpha, const double* a, size_t lda,
                                double* b, size_t ldb, bool unit_diag) {
  if (m == 0 || n == 0) return finish(FERRAY_OK);
  return finish(staged(a, m * lda * 8, nullptr, 0, b, m * ldb * 8, true,
                       [&](void* da, void*, void* db, hipStream_t st) {
                         return ferray_trsm_dev(FERRAY_DTYPE_F64, 0, m, n, alpha, da, lda, db, ldb,
                                                unit_diag, st);
                       }));
}

bool ferray_trsm_left_upper_f64(size_t m, size_t n, double alpha, const double* a, size_t lda,
                                double* b, size_t ldb, bool unit_diag) {
  if (m == 0 || n == 0) return finish(FERRAY_OK);
  return finish(staged(a, m * lda * 8, nullptr, 0, b, m * ldb * 8, true,
                       [&](void* da, void*, void* db, hipStream_t st) {
                         return ferray_trsm_dev(FERRAY_DTYPE_F64, 1, m, n, alpha, da, lda, db, ldb,
                                                unit_diag, st);
                       }));
}

bool ferray_trsm_left_lower_f32(size_t m, size_t n, float alpha, const float* a, size_t lda,
                                float* b, size_t ldb, bool unit_diag) {
  if (m == 0 || n == 0) return finish(FERRAY_OK);
  return finish(staged(a, m * lda * 4, nullptr, 0, b, m * ldb * 4, true,
                       [&](void* da, void*, void* db, hipStream_t st) {
                         return ferray_trsm_dev(FERRAY_DTYPE_F32, 0, m, n, alpha, da, lda, db, ldb,
                                                unit_diag, st);
                       }));
}

bool ferray_trsm_left_upper_f32(size_t m, size_t n, float alpha, const float* a, size_t lda,
                                float* b, size_t ldb, bool unit_diag) {
  if (m == 0 || n == 0) return finish(FERRAY_OK);
  return finish(staged(a, m * lda * 4, nullptr, 0, b, m * ldb * 4, true,
                       [&](void* da, void*, void* db, hipStream_t st) {
                         return ferray_trsm_dev(FERRAY_DTYPE_F32, 1, m, n, alpha, da, lda, db, ldb,
                                                unit_diag, st);
                       }));
}

bool ferray_sgemm_bf16x3(size_t m, size_t n, size_t k, float alpha, const float* a,
                         const float* b, float beta, float* c, int terms) {
  if (m == 0 || n == 0) return finish(FERRAY_OK);
  return finish(staged(a, m * k * 4, b, k * n * 4, c, m * n * 4, beta != 0.0f,
                       [&](void* da, void* db, void* dc, hipStream_t st) {
                         return ferray_sgemm_bf16x3_dev(m, n, k, alpha, (const float*)da,
                                                        (const float*)db, beta, (float*)dc, terms,
                                                        st);
                       }));
}

bool ferray_dgemm_ozaki(size_t m, size_t n, size_t k, double alpha, const double* a,
                        const double* b, double beta, double* c, int moduli) {
  if (m == 0 || n == 0) return finish(FERRAY_OK);
  return finish(staged(a, m * k * 8, b, k * n * 8, c, m * n * 8, beta != 0.0,
                       [&](void* da, void* db, void* dc, hipStream_t st) {
                         return ferray_dgemm_ozaki_dev(m, n, k, alpha, (const double*)da,
                                                       (const double*)db, beta, (double*)dc,
                                                       moduli, st);
                       }));
}

bool ferray_gemm_f32(size_t m, size_t n, size_t k, float alpha, const float* a, const float* b,
                     float beta, float* c) {
  if (m == 0 || n == 0) return finish(FERRAY_OK);
  return finish(staged(a, m * k * 4, b, k * n * 4, c, m * n * 4, beta != 0.0f,
                       [&](void* da, void* db, void* dc, hipStream_t st) {
                         return ferray_gemm_f32_dev(m, n, k, alpha, (const float*)da,
                                                    (const float*)db, beta, (float*)dc, st);
                       }));
}

bool ferray_gemm_c64(size_t m, size_t n, size_t k, double alpha_re, double alpha_im,
                     const double* a, const double* b, double beta_re, double beta_im,
                     double* c) {
  if (m == 0 || n == 0) return finish(FERRAY_OK);
  bool c_in = !(beta_re == 0.0 && beta_im == 0.0);
  return finish(staged(a, m * k * 16, b, k * n * 16, c, m * n * 16, c_in,
                       [&](void* da, void* db, void* dc, hipStream_t st) {
                         return ferray_gemm_c64_dev(m, n, k, alpha_re, alpha_im,
                                                    (const double*)da, (const double*)db,
                                                    beta_re, beta_im, (double*)dc, st);
                       }));
}

bool ferray_gemm_c32(size_t m, size_t n, size_t k, float alpha_re, float alpha_im,
                     const float* a, const float* b, float beta_re, float beta_im, float* c) {
  if (m == 0 || n == 0) return finish(FERRAY_OK);
  bool c_in = !(beta_re == 0.0f && beta_im == 0.0f);
  return finish(staged(a, m * k * 8, b, k * n * 8, c, m * n * 8, c_in,
                       [&](void* da, void* db, void* dc, hipStream_t st) {
                         return ferray_gemm_c32_dev(m, n, k, alpha_re, alpha_im,
                                                    (const float*)da, (const float*)db, beta_re,
                                                    beta_im, (float*)dc, st);
                       }));
}

bool ferray_gemm_i8(size_t m, size_t n, size_t k, const uint8_t* a, const int8_t* b, int32_t* c,
                    bool accumulate) {
  if (m == 0 || n == 0) return finish(FERRAY_OK);
  return finish(staged(a, m * k, b, k * n, c, m * n * 4, accumulate,
                       [&](void* da, void* db, void* dc, hipStream_t st) {
                         return ferray_gemm_i8_dev(m, n, k, (const uint8_t*)da,
                                                   (const int8_t*)db, (int32_t*)dc, accumulate,
                                                   st);
                       }));
}

bool ferray_gemm_i8_signed(size_t m, size_t n, size_t k, const int8_t* a, const int8_t* b,
                           int32_t* c, bool accumulate) {
  if (m == 0 || n == 0) return finish(FERRAY_OK);
  return finish(staged(a, m * k, b, k * n, c, m * n * 4, accumulate,
                       [&](void* da, void* db, void* dc, hipStream_t st) {
                         return ferray_gemm_i8_signed_dev(m, n, k, (const int8_t*)da,
                                                          (const int8_t*)db, (int32_t*)dc,
                                                          accumulate, st);
                       }));
}

bool ferray_gemm_i16(size_t m, size_t n, size_t k, const int16_t* a, const int16_t* b,
                     int32_t* c, bool accumulate) {
  if (m == 0 || n == 0) return finish(FERRAY_OK);
  return finish(staged(a, m * k * 2, b, k * n * 2, c, m * n * 4, accumulate,
                       [&](void* da, void* db, void* dc, hipStream_t st) {
                         return ferray_gemm_i16_dev(m, n, k, (const int16_t*)da,
                                                    (const int16_t*)db, (int32_t*)dc,
                                                    accumulate, st);
                       }));
}

bool ferray_quantized_matmul(size_t m, size_t n, size_t k, const uint8_t* a, uint8_t a_zp,
                             float scale_a, const int8_t* b, int8_t b_zp, float scale_b,
                             float* c, bool accumulate) {
  if (m == 0 || n == 0) return finish(FERRAY_OK);
  return finish(staged(a, m * k, b, k * n, c, m * n * 4, accumulate,
                       [&](void* da, void* db, void* dc, hipStream_t st) {
                         return ferray_quantized_matmul_dev(m, n, k, (const uint8_t*)da, a_zp,
                                                            scale_a, (const int8_t*)db, b_zp,
                                                            scale_b, (float*)dc, accumulate,
                                                            st);
                       }));
}

bool ferray_gemm_bf16_f32(size_t m, size_t n, size_t k, const uint16_t* a, const uint16_t* b,
                          float* c, bool accumulate) {
  if (m == 0 || n == 0) return finish(FERRAY_OK);
  return finish(staged(a, m * k * 2, b, k * n * 2, c, m * n * 4, accumulate,
                       [&](void* da, void* db, void* dc, hipStream_t st) {
                         return ferray_gemm_bf16_f32_dev(m, n, k, (const uint16_t*)da,
                                                         (const uint16_t*)db, (float*)dc,
                                                         accumulate, st);
                       }));
}

bool ferray_gemm_f16_f32(size_t m, size_t n, size_t k, const uint16_t* a, const uint16_t* b,
                         float* c, bool accumulate) {
  if (m == 0 || n == 0) return finish(FERRAY_OK);
  return finish(staged(a, m * k * 2, b, k * n * 2, c, m * n * 4, accumulate,
                       [&](void* da, void* db, void* dc, hipStream_t st) {
                         return ferray_gemm_f16_f32_dev(m, n, k, (const uint16_t*)da,
                                                        (const uint16_t*)db, (float*)dc,
                                                        accumulate, st);
                       }));
}

bool ferray_gemm_e4m3_f32(size_t m, size_t n, size_t k, const uint8_t* a, const uint8_t* b,
                          float* c, bool accumulate) {
  if (m == 0 || n == 0) return finish(FERRAY_OK);
  return finish(staged(a, m * k, b, k * n, c, m * n * 4, accumulate,
                       [&](void* da, void* db, void* dc, hipStream_t st) {
                         return ferray_gemm_e4m3_f32_dev(m, n, k, (const uint8_t*)da,
                                                         (const uint8_t*)db, (float*)dc,
                                                         accumulate, st);
                       }));
}

bool ferray_gemm_e5m2_f32(size_t m, size_t n, size_t k, const uint8_t* a, const uint8_t* b,
                          float* c, bool accumulate) {
  if (m == 0 || n == 0) return finish(FERRAY_OK);
  return finish(staged(a, m * k, b, k * n, c, m * n * 4, accumulate,
                       [&](void* da, void* db, void* dc, hipStream_t st) {
                         return ferray_gemm_e5m2_f32_dev(m, n, k, (const uint8_t*)da,
                                                         (const uint8_t*)db, (float*)dc,
                                                         accumulate, st);
                       }));
}

}  // extern "C"
