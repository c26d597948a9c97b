// Drop-in entry points with ferray's host-pointer contract: operands are
// staged into per-thread device buffers, the device entry point runs on the
// thread's stream, and the call returns once C is back in host memory.
// Any failure returns false with the reason in ferray_gemm_last_status().
//
// Large operands (8 MiB and up) move through pinned bounce buffers: a ring
// of four 16 MiB pinned slots per thread, filled and drained by a small pool
// of host threads while the DMA engine moves the neighbouring chunk.  The
// caller's buffers are ordinary pageable memory, which the DMA engine reads
// and writes at roughly a third of the pinned rate.
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "common.hpp"

extern "C" void ferray_set_last_status_internal(int s);

namespace fg {
namespace {

struct Staging {
  void* buf[3] = {nullptr, nullptr, nullptr};
  size_t cap[3] = {0, 0, 0};
  ~Staging() {
    if (on_main_thread()) return;  // left to the process's end
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

// ---------------------------------------------------------------------------
// Pinned bounce buffers
// ---------------------------------------------------------------------------

constexpr size_t kChunk = (size_t)16 << 20;        // bytes per pinned slot
constexpr int kSlots = 4;                          // slots per thread
constexpr size_t kPipelineMin = (size_t)8 << 20;   // smaller copies go direct

// Process-wide helper threads for large host memcpy.  copy() splits a range
// into pieces of at least 1 MiB, runs one piece on the caller and the rest
// on the helpers, and returns when all are done.  Concurrent callers share
// the queue.  The pool is never destroyed (no joins at process exit).
class CopyPool {
 public:
  static CopyPool& get() {
    static CopyPool* pool = new CopyPool();
    return *pool;
  }
  void copy(void* dst, const void* src, size_t n) {
    const size_t pieces = std::min<size_t>(workers_ + 1, std::max<size_t>(1, n >> 20));
    if (pieces <= 1) {
      std::memcpy(dst, src, n);
      return;
    }
    const size_t step = (n / pieces + 4095) & ~(size_t)4095;
    std::atomic<size_t> left{0};
    size_t off = step;
    {
      std::lock_guard<std::mutex> g(mu_);
      for (; off < n; off += step) {
        q_.push_back(Job{(char*)dst + off, (const char*)src + off, std::min(step, n - off), &left});
        left.fetch_add(1, std::memory_order_relaxed);
      }
    }
    cv_.notify_all();
    std::memcpy(dst, src, std::min(step, n));
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [&] { return left.load(std::memory_order_acquire) == 0; });
  }

 private:
  struct Job {
    char* dst;
    const char* src;
    size_t n;
    std::atomic<size_t>* left;
  };
  CopyPool() {
    const char* env = std::getenv("FERRAY_HOST_COPY_THREADS");
    unsigned hw = std::thread::hardware_concurrency();
    long want = env ? std::atol(env) : (long)std::min(8u, hw ? hw : 1u);
    workers_ = (size_t)std::max(0L, std::min(want, 64L) - 1);
    for (size_t i = 0; i < workers_; ++i) std::thread([this] { run(); }).detach();
  }
  void run() {
    for (;;) {
      Job j;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return !q_.empty(); });
        j = q_.front();
        q_.pop_front();
      }
      std::memcpy(j.dst, j.src, j.n);
      {
        std::lock_guard<std::mutex> g(mu_);
        j.left->fetch_sub(1, std::memory_order_release);
      }
      done_.notify_all();
    }
  }
  size_t workers_ = 0;
  std::mutex mu_;
  std::condition_variable cv_, done_;
  std::deque<Job> q_;
};

// A thread's ring of pinned slots.  ev[i] is recorded after the last DMA
// that touches slot i, so host code waits on it before reusing the slot.
struct PinnedRing {
  void* slot[kSlots] = {};
  hipEvent_t ev[kSlots] = {};
  unsigned next = 0;
  bool ready = false;
  ~PinnedRing() {
    if (on_main_thread()) return;  // left to the process's end
    for (int i = 0; i < kSlots; ++i) {
      if (ev[i]) (void)hipEventDestroy(ev[i]);
      if (slot[i]) (void)hipHostFree(slot[i]);
    }
  }
  int init() {
    if (ready) return FERRAY_OK;
    for (int i = 0; i < kSlots; ++i) {
      if (!slot[i]) FG_CHECK(hipHostMalloc(&slot[i], kChunk, hipHostMallocDefault));
      if (!ev[i]) FG_CHECK(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
    }
    ready = true;
    return FERRAY_OK;
  }
};
thread_local PinnedRing t_ring;

}  // namespace

// Host -> device, enqueued on `st`; the host buffer may be reused on return.
int host_upload(void* dst, const void* src, size_t bytes, hipStream_t st) {
  if (bytes < kPipelineMin) {
    FG_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
    return FERRAY_OK;
  }
  int s = t_ring.init();
  if (s) return s;
  CopyPool& pool = CopyPool::get();
  for (size_t off = 0; off < bytes; off += kChunk) {
    const int sl = (int)(t_ring.next++ % kSlots);
    const size_t len = std::min(kChunk, bytes - off);
    FG_CHECK(hipEventSynchronize(t_ring.ev[sl]));  // the slot's previous DMA is done
    pool.copy(t_ring.slot[sl], (const char*)src + off, len);
    FG_CHECK(hipMemcpyAsync((char*)dst + off, t_ring.slot[sl], len, hipMemcpyHostToDevice, st));
    FG_CHECK(hipEventRecord(t_ring.ev[sl], st));
  }
  return FERRAY_OK;
}

// Device -> host after the work already on `st`; returns once dst holds the
// data.  Small copies are enqueued only (the caller synchronises).
int host_download(void* dst, const void* src, size_t bytes, hipStream_t st) {
  if (bytes < kPipelineMin) {
    FG_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
    return FERRAY_OK;
  }
  int s = t_ring.init();
  if (s) return s;
  CopyPool& pool = CopyPool::get();
  const size_t n = (bytes + kChunk - 1) / kChunk;
  const unsigned base = t_ring.next;
  t_ring.next += (unsigned)n;
  auto enqueue = [&](size_t i) -> int {
    const int sl = (int)((base + i) % kSlots);
    const size_t off = i * kChunk, len = std::min(kChunk, bytes - off);
    FG_CHECK(hipEventSynchronize(t_ring.ev[sl]));  // prior use, possibly on another stream
    FG_CHECK(hipMemcpyAsync(t_ring.slot[sl], (const char*)src + off, len, hipMemcpyDeviceToHost,
                            st));
    FG_CHECK(hipEventRecord(t_ring.ev[sl], st));
    return FERRAY_OK;
  };
  for (size_t i = 0; i < n && i < (size_t)kSlots; ++i)
    if ((s = enqueue(i))) return s;
  for (size_t i = 0; i < n; ++i) {
    const int sl = (int)((base + i) % kSlots);
    const size_t off = i * kChunk, len = std::min(kChunk, bytes - off);
    FG_CHECK(hipEventSynchronize(t_ring.ev[sl]));
    pool.copy((char*)dst + off, t_ring.slot[sl], len);
    if (i + kSlots < n && (s = enqueue(i + kSlots))) return s;
  }
  return FERRAY_OK;
}

namespace {

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
  auto fail = [&](int e) {
    (void)hipStreamSynchronize(st);
    return e;
  };
  if (abytes && (s = host_upload(da, a, abytes, st))) return fail(s);
  if (bbytes && (s = host_upload(db, b, bbytes, st))) return fail(s);
  if (c_in && cbytes && (s = host_upload(dc, c, cbytes, st))) return fail(s);
  if ((s = launch(da, db, dc, st))) return fail(s);
  if (cbytes && (s = host_download(c, dc, cbytes, st))) return fail(s);
  FG_CHECK(hipStreamSynchronize(st));
  return FERRAY_OK;
}

// A thread's second and third streams and a pool of events for the
// row-panel pipeline below.
struct PanelStreams {
  hipStream_t comp = nullptr, down = nullptr;
  std::vector<hipEvent_t> ev;
  ~PanelStreams() {
    if (on_main_thread()) return;  // left to the process's end
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    if (comp) {
      release_workspace(comp);
      (void)hipStreamDestroy(comp);
    }
    if (down) (void)hipStreamDestroy(down);
  }
  int init(size_t events) {
    if (!comp) FG_CHECK(hipStreamCreateWithFlags(&comp, hipStreamNonBlocking));
    if (!down) FG_CHECK(hipStreamCreateWithFlags(&down, hipStreamNonBlocking));
    while (ev.size() < events) {
      hipEvent_t e;
      FG_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      ev.push_back(e);
    }
    return FERRAY_OK;
  }
};
thread_local PanelStreams t_panels;

// Number of row panels for a product with an m x n C: one per 4 Mi elements
// of C, at most 8, so each panel still fills the chip with tiles.
inline size_t row_panels(size_t m, size_t n) {
  return std::min<size_t>(8, (m * n) >> 22);
}

// Like staged(), for products whose rows of C depend only on the same rows
// of A (and of C when it is read): B goes up first, then A (with C) in row
// panels, and each panel's product is launched on a second stream as soon
// as its rows have landed, so the uploads of later panels and the downloads
// of earlier ones (a third stream) overlap the products.  The panel products
// are the same kernels' results as one call over all rows (every element's
// k-sequence is fixed by the kernels, not by the row count), so this is used
// only where that holds.  Falls back to staged() below 64 MiB of traffic or
// with fewer than two panels.
// launch(row0, rows, dA_rows, dB, dC_rows, stream) runs one panel.
template <typename F>
int staged_rows(const void* a, size_t a_row, const void* b, size_t bbytes, void* c, size_t c_row,
                size_t m, size_t n, bool c_in, F&& launch) {
  const size_t panels = row_panels(m, n);
  const size_t traffic = m * a_row + bbytes + m * c_row * (c_in ? 2 : 1);
  if (panels < 2 || traffic < ((size_t)64 << 20))
    return staged(a, m * a_row, b, bbytes, c, m * c_row, c_in,
                  [&](void* da, void* db, void* dc, hipStream_t st) {
                    return launch((size_t)0, m, da, db, dc, st);
                  });
  int s = ensure_device();
  if (s) return s;
  hipStream_t st = resolve_stream(nullptr);
  if (!st) return FERRAY_ERR_HIP;
  if ((s = t_panels.init(2 * panels))) return s;
  hipStream_t sc = t_panels.comp, sd = t_panels.down;
  void *da = nullptr, *db = nullptr, *dc = nullptr;
  if ((s = t_staging.get(0, m * a_row, &da))) return s;
  if ((s = t_staging.get(1, bbytes, &db))) return s;
  if ((s = t_staging.get(2, m * c_row, &dc))) return s;
  auto fail = [&](int e) {
    (void)hipStreamSynchronize(st);
    (void)hipStreamSynchronize(sc);
    (void)hipStreamSynchronize(sd);
    return e;
  };
  const size_t rows = cdiv(cdiv(m, panels), 128) * 128;
  auto A = [&](size_t r) { return (char*)da + r * a_row; };
  auto C = [&](size_t r) { return (char*)dc + r * c_row; };
  if (bbytes && (s = host_upload(db, b, bbytes, st))) return fail(s);
  size_t np = 0;
  for (size_t r0 = 0; r0 < m; r0 += rows, ++np) {
    const size_t rr = std::min(rows, m - r0);
    hipEvent_t up = t_panels.ev[2 * np], done = t_panels.ev[2 * np + 1];
    if ((s = host_upload(A(r0), (const char*)a + r0 * a_row, rr * a_row, st))) return fail(s);
    if (c_in && (s = host_upload(C(r0), (const char*)c + r0 * c_row, rr * c_row, st)))
      return fail(s);
    if (hipEventRecord(up, st) != hipSuccess || hipStreamWaitEvent(sc, up, 0) != hipSuccess)
      return fail(FERRAY_ERR_HIP);
    if ((s = launch(r0, rr, (void*)A(r0), db, (void*)C(r0), sc))) return fail(s);
    if (hipEventRecord(done, sc) != hipSuccess) return fail(FERRAY_ERR_HIP);
  }
  np = 0;
  for (size_t r0 = 0; r0 < m; r0 += rows, ++np) {
    const size_t rr = std::min(rows, m - r0);
    if (hipStreamWaitEvent(sd, t_panels.ev[2 * np + 1], 0) != hipSuccess)
      return fail(FERRAY_ERR_HIP);
    if ((s = host_download((char*)c + r0 * c_row, C(r0), rr * c_row, sd))) return fail(s);
  }
  if (hipStreamSynchronize(sd) != hipSuccess || hipStreamSynchronize(sc) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return FERRAY_ERR_HIP;
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
  return finish(staged_rows(a, k * 8, b, k * n * 8, c, n * 8, m, n, beta != 0.0,
                            [&](size_t, size_t rows, void* da, void* db, void* dc,
                                hipStream_t st) {
                              return ferray_gemm_f64_dev(rows, n, k, alpha, (const double*)da,
                                                         (const double*)db, beta, (double*)dc,
                                                         st);
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
  return finish(staged_rows(a, k * 4, b, k * n * 4, c, n * 4, m, n, beta != 0.0f,
                            [&](size_t, size_t rows, void* da, void* db, void* dc,
                                hipStream_t st) {
                              return ferray_gemm_f32_dev(rows, n, k, alpha, (const float*)da,
                                                         (const float*)db, beta, (float*)dc, st);
                            }));
}

bool ferray_gemm_c64(size_t m, size_t n, size_t k, double alpha_re, double alpha_im,
                     const double* a, const double* b, double beta_re, double beta_im,
                     double* c) {
  if (m == 0 || n == 0) return finish(FERRAY_OK);
  bool c_in = !(beta_re == 0.0 && beta_im == 0.0);
  return finish(staged_rows(a, k * 16, b, k * n * 16, c, n * 16, m, n, c_in,
                            [&](size_t, size_t rows, void* da, void* db, void* dc,
                                hipStream_t st) {
                              return ferray_gemm_c64_dev(rows, n, k, alpha_re, alpha_im,
                                                         (const double*)da, (const double*)db,
                                                         beta_re, beta_im, (double*)dc, st);
                            }));
}

bool ferray_gemm_c32(size_t m, size_t n, size_t k, float alpha_re, float alpha_im,
                     const float* a, const float* b, float beta_re, float beta_im, float* c) {
  if (m == 0 || n == 0) return finish(FERRAY_OK);
  bool c_in = !(beta_re == 0.0f && beta_im == 0.0f);
  return finish(staged_rows(a, k * 8, b, k * n * 8, c, n * 8, m, n, c_in,
                            [&](size_t, size_t rows, void* da, void* db, void* dc,
                                hipStream_t st) {
                              return ferray_gemm_c32_dev(rows, n, k, alpha_re, alpha_im,
                                                         (const float*)da, (const float*)db,
                                                         beta_re, beta_im, (float*)dc, st);
                            }));
}

bool ferray_gemm_i8(size_t m, size_t n, size_t k, const uint8_t* a, const int8_t* b, int32_t* c,
                    bool accumulate) {
  if (m == 0 || n == 0) return finish(FERRAY_OK);
  return finish(staged_rows(a, k, b, k * n, c, n * 4, m, n, accumulate,
                            [&](size_t, size_t rows, void* da, void* db, void* dc,
                                hipStream_t st) {
                              return ferray_gemm_i8_dev(rows, n, k, (const uint8_t*)da,
                                                        (const int8_t*)db, (int32_t*)dc,
                                                        accumulate, st);
                            }));
}

bool ferray_gemm_i8_signed(size_t m, size_t n, size_t k, const int8_t* a, const int8_t* b,
                           int32_t* c, bool accumulate) {
  if (m == 0 || n == 0) return finish(FERRAY_OK);
  return finish(staged_rows(a, k, b, k * n, c, n * 4, m, n, accumulate,
                            [&](size_t, size_t rows, void* da, void* db, void* dc,
                                hipStream_t st) {
                              return ferray_gemm_i8_signed_dev(rows, n, k, (const int8_t*)da,
                                                               (const int8_t*)db, (int32_t*)dc,
                                                               accumulate, st);
                            }));
}

bool ferray_gemm_i16(size_t m, size_t n, size_t k, const int16_t* a, const int16_t* b,
                     int32_t* c, bool accumulate) {
  if (m == 0 || n == 0) return finish(FERRAY_OK);
  return finish(staged_rows(a, k * 2, b, k * n * 2, c, n * 4, m, n, accumulate,
                            [&](size_t, size_t rows, void* da, void* db, void* dc,
                                hipStream_t st) {
                              return ferray_gemm_i16_dev(rows, n, k, (const int16_t*)da,
                                                         (const int16_t*)db, (int32_t*)dc,
                                                         accumulate, st);
                            }));
}

bool ferray_quantized_matmul(size_t m, size_t n, size_t k, const uint8_t* a, uint8_t a_zp,
                             float scale_a, const int8_t* b, int8_t b_zp, float scale_b,
                             float* c, bool accumulate) {
  if (m == 0 || n == 0) return finish(FERRAY_OK);
  return finish(staged_rows(a, k, b, k * n, c, n * 4, m, n, accumulate,
                            [&](size_t, size_t rows, void* da, void* db, void* dc,
                                hipStream_t st) {
                              return ferray_quantized_matmul_dev(rows, n, k, (const uint8_t*)da,
                                                                 a_zp, scale_a, (const int8_t*)db,
                                                                 b_zp, scale_b, (float*)dc,
                                                                 accumulate, st);
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
