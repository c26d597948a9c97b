/*
 * ferray_gemm.h — C ABI of the MI355X (gfx950) GEMM library.
 *
 * The first block mirrors ferray-linalg's BLAS-style GEMM entry points one for
 * one (ferray-linalg/src/gemm/mod.rs): same parameter order and meaning,
 * tightly packed row-major operands (lda = k, ldb = ldc = n), interleaved
 * [re, im] complex storage, host pointers borrowed for the duration of a
 * synchronous call, and a `bool` result that is false iff the backend could
 * not run (no device, allocation failure, launch failure), in which case the
 * caller falls back to its CPU path exactly as ferray's callers do when a
 * hand-tuned kernel returns false (products/mod.rs:509-520, complex.rs:189).
 * Buffers must not overlap (mod.rs:342-346).  All calls are thread safe; each
 * calling thread gets its own HIP stream.
 *
 * The `_dev` block takes device pointers and a hipStream_t (passed as void*,
 * NULL = the calling thread's library stream), enqueues work and returns a
 * status code without synchronising.
 */
#ifndef FERRAY_GEMM_H
#define FERRAY_GEMM_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes of the `_dev` entry points and ferray_gemm_last_status(). */
#define FERRAY_OK 0
#define FERRAY_ERR_NO_DEVICE 1
#define FERRAY_ERR_OOM 2
#define FERRAY_ERR_INVALID 3
#define FERRAY_ERR_HIP 4
#define FERRAY_ERR_UNSUPPORTED 5
#define FERRAY_ERR_SCRATCH 6 /* problem needs more than a fixed scratch holds */

/* ---------------- drop-in entry points (host pointers) ---------------- */

/* C = alpha*A@B + beta*C, f64.  Replaces gemm::gemm_f64 (mod.rs:197-328). */
bool ferray_gemm_f64(size_t m, size_t n, size_t k, double alpha,
                     const double *a, const double *b, double beta, double *c);
/* f32 sibling.  Replaces gemm::gemm_f32 (mod.rs:940-1062). */
bool ferray_gemm_f32(size_t m, size_t n, size_t k, float alpha, const float *a,
                     const float *b, float beta, float *c);
/* Complex<f64>, interleaved.  Replaces gemm::gemm_c64 (mod.rs:757-826). */
bool ferray_gemm_c64(size_t m, size_t n, size_t k, double alpha_re,
                     double alpha_im, const double *a, const double *b,
                     double beta_re, double beta_im, double *c);
/* Complex<f32>, interleaved.  Replaces gemm::gemm_c32 (mod.rs:669-736). */
bool ferray_gemm_c32(size_t m, size_t n, size_t k, float alpha_re,
                     float alpha_im, const float *a, const float *b,
                     float beta_re, float beta_im, float *c);
/* u8 x i8 -> i32, exact wrapping (the VNNI / NEON semantics).
 * Replaces gemm::gemm_i8 (mod.rs:601-651). */
bool ferray_gemm_i8(size_t m, size_t n, size_t k, const uint8_t *a,
                    const int8_t *b, int32_t *c, bool accumulate);
/* i8 x i8 -> i32 wrapping.  Replaces gemm::gemm_i8_signed (mod.rs:414-448). */
bool ferray_gemm_i8_signed(size_t m, size_t n, size_t k, const int8_t *a,
                           const int8_t *b, int32_t *c, bool accumulate);
/* i16 x i16 -> i32 wrapping (mod 2^32).  Replaces gemm::gemm_i16
 * (mod.rs:348-392). */
bool ferray_gemm_i16(size_t m, size_t n, size_t k, const int16_t *a,
                     const int16_t *b, int32_t *c, bool accumulate);
/* Zero-point / scale quantized matmul, f32 out.  Replaces
 * gemm::quantized_matmul (mod.rs:484-575). */
bool ferray_quantized_matmul(size_t m, size_t n, size_t k, const uint8_t *a,
                             uint8_t a_zp, float scale_a, const int8_t *b,
                             int8_t b_zp, float scale_b, float *c,
                             bool accumulate);
/* bf16 x bf16 -> f32 (raw bf16 bits).  Replaces gemm::gemm_bf16_f32
 * (mod.rs:851-882). */
bool ferray_gemm_bf16_f32(size_t m, size_t n, size_t k, const uint16_t *a,
                          const uint16_t *b, float *c, bool accumulate);
/* f16 x f16 -> f32 (raw binary16 bits).  Replaces gemm::gemm_f16_f32
 * (mod.rs:906-933). */
bool ferray_gemm_f16_f32(size_t m, size_t n, size_t k, const uint16_t *a,
                         const uint16_t *b, float *c, bool accumulate);
/* fp8 x fp8 -> f32 on the matrix cores' 8-bit float path: OCP E4M3 (no
 * infinities, 0x7f/0xff NaN, max 448) or OCP E5M2 bytes, row-major, C
 * written or (accumulate) added.  No counterpart in the reference (its 8-bit
 * entries are the integer ones); it is the fp8 member of the family, with
 * the argument order and ownership of ferray_gemm_bf16_f32. */
bool ferray_gemm_e4m3_f32(size_t m, size_t n, size_t k, const uint8_t *a,
                          const uint8_t *b, float *c, bool accumulate);
bool ferray_gemm_e5m2_f32(size_t m, size_t n, size_t k, const uint8_t *a,
                          const uint8_t *b, float *c, bool accumulate);

/* Size-threshold policy for host callers: the smallest max(m, n, k) from
 * which a host-pointer call (upload, product, download) beats ferray's CPU
 * GEMM, measured on MI355X against the CPU restatement under ferray's own
 * dispatch (one thread below max-dim 256, the pool from 256): 128 for
 * FERRAY_DTYPE_F64, 96 for F32 and C64, 64 for C32; I8 (not measured
 * separately) uses the F32 figure.  The environment variable
 * FERRAY_OFFLOAD_MIN_DIM, when set to a positive integer, overrides it for
 * every dtype.  The analogue of the reference's size dispatch
 * (products/mod.rs:445, :478-499 naive / faer / hand-tuned; complex.rs:160):
 * callers keep ferray's CPU kernels below it.  Returns 0 for an unknown
 * dtype. */
size_t ferray_gemm_offload_min_dim(int dtype);

/* Device probe: the analogue of gemm::cpu_supports_avx2_fma
 * (mod.rs:118-188) — true iff a gfx950 device is usable. */
bool ferray_gpu_available(void);
/* Status of the calling thread's last drop-in call (FERRAY_OK ...). */
int ferray_gemm_last_status(void);
/* Human-readable text for a status code. */
const char *ferray_status_string(int status);

/* ---------------- device-pointer entry points ---------------- */

int ferray_gemm_f64_dev(size_t m, size_t n, size_t k, double alpha,
                        const double *a, const double *b, double beta,
                        double *c, void *stream);
int ferray_gemm_f32_dev(size_t m, size_t n, size_t k, float alpha,
                        const float *a, const float *b, float beta, float *c,
                        void *stream);
int ferray_gemm_c64_dev(size_t m, size_t n, size_t k, double alpha_re,
                        double alpha_im, const double *a, const double *b,
                        double beta_re, double beta_im, double *c,
                        void *stream);
int ferray_gemm_c32_dev(size_t m, size_t n, size_t k, float alpha_re,
                        float alpha_im, const float *a, const float *b,
                        float beta_re, float beta_im, float *c, void *stream);
int ferray_gemm_i8_dev(size_t m, size_t n, size_t k, const uint8_t *a,
                       const int8_t *b, int32_t *c, bool accumulate,
                       void *stream);
int ferray_gemm_i8_signed_dev(size_t m, size_t n, size_t k, const int8_t *a,
                              const int8_t *b, int32_t *c, bool accumulate,
                              void *stream);
/* Allocation-free form of ferray_gemm_i8_signed_dev, with the same
 * arguments, semantics and bits.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  It calls no
 * hipMalloc / hipFree; the packed B^T planes (and a packed copy of A when A
 * is not a whole number of tiles) live in one fixed 128 MiB device array per
 * GPU, held until the process ends.  A problem needs
 *   Np*Kp + 4*(Mp + Np) (+ Mp*Kp when A is packed) bytes, rounded up to 256
 * each, where Mp, Np round m, n up to the tile (256 when the product has at
 * least 128 tiles of 256 x 256, else 128) and Kp rounds k up to 128 (256-row
 * tiles) or 64; A is used in place when m is a whole number of tiles, k a
 * multiple of the k step and A 16-byte aligned.  Larger problems return
 * FERRAY_ERR_SCRATCH before anything is enqueued.  `stream` must not be NULL.
 * Every call on one GPU shares that GPU's array: calls that overlap in time
 * on one GPU (different streams or threads) race and must be serialised by
 * the caller. */
int ferray_gemm_i8_signed_dev_noalloc(size_t m, size_t n, size_t k,
                                      const int8_t *a, const int8_t *b,
                                      int32_t *c, bool accumulate,
                                      void *stream);
int ferray_gemm_i16_dev(size_t m, size_t n, size_t k, const int16_t *a,
                        const int16_t *b, int32_t *c, bool accumulate,
                        void *stream);
int ferray_quantized_matmul_dev(size_t m, size_t n, size_t k,
                                const uint8_t *a, uint8_t a_zp, float scale_a,
                                const int8_t *b, int8_t b_zp, float scale_b,
                                float *c, bool accumulate, void *stream);
int ferray_gemm_bf16_f32_dev(size_t m, size_t n, size_t k, const uint16_t *a,
                             const uint16_t *b, float *c, bool accumulate,
                             void *stream);
int ferray_gemm_f16_f32_dev(size_t m, size_t n, size_t k, const uint16_t *a,
                            const uint16_t *b, float *c, bool accumulate,
                            void *stream);
int ferray_gemm_e4m3_f32_dev(size_t m, size_t n, size_t k, const uint8_t *a,
                             const uint8_t *b, float *c, bool accumulate,
                             void *stream);
int ferray_gemm_e5m2_f32_dev(size_t m, size_t n, size_t k, const uint8_t *a,
                             const uint8_t *b, float *c, bool accumulate,
                             void *stream);

/* Row-strided forms of the four floating-point device entries, the
 * drivers' own signatures: C = alpha*A@B + beta*C with A m x k (row stride
 * lda >= k), B k x n (ldb >= n), C m x n (ldc >= n), strides in elements
 * (complex elements for c64/c32); elements of C outside its m x n block are
 * not touched.  Replace outer::gemm_avx2_f64 (outer.rs:198-210),
 * gemm_avx2_f32 (outer.rs:499-511), gemm_avx2_c64 (outer.rs:1231-1245) and
 * gemm_avx2_c32 (outer.rs:1484-1498), which the TRSM trailing update calls
 * on strided blocks (trsm.rs:205-214).  Results equal the tight entries'
 * bit for bit. */
int ferray_gemm_f64_strided_dev(size_t m, size_t n, size_t k, double alpha,
                                const double *a, size_t lda, const double *b,
                                size_t ldb, double beta, double *c, size_t ldc,
                                void *stream);
int ferray_gemm_f32_strided_dev(size_t m, size_t n, size_t k, float alpha,
                                const float *a, size_t lda, const float *b,
                                size_t ldb, float beta, float *c, size_t ldc,
                                void *stream);
int ferray_gemm_c64_strided_dev(size_t m, size_t n, size_t k, double alpha_re,
                                double alpha_im, const double *a, size_t lda,
                                const double *b, size_t ldb, double beta_re,
                                double beta_im, double *c, size_t ldc,
                                void *stream);
int ferray_gemm_c32_strided_dev(size_t m, size_t n, size_t k, float alpha_re,
                                float alpha_im, const float *a, size_t lda,
                                const float *b, size_t ldb, float beta_re,
                                float beta_im, float *c, size_t ldc,
                                void *stream);

/* ---------------- triangular solve (GEMM consumer) ---------------- */

/* Solves L X = alpha B in place: L lower triangular m x m (row stride
 * lda >= m), B m x n (row stride ldb >= n); unit_diag ignores the stored
 * diagonal.  Replaces trsm::trsm_left_lower_f64 (trsm.rs:46-68). */
bool ferray_trsm_left_lower_f64(size_t m, size_t n, double alpha,
                                const double *a, size_t lda, double *b,
                                size_t ldb, bool unit_diag);
/* U X = alpha B, U upper triangular.  Replaces trsm::trsm_left_upper_f64
 * (trsm.rs:70-92). */
bool ferray_trsm_left_upper_f64(size_t m, size_t n, double alpha,
                                const double *a, size_t lda, double *b,
                                size_t ldb, bool unit_diag);
/* f32 siblings.  Replace trsm::trsm_left_lower_f32 / trsm_left_upper_f32
 * (trsm.rs:94-140). */
bool ferray_trsm_left_lower_f32(size_t m, size_t n, float alpha,
                                const float *a, size_t lda, float *b,
                                size_t ldb, bool unit_diag);
bool ferray_trsm_left_upper_f32(size_t m, size_t n, float alpha,
                                const float *a, size_t lda, float *b,
                                size_t ldb, bool unit_diag);
/* Device form: dtype FERRAY_DTYPE_F64 or F32, upper 0/1. */
int ferray_trsm_dev(int dtype, int upper, size_t m, size_t n, double alpha,
                    const void *a, size_t lda, void *b, size_t ldb,
                    bool unit_diag, void *stream);

/* ---------------- split-precision DGEMM ---------------- */

/* C = alpha*A@B + beta*C in f64 computed on INT8 matrix cores by the
 * integer-modular Ozaki scheme: A and B are scaled per row / per column by
 * powers of two to integers, reduced modulo `moduli` pairwise-coprime
 * moduli <= 256 (INT8 residues), multiplied exactly on INT8 MFMA, and
 * recombined by Chinese remaindering in FP64.  moduli in [2, 16] (0 = 16)
 * trades speed for accuracy: each modulus adds ~4 bits to the product's
 * significand.  k up to 2^30: the INT8 products run in k chunks of 65536,
 * each added to the residues mod the modulus (exact: same result as one
 * pass).  Rows/columns holding Inf or NaN give NaN.
 * No counterpart in the reference (its DGEMM is gemm_f64, mod.rs:197-328);
 * same argument order and meaning as ferray_gemm_f64 plus `moduli`. */
bool ferray_dgemm_ozaki(size_t m, size_t n, size_t k, double alpha,
                        const double *a, const double *b, double beta,
                        double *c, int moduli);
int ferray_dgemm_ozaki_dev(size_t m, size_t n, size_t k, double alpha,
                           const double *a, const double *b, double beta,
                           double *c, int moduli, void *stream);
/* Largest supported number of moduli. */
int ferray_ozaki_max_moduli(void);

/* C = alpha*A@B + beta*C in f32 computed on BF16 matrix cores: each f32
 * operand is split into bf16 pieces (x = p0 + p1 + p2 exactly) and the piece
 * products are summed in one f32-accumulating GEMM of depth terms*k.
 * terms = 3 (0 = 3): p0p0 + p0p1 + p1p0, ~2^-16 relative error per
 * product; terms = 6: adds p0p2 + p1p1 + p2p0, f32-level accuracy.
 * No counterpart in the reference (its SGEMM is gemm_f32, mod.rs:940-1062);
 * same argument order and meaning as ferray_gemm_f32 plus `terms`. */
bool ferray_sgemm_bf16x3(size_t m, size_t n, size_t k, float alpha,
                         const float *a, const float *b, float beta, float *c,
                         int terms);
int ferray_sgemm_bf16x3_dev(size_t m, size_t n, size_t k, float alpha,
                            const float *a, const float *b, float beta,
                            float *c, int terms, void *stream);

/* ---------------- grouped (batched, mixed-dtype) GEMM ---------------- */

/* Element types of a grouped problem. */
#define FERRAY_DTYPE_F64 0 /* gemm_f64: C = alpha*A@B + beta*C           */
#define FERRAY_DTYPE_F32 1 /* gemm_f32                                   */
#define FERRAY_DTYPE_C64 2 /* gemm_c64: interleaved complex, complex a/b */
#define FERRAY_DTYPE_C32 3 /* gemm_c32                                   */
#define FERRAY_DTYPE_I8 4  /* gemm_i8_signed: i8 x i8 -> i32 wrapping    */

/* One independent GEMM of a grouped launch: device pointers, tightly packed
 * row-major operands, the semantics of the per-dtype entry point (floating
 * dtypes use alpha/beta, the real part only for real dtypes; I8 uses
 * `accumulate`). */
typedef struct ferray_gemm_problem {
  int dtype;
  int accumulate;
  size_t m, n, k;
  const void *a;
  const void *b;
  void *c;
  double alpha_re, alpha_im, beta_re, beta_im;
} ferray_gemm_problem;

/* Many problems of uneven shape and mixed dtype in one persistent-kernel
 * launch; each problem's result equals its single-problem entry point's bit
 * for bit.  Uses a per-thread plan; returns once enqueued on `stream`.
 * No counterpart in the reference: its only many-GEMM path is
 * matmul_batched (products/mod.rs:579-659), same-shape problems on rayon. */
int ferray_gemm_grouped_dev(const ferray_gemm_problem *problems, size_t count,
                            void *stream);

/* Reusable launch plans for callers with several batches in flight: each
 * plan owns its descriptor buffers and a captured hipGraph of its launch
 * (descriptor upload + persistent kernel) that is replayed per batch.
 * plan_launch waits for the plan's previous launch before refilling it. */
typedef struct ferray_grouped_plan ferray_grouped_plan;
int ferray_grouped_plan_create(size_t max_problems, size_t max_tiles,
                               ferray_grouped_plan **out);
int ferray_grouped_plan_launch(ferray_grouped_plan *plan,
                               const ferray_gemm_problem *problems,
                               size_t count, void *stream);
int ferray_grouped_plan_sync(ferray_grouped_plan *plan);
int ferray_grouped_plan_destroy(ferray_grouped_plan *plan);

/* Batched product with ferray's batch broadcasting: for i < batch,
 *   C[i] = A[i % a_batch] @ B[i % b_batch]   (m x k @ k x n, row-major,
 * matrices stored back to back), as in matmul_batched
 * (products/mod.rs:579-659, indexing at :625-650), all in one grouped
 * launch.  dtype: FERRAY_DTYPE_F64 / F32 / C64 / C32 / I8.  The host form
 * follows the drop-in contract (synchronous, false on failure). */
bool ferray_matmul_batched(int dtype, size_t batch, size_t a_batch,
                           size_t b_batch, size_t m, size_t k, size_t n,
                           const void *a, const void *b, void *c);
int ferray_matmul_batched_dev(int dtype, size_t batch, size_t a_batch,
                              size_t b_batch, size_t m, size_t k, size_t n,
                              const void *a, const void *b, void *c,
                              void *stream);

/* Tile-shape table of the grouped kernel, written by the package's
 * autotuner: entry [((dtype*8 + b)*3 + s)*3 + q] (len 360) is 0 (64x64
 * tiles) or 1 (128x128) for problems with clamp(ceil(log2(max(m, n_real))),
 * 4, 11) = b + 4, min(m, n_real) in [0,64) / [64,128) / [128,..) for
 * s = 0 / 1 / 2 and k_real in [0,128) / [128,512) / [512,..) for q = 0 / 1 / 2
 * (n_real = 2n, k_real = 2k for complex).  A table of len 40, entry
 * [dtype*8 + b], sets every s and q of its bucket; get with len 40 returns
 * the s = 2, q = 2 entries. */
int ferray_set_tile_table(const int *table, size_t len);
int ferray_get_tile_table(int *table, size_t len);

/* ---------------- serving loop ----------------
 * No counterpart in the reference (its GEMM calls are synchronous library
 * calls).  A server owns an HBM arena, a worker thread and `lanes` launch
 * lanes (a stream and a grouped plan each).  Requests are admitted in
 * arrival order while their footprint (C, plus A and B when the operands
 * come from the host) fits in the arena and fewer than `max_resident` are
 * resident.  In GROUPED mode every idle lane takes up to `max_batch`
 * admitted requests into one grouped launch, so requests that arrive while
 * others run are batched into the next launch on another lane and overlap on
 * the device (continuous batching).  PER_REQUEST mode launches each admitted
 * request through its single-problem entry point at once (the baseline).
 * Each request computes C = alpha * A.B (C in the arena); its result is
 * bit-identical in either mode and alone or in any mix. */
#define FERRAY_SERVE_GROUPED 0
#define FERRAY_SERVE_PER_REQUEST 1

typedef struct ferray_server ferray_server;

typedef struct ferray_server_config {
  size_t hbm_bytes;    /* arena size (admission budget), > 0 */
  size_t max_resident; /* admitted-and-unreleased requests at most, > 0 */
  size_t max_batch;    /* requests per grouped launch at most, > 0 */
  int lanes;           /* launch lanes (streams), 1..16 */
  int mode;            /* FERRAY_SERVE_GROUPED or FERRAY_SERVE_PER_REQUEST */
  int keep_results;    /* 1: C stays until ferray_server_result/release;
                          0: C is dropped on completion (benchmarks) */
} ferray_server_config;

typedef struct ferray_request {
  int dtype;            /* FERRAY_DTYPE_*; I8 means i8 x i8 -> i32 */
  int operands_on_host; /* 0: a, b are device-resident; 1: host memory that
                           is copied in (must stay valid until completion) */
  size_t m, n, k;
  double alpha_re, alpha_im; /* I8 ignores alpha */
  const void *a;
  const void *b;
} ferray_request;

typedef struct ferray_server_stats {
  uint64_t submitted, completed, launches, launched_requests;
  size_t arena_bytes, arena_peak;
} ferray_server_stats;

int ferray_server_create(const ferray_server_config *cfg, ferray_server **out);
/* Drains every submitted request, then frees the server. */
int ferray_server_destroy(ferray_server *server);
int ferray_server_submit(ferray_server *server, const ferray_request *req,
                         uint64_t *ticket);
/* state: 0 queued, 1 admitted, 2 running, 3 done. */
int ferray_server_poll(ferray_server *server, uint64_t ticket, int *state);
/* Blocks until done; latency = completion - submission, milliseconds. */
int ferray_server_wait(ferray_server *server, uint64_t ticket,
                       double *latency_ms);
/* keep_results servers: copies C to host memory (`bytes` must equal C's
 * size) and releases the request. */
int ferray_server_result(ferray_server *server, uint64_t ticket, void *c_host,
                         size_t bytes);
/* keep_results servers: device address of C, valid until release. */
int ferray_server_result_dev(ferray_server *server, uint64_t ticket,
                             const void **c_dev);
int ferray_server_release(ferray_server *server, uint64_t ticket);
int ferray_server_stats_get(ferray_server *server, ferray_server_stats *out);

/* Replays a trace through a fresh server (keep_results forced to 0): request
 * i is submitted arrival_s[i] seconds after the start (non-decreasing), and
 * latency_ms[i] = completion - scheduled arrival.  *wall_s = last completion
 * - start.  Returns the first failing status, if any. */
int ferray_serve_trace(const ferray_server_config *cfg, const ferray_request *reqs,
                       const double *arrival_s, size_t count, double *latency_ms,
                       double *wall_s, ferray_server_stats *stats);

/* ---------------- device memory, streams, timing ---------------- */

int ferray_dev_malloc(size_t bytes, void **out);
int ferray_dev_free(void *ptr);
int ferray_dev_upload(void *dst, const void *src, size_t bytes, void *stream);
int ferray_dev_download(void *dst, const void *src, size_t bytes, void *stream);
int ferray_dev_memset(void *dst, int value, size_t bytes, void *stream);
/* Device-to-device copy by a plain copy kernel (enqueued, not synchronised). */
int ferray_dev_copy(void *dst, const void *src, size_t bytes, void *stream);
int ferray_stream_sync(void *stream);
/* The calling thread's library stream (what NULL resolves to). */
void *ferray_default_stream(void);

/* HIP events for timing work on a stream (elapsed time in milliseconds). */
int ferray_event_create(void **out);
int ferray_event_record(void *event, void *stream);
int ferray_event_elapsed_ms(void *start, void *stop, float *ms);
int ferray_event_destroy(void *event);
int ferray_device_sync(void);

/* Seeded device generators (SplitMix64 counter stream, element i of stream
 * `seed` = mix(seed + (i+1)*0x9E3779B97F4A7C15)).
 *   kind 0: f64 uniform [-1,1)          kind 1: f64 normal(0,1) (Box-Muller)
 *   kind 2: f64 u*2^e, u uniform[-1,1), e uniform integer in [lo, hi]
 *   kind 3: f32 uniform [-1,1)          kind 4: f32 normal(0,1)
 *   kind 5: i8  uniform integer [lo, hi]   kind 6: i16 uniform integer [lo, hi]
 *   kind 7: u8  uniform integer [lo, hi]
 * `count` counts scalar elements (complex: 2 per element). */
int ferray_dev_fill(void *dst, size_t count, int kind, uint64_t seed,
                    int64_t lo, int64_t hi, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* FERRAY_GEMM_H */
