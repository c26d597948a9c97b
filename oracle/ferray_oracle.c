/*
 * ferray GEMM oracle — TEST INFRASTRUCTURE ONLY.
 *
 * A CPU restatement, in plain C, of the arithmetic that ferray-linalg's
 * hand-tuned GEMM family performs (ferray-linalg/src/gemm/).  Only the test
 * suite, the package's smoke check and the benchmark's CPU-baseline leg load
 * this library, and only as the checker / the reference CPU timing.  The
 * product library never links or calls it.
 *
 * What each function follows (reference file:line) and what pins it is listed
 * beside the function.  Pins live in tests/test_oracle.py:
 *   - reference known-answer tests and golden vectors (tests/golden/),
 *   - closed forms / invariants (integer linearity, exact small-integer
 *     products, the KC-block formula against an independent long-double sum),
 *   - brute force on tiny inputs.
 *
 * Real f64/f32 ("ferray formula"): for max(m,n,k) of any size, every element
 * of C produced by gemm_avx2_f64 / gemm_avx2_f32 is
 *     acc_p = fma-chain over k in [256p, 256p+256)      (kernel_avx2.rs:92-142)
 *     acc_p *= alpha   if alpha != 1                   (kernel_avx2.rs:165-175)
 *     C = store(C, acc_p, beta_p)  beta_0 = beta, beta_p = 1 for p > 0
 *                                                      (outer.rs:416-421)
 * with store(): beta==0 -> acc; beta==1 -> C+acc; else fma(C,beta,acc)
 * (kernel_avx2.rs:198-216; the scalar edge tile uses beta*C+acc unfused,
 * kernel_avx2.rs:280-287, so general beta is placement dependent and is
 * compared by tolerance only).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>
#if defined(__x86_64__)
#include <immintrin.h>
#endif

#define ORC_KC 256 /* outer.rs:67 */
#define ORC_MC 64  /* outer.rs:66 */
#define ORC_NC 512 /* outer.rs:68 */

/* ------------------------------------------------------------------ */
/* Input generators                                                     */
/* ------------------------------------------------------------------ */

/* ferray-linalg/examples/bench_matmul.rs:21-34 (make_matrix): LCG with
 * multiplier 6364136223846793005, increment 1, x = (state>>33)/2^31 - 1. */
void orc_lcg_matrix(size_t count, uint64_t seed, double *out) {
  uint64_t state = seed * 6364136223846793005ULL + 1ULL;
  for (size_t i = 0; i < count; ++i) {
    state = state * 6364136223846793005ULL + 1ULL;
    out[i] = (double)(state >> 33) / (double)(1ULL << 31) - 1.0;
  }
}

/* SplitMix64 (Steele, Lea, Flood 2014), counter form: value i of stream
 * `seed` is mix(seed + (i+1)*0x9E3779B97F4A7C15).  The device generators
 * in the product library use the same function so that the CPU can
 * regenerate any element. */
static inline uint64_t orc_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
uint64_t orc_splitmix64(uint64_t seed, uint64_t i) {
  return orc_mix64(seed + (i + 1) * 0x9E3779B97F4A7C15ULL);
}
/* uniform in [-1,1): 53 random bits, exact in double */
void orc_uniform_f64(size_t count, uint64_t seed, double *out) {
  for (size_t i = 0; i < count; ++i) {
    uint64_t r = orc_splitmix64(seed, i) >> 11;
    out[i] = (double)r * 0x1.0p-52 - 1.0;
  }
}
/* integers uniform in [lo, hi] (inclusive), range < 2^32 */
void orc_uniform_int(size_t count, uint64_t seed, int64_t lo, int64_t hi,
                     int64_t *out) {
  uint64_t span = (uint64_t)(hi - lo + 1);
  for (size_t i = 0; i < count; ++i) {
    uint64_t r = orc_splitmix64(seed, i) >> 32;
    out[i] = lo + (int64_t)((r * span) >> 32);
  }
}

/* ------------------------------------------------------------------ */
/* Real f64 / f32: the ferray KC-block formula (scalar)                */
/* ------------------------------------------------------------------ */

/* outer.rs:198-231 (m==0||n==0 no-op; k==0 -> C = beta*C with beta==0
 * writing zeros and beta==1 leaving C untouched) and outer.rs:393-488 +
 * kernel_avx2.rs:53-216 for k > 0. */
void orc_gemm_f64(size_t m, size_t n, size_t k, double alpha, const double *a,
                  const double *b, double beta, double *c) {
  if (m == 0 || n == 0) return;
  if (k == 0) {
    for (size_t i = 0; i < m * n; ++i) {
      if (beta == 0.0) c[i] = 0.0;
      else if (beta != 1.0) c[i] *= beta;
    }
    return;
  }
  for (size_t i = 0; i < m; ++i) {
    for (size_t j = 0; j < n; ++j) {
      double cv = c[i * n + j];
      for (size_t p0 = 0; p0 < k; p0 += ORC_KC) {
        size_t p1 = p0 + ORC_KC < k ? p0 + ORC_KC : k;
        double acc = 0.0;
        for (size_t p = p0; p < p1; ++p) acc = fma(a[i * k + p], b[p * n + j], acc);
        if (alpha != 1.0) acc *= alpha;
        double bt = p0 == 0 ? beta : 1.0;
        if (bt == 0.0) cv = acc;
        else if (bt == 1.0) cv = cv + acc;
        else cv = fma(cv, bt, acc);
      }
      c[i * n + j] = cv;
    }
  }
}

/* outer.rs:499-687 + kernel_avx2_f32.rs:56-230: identical structure, f32. */
void orc_gemm_f32(size_t m, size_t n, size_t k, float alpha, const float *a,
                  const float *b, float beta, float *c) {
  if (m == 0 || n == 0) return;
  if (k == 0) {
    for (size_t i = 0; i < m * n; ++i) {
      if (beta == 0.0f) c[i] = 0.0f;
      else if (beta != 1.0f) c[i] *= beta;
    }
    return;
  }
  for (size_t i = 0; i < m; ++i) {
    for (size_t j = 0; j < n; ++j) {
      float cv = c[i * n + j];
      for (size_t p0 = 0; p0 < k; p0 += ORC_KC) {
        size_t p1 = p0 + ORC_KC < k ? p0 + ORC_KC : k;
        float acc = 0.0f;
        for (size_t p = p0; p < p1; ++p) acc = fmaf(a[i * k + p], b[p * n + j], acc);
        if (alpha != 1.0f) acc *= alpha;
        float bt = p0 == 0 ? beta : 1.0f;
        if (bt == 0.0f) cv = acc;
        else if (bt == 1.0f) cv = cv + acc;
        else cv = fmaf(cv, bt, acc);
      }
      c[i * n + j] = cv;
    }
  }
}

/* ------------------------------------------------------------------ */
/* Complex c64 / c32 (interleaved [re, im], row-major)                  */
/* ------------------------------------------------------------------ */

/* kernel_avx2_complex_f64.rs:4-28,110-230 (interior tile): per KC block,
 * four fma chains rr=sum ar*br, ri=sum ar*bi, ii=sum ai*bi, ir=sum ai*br,
 * combined re = rr - ii, im = ri + ir (vaddsubpd), scaled by alpha with the
 * same split, stored with beta on the first block and 1 after it.
 * k == 0 follows outer.rs:1249-1266.  Edge tiles (kernel_avx2_complex_f64.rs
 * :259-300) round differently, so complex parity is by tolerance. */
void orc_gemm_c64(size_t m, size_t n, size_t k, double alpha_re,
                  double alpha_im, const double *a, const double *b,
                  double beta_re, double beta_im, double *c) {
  if (m == 0 || n == 0) return;
  if (k == 0) {
    for (size_t i = 0; i < m * n; ++i) {
      double *p = c + 2 * i;
      if (beta_re == 0.0 && beta_im == 0.0) { p[0] = 0.0; p[1] = 0.0; }
      else if (!(beta_re == 1.0 && beta_im == 0.0)) {
        double cr = p[0], ci = p[1];
        p[0] = beta_re * cr - beta_im * ci;
        p[1] = beta_re * ci + beta_im * cr;
      }
    }
    return;
  }
  for (size_t i = 0; i < m; ++i) {
    for (size_t j = 0; j < n; ++j) {
      double cr = c[2 * (i * n + j)], ci = c[2 * (i * n + j) + 1];
      for (size_t p0 = 0; p0 < k; p0 += ORC_KC) {
        size_t p1 = p0 + ORC_KC < k ? p0 + ORC_KC : k;
        double rr = 0, ri = 0, ii = 0, ir = 0;
        for (size_t p = p0; p < p1; ++p) {
          double ar = a[2 * (i * k + p)], ai = a[2 * (i * k + p) + 1];
          double br = b[2 * (p * n + j)], bi = b[2 * (p * n + j) + 1];
          rr = fma(ar, br, rr); ri = fma(ar, bi, ri);
          ii = fma(ai, bi, ii); ir = fma(ai, br, ir);
        }
        double vr = rr - ii, vi = ri + ir;
        if (!(alpha_re == 1.0 && alpha_im == 0.0)) {
          double tr = alpha_re * vr, ti = alpha_re * vi;
          double sr = alpha_im * vi, si = alpha_im * vr;
          vr = tr - sr; vi = ti + si;
        }
        double bre = p0 == 0 ? beta_re : 1.0, bim = p0 == 0 ? beta_im : 0.0;
        if (bre == 0.0 && bim == 0.0) { cr = vr; ci = vi; }
        else if (bre == 1.0 && bim == 0.0) { cr = cr + vr; ci = ci + vi; }
        else {
          double sr = bre * cr - bim * ci, si = bre * ci + bim * cr;
          cr = sr + vr; ci = si + vi;
        }
      }
      c[2 * (i * n + j)] = cr; c[2 * (i * n + j) + 1] = ci;
    }
  }
}

/* kernel_avx2_complex_f32.rs: same scheme in f32. */
void orc_gemm_c32(size_t m, size_t n, size_t k, float alpha_re, float alpha_im,
                  const float *a, const float *b, float beta_re, float beta_im,
                  float *c) {
  if (m == 0 || n == 0) return;
  if (k == 0) {
    for (size_t i = 0; i < m * n; ++i) {
      float *p = c + 2 * i;
      if (beta_re == 0.0f && beta_im == 0.0f) { p[0] = 0.0f; p[1] = 0.0f; }
      else if (!(beta_re == 1.0f && beta_im == 0.0f)) {
        float cr = p[0], ci = p[1];
        p[0] = beta_re * cr - beta_im * ci;
        p[1] = beta_re * ci + beta_im * cr;
      }
    }
    return;
  }
  for (size_t i = 0; i < m; ++i) {
    for (size_t j = 0; j < n; ++j) {
      float cr = c[2 * (i * n + j)], ci = c[2 * (i * n + j) + 1];
      for (size_t p0 = 0; p0 < k; p0 += ORC_KC) {
        size_t p1 = p0 + ORC_KC < k ? p0 + ORC_KC : k;
        float rr = 0, ri = 0, ii = 0, ir = 0;
        for (size_t p = p0; p < p1; ++p) {
          float ar = a[2 * (i * k + p)], ai = a[2 * (i * k + p) + 1];
          float br = b[2 * (p * n + j)], bi = b[2 * (p * n + j) + 1];
          rr = fmaf(ar, br, rr); ri = fmaf(ar, bi, ri);
          ii = fmaf(ai, bi, ii); ir = fmaf(ai, br, ir);
        }
        float vr = rr - ii, vi = ri + ir;
        if (!(alpha_re == 1.0f && alpha_im == 0.0f)) {
          float tr = alpha_re * vr, ti = alpha_re * vi;
          float sr = alpha_im * vi, si = alpha_im * vr;
          vr = tr - sr; vi = ti + si;
        }
        float bre = p0 == 0 ? beta_re : 1.0f, bim = p0 == 0 ? beta_im : 0.0f;
        if (bre == 0.0f && bim == 0.0f) { cr = vr; ci = vi; }
        else if (bre == 1.0f && bim == 0.0f) { cr = cr + vr; ci = ci + vi; }
        else {
          float sr = bre * cr - bim * ci, si = bre * ci + bim * cr;
          cr = sr + vr; ci = si + vi;
        }
      }
      c[2 * (i * n + j)] = cr; c[2 * (i * n + j) + 1] = ci;
    }
  }
}

/* ------------------------------------------------------------------ */
/* Integer GEMM: wrapping i32 (mod 2^32)                                */
/* ------------------------------------------------------------------ */

/* Products are accumulated in uint32 so that overflow wraps exactly like
 * Rust's wrapping_add (faer_fallback.rs:245-341, integer_tests.rs:15-37).
 * `accumulate` adds onto C (mod 2^32); k == 0 with !accumulate zeroes C
 * (outer.rs:2243-2252). */
#define ORC_INT_GEMM(NAME, TA, TB)                                           \
  void NAME(size_t m, size_t n, size_t k, const TA *a, const TB *b,          \
            int32_t *c, int accumulate) {                                    \
    if (m == 0 || n == 0) return;                                            \
    for (size_t i = 0; i < m; ++i) {                                         \
      for (size_t j = 0; j < n; ++j) {                                       \
        uint32_t acc = 0;                                                    \
        for (size_t p = 0; p < k; ++p)                                       \
          acc += (uint32_t)((int32_t)a[i * k + p] * (int32_t)b[p * n + j]);  \
        if (accumulate) acc += (uint32_t)c[i * n + j];                       \
        c[i * n + j] = (int32_t)acc;                                         \
      }                                                                      \
    }                                                                        \
  }
ORC_INT_GEMM(orc_gemm_i8, uint8_t, int8_t)        /* mod.rs:601-651  */
ORC_INT_GEMM(orc_gemm_i8_signed, int8_t, int8_t)  /* mod.rs:414-448  */
ORC_INT_GEMM(orc_gemm_i16, int16_t, int16_t)      /* mod.rs:348-392  */

/* Documentation only: the default-AVX2 u8xi8 interior tile
 * (kernel_avx2_i8.rs:1-36,61-136) sums adjacent k pairs with vpmaddubsw,
 * which saturates the pair sum to i16 before widening.  The product library
 * implements the exact (VNNI / NEON / fallback) semantics; this function lets
 * a test show where the two differ. */
void orc_gemm_i8_avx2_saturating(size_t m, size_t n, size_t k,
                                 const uint8_t *a, const int8_t *b, int32_t *c,
                                 int accumulate) {
  if (m == 0 || n == 0) return;
  for (size_t i = 0; i < m; ++i)
    for (size_t j = 0; j < n; ++j) {
      uint32_t acc = 0;
      for (size_t p = 0; p < k; p += 2) {
        int32_t s = (int32_t)a[i * k + p] * (int32_t)b[p * n + j];
        if (p + 1 < k) s += (int32_t)a[i * k + p + 1] * (int32_t)b[(p + 1) * n + j];
        if (s > 32767) s = 32767;
        if (s < -32768) s = -32768;
        acc += (uint32_t)s;
      }
      if (accumulate) acc += (uint32_t)c[i * n + j];
      c[i * n + j] = (int32_t)acc;
    }
}

/* mod.rs:484-575: raw = gemm_i8; row sums of A and column sums of B in
 * wrapping i32; corrected = raw - azp*csb - bzp*rsa + k*azp*bzp (wrapping);
 * out = (scale_a*scale_b) * (float)corrected, written or added. */
void orc_quantized_matmul(size_t m, size_t n, size_t k, const uint8_t *a,
                          uint8_t a_zp, float scale_a, const int8_t *b,
                          int8_t b_zp, float scale_b, float *c,
                          int accumulate) {
  if (m == 0 || n == 0) return;
  if (k == 0) {
    if (!accumulate)
      for (size_t i = 0; i < m * n; ++i) c[i] = 0.0f;
    return;
  }
  int32_t *raw = (int32_t *)malloc(m * n * sizeof(int32_t));
  uint32_t *rsa = (uint32_t *)calloc(m, sizeof(uint32_t));
  uint32_t *csb = (uint32_t *)calloc(n, sizeof(uint32_t));
  orc_gemm_i8(m, n, k, a, b, raw, 0);
  for (size_t i = 0; i < m; ++i)
    for (size_t p = 0; p < k; ++p) rsa[i] += (uint32_t)a[i * k + p];
  for (size_t p = 0; p < k; ++p)
    for (size_t j = 0; j < n; ++j) csb[j] += (uint32_t)(int32_t)b[p * n + j];
  uint32_t azp = (uint32_t)(int32_t)a_zp, bzp = (uint32_t)(int32_t)b_zp;
  uint32_t kab = (uint32_t)k * azp * bzp;
  float scale = scale_a * scale_b;
  for (size_t i = 0; i < m; ++i)
    for (size_t j = 0; j < n; ++j) {
      uint32_t v = (uint32_t)raw[i * n + j] - azp * csb[j] - bzp * rsa[i] + kab;
      float out = scale * (float)(int32_t)v;
      if (accumulate) c[i * n + j] += out;
      else c[i * n + j] = out;
    }
  free(raw); free(rsa); free(csb);
}

/* ------------------------------------------------------------------ */
/* bf16 / f16 -> f32 (mod.rs:851-933): exact widening, then gemm_f32    */
/* with alpha = 1 and beta = accumulate ? 1 : 0.                        */
/* ------------------------------------------------------------------ */
float orc_bf16_to_f32(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}
float orc_f16_to_f32(uint16_t h) {
  uint32_t sign = (uint32_t)(h >> 15) << 31;
  uint32_t e = (h >> 10) & 0x1f, mant = h & 0x3ff;
  uint32_t u;
  if (e == 0) {
    if (mant == 0) u = sign;
    else { /* subnormal: value = mant * 2^-24, exact in f32 */
      float f = (float)mant * 0x1.0p-24f;
      memcpy(&u, &f, 4);
      u |= sign;
    }
  } else if (e == 31) {
    u = sign | 0x7f800000u | (mant << 13);
  } else {
    u = sign | ((e - 15 + 127) << 23) | (mant << 13);
  }
  float f;
  memcpy(&f, &u, 4);
  return f;
}
static void orc_half_gemm(size_t m, size_t n, size_t k, const uint16_t *a,
                          const uint16_t *b, float *c, int accumulate,
                          float (*widen)(uint16_t)) {
  if (m == 0 || n == 0) return;
  float *aw = (float *)malloc((m * k + 1) * sizeof(float));
  float *bw = (float *)malloc((k * n + 1) * sizeof(float));
  for (size_t i = 0; i < m * k; ++i) aw[i] = widen(a[i]);
  for (size_t i = 0; i < k * n; ++i) bw[i] = widen(b[i]);
  orc_gemm_f32(m, n, k, 1.0f, aw, bw, accumulate ? 1.0f : 0.0f, c);
  free(aw); free(bw);
}
void orc_gemm_bf16_f32(size_t m, size_t n, size_t k, const uint16_t *a,
                       const uint16_t *b, float *c, int accumulate) {
  orc_half_gemm(m, n, k, a, b, c, accumulate, orc_bf16_to_f32);
}
void orc_gemm_f16_f32(size_t m, size_t n, size_t k, const uint16_t *a,
                      const uint16_t *b, float *c, int accumulate) {
  orc_half_gemm(m, n, k, a, b, c, accumulate, orc_f16_to_f32);
}

/* OCP 8-bit floats (OCP "8-bit Floating Point Specification", v1.0, 2023).
 * E4M3: bias 7, no infinities, S.1111.111 is NaN, largest finite 448.
 * E5M2: bias 15, IEEE-style: exponent 31 is Inf (mantissa 0) or NaN.
 * Subnormals (exponent 0) are mantissa * 2^(1 - bias - mantissa bits). */
double orc_e4m3_to_f64(uint8_t x) {
  const int sign = x >> 7, e = (x >> 3) & 15, mant = x & 7;
  double v;
  if (e == 15 && mant == 7) return NAN;
  if (e == 0) v = ldexp((double)mant, -9);
  else v = ldexp(1.0 + mant / 8.0, e - 7);
  return sign ? -v : v;
}
double orc_e5m2_to_f64(uint8_t x) {
  const int sign = x >> 7, e = (x >> 2) & 31, mant = x & 3;
  double v;
  if (e == 31) v = mant ? NAN : INFINITY;
  else if (e == 0) v = ldexp((double)mant, -16);
  else v = ldexp(1.0 + mant / 4.0, e - 15);
  return sign ? -v : v;
}
/* C (+)= A B with A, B fp8 bytes (format 0 = E4M3, 1 = E5M2), summed in
 * double: every product of two fp8 values is exact in double, so the result
 * is the exact product to double rounding of the sums. */
void orc_gemm_fp8_f64(size_t m, size_t n, size_t k, const uint8_t *a, const uint8_t *b,
                      int format, double *c, int accumulate) {
  double lut[256];
  for (int i = 0; i < 256; ++i)
    lut[i] = format == 0 ? orc_e4m3_to_f64((uint8_t)i) : orc_e5m2_to_f64((uint8_t)i);
  for (size_t i = 0; i < m; ++i) {
    double *ci = c + i * n;
    if (!accumulate)
      for (size_t j = 0; j < n; ++j) ci[j] = 0.0;
    for (size_t p = 0; p < k; ++p) {
      const double av = lut[a[i * k + p]];
      const uint8_t *bp = b + p * n;
      for (size_t j = 0; j < n; ++j) ci[j] += av * lut[bp[j]];
    }
  }
}

/* ------------------------------------------------------------------ */
/* Double-double reference for error measurement                        */
/* ------------------------------------------------------------------ */

/* Accurate dot products with TwoProd (fma) + TwoSum (Knuth), i.e. the
 * compensated "Dot2" of Ogita, Rump & Oishi (2005): the result is as if
 * computed in twice the working precision and then rounded.  Computes
 * C_ref[r, c] for the selected rows and columns (hi and lo parts).  Rows are
 * split over up to min(nproc, 16) threads; each element's arithmetic is the
 * same sequential loop whatever the split, so the result does not depend on
 * the thread count. */
typedef struct {
  int cplx;
  size_t n, k, r0, r1, ncols;
  const double *a, *b;
  const int64_t *rows, *cols;
  double *out_hi, *out_lo;
} orc_dd_task;

/* One row at a time, all selected columns together: step p updates every
 * column's (s, e) pair, so B is read along its rows.  Each element still sees
 * the terms p = 0, 1, ..., k-1 in order with the same operations. */
static void orc_dd_rows_f64(const orc_dd_task *t) {
  double *s = (double *)malloc(2 * t->ncols * sizeof(double));
  double *e = s + t->ncols;
  for (size_t ri = t->r0; ri < t->r1; ++ri) {
    const double *ar = t->a + (size_t)t->rows[ri] * t->k;
    for (size_t ci = 0; ci < t->ncols; ++ci) s[ci] = e[ci] = 0.0;
    for (size_t p = 0; p < t->k; ++p) {
      const double x = ar[p];
      const double *brow = t->b + p * t->n;
      for (size_t ci = 0; ci < t->ncols; ++ci) {
        double y = brow[t->cols[ci]];
        double pr = x * y;
        double pe = fma(x, y, -pr);
        double u = s[ci] + pr;
        double z = u - s[ci];
        double se = (s[ci] - (u - z)) + (pr - z);
        s[ci] = u;
        e[ci] += se + pe;
      }
    }
    for (size_t ci = 0; ci < t->ncols; ++ci) {
      double hi = s[ci] + e[ci];
      double lo = e[ci] - (hi - s[ci]);
      t->out_hi[ri * t->ncols + ci] = hi;
      t->out_lo[ri * t->ncols + ci] = lo;
    }
  }
  free(s);
}

/* Complex f64: C = A.B with the real part sum(xr yr) + sum(-xi yi) and the
 * imaginary part sum(xr yi) + sum(xi yr), each one compensated sum (terms in
 * the order xr yr, -xi yi for the real part and xr yi, xi yr for the
 * imaginary part, p increasing). */
static void orc_dd_rows_c64(const orc_dd_task *t) {
  double *st = (double *)malloc(4 * t->ncols * sizeof(double));
  for (size_t ri = t->r0; ri < t->r1; ++ri) {
    const double *ar = t->a + 2 * (size_t)t->rows[ri] * t->k;
    for (size_t ci = 0; ci < 4 * t->ncols; ++ci) st[ci] = 0.0;
    for (size_t p = 0; p < t->k; ++p) {
      const double xr = ar[2 * p], xi = ar[2 * p + 1];
      const double *brow = t->b + 2 * p * t->n;
      for (size_t ci = 0; ci < t->ncols; ++ci) {
        const size_t j = (size_t)t->cols[ci];
        const double yr = brow[2 * j], yi = brow[2 * j + 1];
        const double terms[2][2][2] = {{{xr, yr}, {-xi, yi}}, {{xr, yi}, {xi, yr}}};
        double *sq = st + 4 * ci;  /* s_re, e_re, s_im, e_im */
        for (int q = 0; q < 2; ++q)
          for (int w = 0; w < 2; ++w) {
            double x = terms[q][w][0], y = terms[q][w][1];
            double pr = x * y, pe = fma(x, y, -pr);
            double u = sq[2 * q] + pr, z = u - sq[2 * q];
            double se = (sq[2 * q] - (u - z)) + (pr - z);
            sq[2 * q] = u;
            sq[2 * q + 1] += se + pe;
          }
      }
    }
    for (size_t ci = 0; ci < t->ncols; ++ci) {
      const double *sq = st + 4 * ci;
      t->out_hi[2 * (ri * t->ncols + ci)] = sq[0] + sq[1];
      t->out_hi[2 * (ri * t->ncols + ci) + 1] = sq[2] + sq[3];
    }
  }
  free(st);
}

static void *orc_dd_worker(void *arg) {
  orc_dd_task *t = (orc_dd_task *)arg;
  if (t->cplx) orc_dd_rows_c64(t);
  else orc_dd_rows_f64(t);
  return NULL;
}

static void orc_dd_run(orc_dd_task base, size_t nrows) {
  long np = sysconf(_SC_NPROCESSORS_ONLN);
  size_t workers = np > 16 ? 16 : (size_t)(np > 0 ? np : 1);
  if (workers > nrows) workers = nrows ? nrows : 1;
  orc_dd_task t[16];
  pthread_t th[16];
  size_t per = (nrows + workers - 1) / workers;
  size_t nt = 0;
  for (size_t r0 = 0; r0 < nrows; r0 += per, ++nt) {
    t[nt] = base;
    t[nt].r0 = r0;
    t[nt].r1 = r0 + per < nrows ? r0 + per : nrows;
    if (pthread_create(&th[nt], NULL, orc_dd_worker, &t[nt]) != 0) {
      orc_dd_worker(&t[nt]);
      th[nt] = 0;
    }
  }
  for (size_t i = 0; i < nt; ++i)
    if (th[i]) pthread_join(th[i], NULL);
}

void orc_dd_gemm_f64_sampled(size_t m, size_t n, size_t k, const double *a,
                             const double *b, size_t nrows,
                             const int64_t *rows, size_t ncols,
                             const int64_t *cols, double *out_hi,
                             double *out_lo) {
  (void)m;
  orc_dd_task t = {0, n, k, 0, 0, ncols, a, b, rows, cols, out_hi, out_lo};
  orc_dd_run(t, nrows);
}

/* Interleaved complex output: out[2 (r ncols + c)] = re, [.. + 1] = im. */
void orc_dd_gemm_c64_sampled(size_t m, size_t n, size_t k, const double *a,
                             const double *b, size_t nrows,
                             const int64_t *rows, size_t ncols,
                             const int64_t *cols, double *out) {
  (void)m;
  orc_dd_task t = {1, n, k, 0, 0, ncols, a, b, rows, cols, out, NULL};
  orc_dd_run(t, nrows);
}

/* ------------------------------------------------------------------ */
/* Timed CPU baseline: restatement of ferray's Goto driver              */
/* ------------------------------------------------------------------ */
/*
 * Blocking MC/KC/NC = 64/256/512 (outer.rs:66-68), B packed per (jc,pc)
 * into NR-wide strips, A per (jc,pc,ic) into MR-tall strips (pack.rs:28-131),
 * AVX2+FMA 4x8 micro-kernel for f64 and 4x16 for f32 (kernel_avx2.rs:53-216,
 * kernel_avx2_f32.rs:56), scalar edge tile (kernel_avx2.rs:251-291), and the
 * 2-D worker grid of pick_grid with min(nproc, 8) workers (outer.rs:142-155,
 * :233-308, :343-383).  Every element goes through the same fma chain as the
 * scalar formula above, so the result is bit-identical to orc_gemm_f64.
 */
#define D_MR 4
#define D_NR 8
#define S_MR 4
#define S_NR 16

typedef struct {
  size_t m, n, k;
  double alpha, beta;
  const double *a, *b;
  double *c;
  size_t lda, ldb, ldc;
} orc_d_job;

typedef struct {
  size_t m, n, k;
  float alpha, beta;
  const float *a, *b;
  float *c;
  size_t lda, ldb, ldc;
} orc_s_job;

static void orc_pick_grid(size_t m, size_t n, size_t k, size_t workers,
                          size_t mr, size_t *mw_out, size_t *nw_out) {
  /* outer.rs:343-383 */
  if (workers <= 1 || m == 0 || n == 0) { *mw_out = workers ? workers : 1; *nw_out = 1; return; }
  size_t kc_eff = k < ORC_KC ? (k ? k : 1) : ORC_KC;
  size_t per_worker_b = kc_eff * n;
  size_t mr_chunks = m / mr;
  if (mr_chunks >= workers && per_worker_b <= 128 * 1024) { *mw_out = workers; *nw_out = 1; return; }
  size_t mw_max = mr_chunks ? mr_chunks : 1;
  size_t best_m = 1, best_n = workers;
  double best = INFINITY;
  for (size_t mw = 1; mw <= mw_max && mw <= workers; ++mw) {
    if (workers % mw) continue;
    size_t nw = workers / mw;
    double score = fabs(log((double)m / (double)mw) - log((double)n / (double)nw));
    if (score < best || (score == best && mw > best_m)) { best = score; best_m = mw; best_n = nw; }
  }
  *mw_out = best_m; *nw_out = best_n;
}

static inline double orc_d_store(double cv, double acc, double beta) {
  if (beta == 0.0) return acc;
  if (beta == 1.0) return cv + acc;
  return fma(cv, beta, acc);
}

#if defined(__x86_64__)
__attribute__((target("avx2,fma")))
static void orc_dk_4x8(size_t kc, double alpha, const double *pa,
                       const double *pb, double beta, double *c, size_t ldc) {
  __m256d c00 = _mm256_setzero_pd(), c01 = _mm256_setzero_pd();
  __m256d c10 = _mm256_setzero_pd(), c11 = _mm256_setzero_pd();
  __m256d c20 = _mm256_setzero_pd(), c21 = _mm256_setzero_pd();
  __m256d c30 = _mm256_setzero_pd(), c31 = _mm256_setzero_pd();
  for (size_t p = 0; p < kc; ++p) {
    __m256d b0 = _mm256_loadu_pd(pb), b1 = _mm256_loadu_pd(pb + 4);
    __m256d a0 = _mm256_broadcast_sd(pa), a1 = _mm256_broadcast_sd(pa + 1);
    __m256d a2 = _mm256_broadcast_sd(pa + 2), a3 = _mm256_broadcast_sd(pa + 3);
    c00 = _mm256_fmadd_pd(a0, b0, c00); c01 = _mm256_fmadd_pd(a0, b1, c01);
    c10 = _mm256_fmadd_pd(a1, b0, c10); c11 = _mm256_fmadd_pd(a1, b1, c11);
    c20 = _mm256_fmadd_pd(a2, b0, c20); c21 = _mm256_fmadd_pd(a2, b1, c21);
    c30 = _mm256_fmadd_pd(a3, b0, c30); c31 = _mm256_fmadd_pd(a3, b1, c31);
    pa += D_MR; pb += D_NR;
  }
  __m256d acc[8] = {c00, c01, c10, c11, c20, c21, c30, c31};
  if (alpha != 1.0) {
    __m256d av = _mm256_set1_pd(alpha);
    for (int q = 0; q < 8; ++q) acc[q] = _mm256_mul_pd(acc[q], av);
  }
  for (int r = 0; r < 4; ++r) {
    double *p = c + r * ldc;
    for (int h = 0; h < 2; ++h) {
      __m256d v = acc[2 * r + h];
      if (beta == 0.0) _mm256_storeu_pd(p + 4 * h, v);
      else if (beta == 1.0) _mm256_storeu_pd(p + 4 * h, _mm256_add_pd(_mm256_loadu_pd(p + 4 * h), v));
      else _mm256_storeu_pd(p + 4 * h, _mm256_fmadd_pd(_mm256_loadu_pd(p + 4 * h), _mm256_set1_pd(beta), v));
    }
  }
}
#endif

static void orc_dk_edge(size_t mt, size_t nt, size_t kc, double alpha,
                        const double *pa, const double *pb, double beta,
                        double *c, size_t ldc) {
  double tile[D_MR][D_NR] = {{0}};
  for (size_t p = 0; p < kc; ++p) {
    for (size_t i = 0; i < mt; ++i)
      for (size_t j = 0; j < nt; ++j) tile[i][j] = fma(pa[i], pb[j], tile[i][j]);
    pa += D_MR; pb += D_NR;
  }
  for (size_t i = 0; i < mt; ++i)
    for (size_t j = 0; j < nt; ++j) {
      double v = alpha != 1.0 ? alpha * tile[i][j] : tile[i][j];
      c[i * ldc + j] = orc_d_store(c[i * ldc + j], v, beta);
    }
}

static void orc_d_seq(const orc_d_job *jb, int use_avx2) {
  size_t m = jb->m, n = jb->n, k = jb->k;
  if (!m || !n || !k) return;
  double *bp = (double *)aligned_alloc(64, ((ORC_NC + D_NR) * ORC_KC) * sizeof(double));
  double *ap = (double *)aligned_alloc(64, ((ORC_MC + D_MR) * ORC_KC) * sizeof(double));
  for (size_t jc = 0; jc < n; jc += ORC_NC) {
    size_t nc = n - jc < ORC_NC ? n - jc : ORC_NC;
    size_t nstr = (nc + D_NR - 1) / D_NR;
    for (size_t pc = 0; pc < k; pc += ORC_KC) {
      size_t kc = k - pc < ORC_KC ? k - pc : ORC_KC;
      double eb = pc == 0 ? jb->beta : 1.0;
      for (size_t s = 0; s < nstr; ++s) /* pack.rs:68-131 */
        for (size_t p = 0; p < kc; ++p)
          for (size_t j = 0; j < D_NR; ++j) {
            size_t col = s * D_NR + j;
            bp[s * D_NR * kc + p * D_NR + j] = col < nc ? jb->b[(pc + p) * jb->ldb + jc + col] : 0.0;
          }
      for (size_t ic = 0; ic < m; ic += ORC_MC) {
        size_t mc = m - ic < ORC_MC ? m - ic : ORC_MC;
        size_t mstr = (mc + D_MR - 1) / D_MR;
        for (size_t s = 0; s < mstr; ++s) /* pack.rs:28-58 */
          for (size_t p = 0; p < kc; ++p)
            for (size_t i = 0; i < D_MR; ++i) {
              size_t row = s * D_MR + i;
              ap[s * D_MR * kc + p * D_MR + i] = row < mc ? jb->a[(ic + row) * jb->lda + pc + p] : 0.0;
            }
        for (size_t jr = 0; jr < nstr; ++jr) {
          size_t nt = nc - jr * D_NR < D_NR ? nc - jr * D_NR : D_NR;
          for (size_t ir = 0; ir < mstr; ++ir) {
            size_t mt = mc - ir * D_MR < D_MR ? mc - ir * D_MR : D_MR;
            double *ct = jb->c + (ic + ir * D_MR) * jb->ldc + jc + jr * D_NR;
            const double *pa = ap + ir * D_MR * kc, *pb = bp + jr * D_NR * kc;
#if defined(__x86_64__)
            if (use_avx2 && mt == D_MR && nt == D_NR) {
              orc_dk_4x8(kc, jb->alpha, pa, pb, eb, ct, jb->ldc);
              continue;
            }
#endif
            orc_dk_edge(mt, nt, kc, jb->alpha, pa, pb, eb, ct, jb->ldc);
          }
        }
      }
    }
  }
  free(bp); free(ap);
}

/* f32 4x16 */
#if defined(__x86_64__)
__attribute__((target("avx2,fma")))
static void orc_sk_4x16(size_t kc, float alpha, const float *pa, const float *pb,
                        float beta, float *c, size_t ldc) {
  __m256 acc[8];
  for (int q = 0; q < 8; ++q) acc[q] = _mm256_setzero_ps();
  for (size_t p = 0; p < kc; ++p) {
    __m256 b0 = _mm256_loadu_ps(pb), b1 = _mm256_loadu_ps(pb + 8);
    for (int r = 0; r < 4; ++r) {
      __m256 ar = _mm256_broadcast_ss(pa + r);
      acc[2 * r] = _mm256_fmadd_ps(ar, b0, acc[2 * r]);
      acc[2 * r + 1] = _mm256_fmadd_ps(ar, b1, acc[2 * r + 1]);
    }
    pa += S_MR; pb += S_NR;
  }
  if (alpha != 1.0f) {
    __m256 av = _mm256_set1_ps(alpha);
    for (int q = 0; q < 8; ++q) acc[q] = _mm256_mul_ps(acc[q], av);
  }
  for (int r = 0; r < 4; ++r) {
    float *p = c + r * ldc;
    for (int h = 0; h < 2; ++h) {
      __m256 v = acc[2 * r + h];
      if (beta == 0.0f) _mm256_storeu_ps(p + 8 * h, v);
      else if (beta == 1.0f) _mm256_storeu_ps(p + 8 * h, _mm256_add_ps(_mm256_loadu_ps(p + 8 * h), v));
      else _mm256_storeu_ps(p + 8 * h, _mm256_fmadd_ps(_mm256_loadu_ps(p + 8 * h), _mm256_set1_ps(beta), v));
    }
  }
}
#endif

static void orc_s_seq(const orc_s_job *jb, int use_avx2) {
  size_t m = jb->m, n = jb->n, k = jb->k;
  if (!m || !n || !k) return;
  float *bp = (float *)aligned_alloc(64, ((ORC_NC + S_NR) * ORC_KC) * sizeof(float));
  float *ap = (float *)aligned_alloc(64, ((ORC_MC + S_MR) * ORC_KC) * sizeof(float));
  for (size_t jc = 0; jc < n; jc += ORC_NC) {
    size_t nc = n - jc < ORC_NC ? n - jc : ORC_NC;
    size_t nstr = (nc + S_NR - 1) / S_NR;
    for (size_t pc = 0; pc < k; pc += ORC_KC) {
      size_t kc = k - pc < ORC_KC ? k - pc : ORC_KC;
      float eb = pc == 0 ? jb->beta : 1.0f;
      for (size_t s = 0; s < nstr; ++s)
        for (size_t p = 0; p < kc; ++p)
          for (size_t j = 0; j < S_NR; ++j) {
            size_t col = s * S_NR + j;
            bp[s * S_NR * kc + p * S_NR + j] = col < nc ? jb->b[(pc + p) * jb->ldb + jc + col] : 0.0f;
          }
      for (size_t ic = 0; ic < m; ic += ORC_MC) {
        size_t mc = m - ic < ORC_MC ? m - ic : ORC_MC;
        size_t mstr = (mc + S_MR - 1) / S_MR;
        for (size_t s = 0; s < mstr; ++s)
          for (size_t p = 0; p < kc; ++p)
            for (size_t i = 0; i < S_MR; ++i) {
              size_t row = s * S_MR + i;
              ap[s * S_MR * kc + p * S_MR + i] = row < mc ? jb->a[(ic + row) * jb->lda + pc + p] : 0.0f;
            }
        for (size_t jr = 0; jr < nstr; ++jr) {
          size_t nt = nc - jr * S_NR < S_NR ? nc - jr * S_NR : S_NR;
          for (size_t ir = 0; ir < mstr; ++ir) {
            size_t mt = mc - ir * S_MR < S_MR ? mc - ir * S_MR : S_MR;
            float *ct = jb->c + (ic + ir * S_MR) * jb->ldc + jc + jr * S_NR;
            const float *pa = ap + ir * S_MR * kc, *pb = bp + jr * S_NR * kc;
#if defined(__x86_64__)
            if (use_avx2 && mt == S_MR && nt == S_NR) {
              orc_sk_4x16(kc, jb->alpha, pa, pb, eb, ct, jb->ldc);
              continue;
            }
#endif
            float tile[S_MR][S_NR] = {{0}};
            for (size_t p = 0; p < kc; ++p)
              for (size_t i = 0; i < mt; ++i)
                for (size_t j = 0; j < nt; ++j)
                  tile[i][j] = fmaf(pa[p * S_MR + i], pb[p * S_NR + j], tile[i][j]);
            for (size_t i = 0; i < mt; ++i)
              for (size_t j = 0; j < nt; ++j) {
                float v = jb->alpha != 1.0f ? jb->alpha * tile[i][j] : tile[i][j];
                float cv = ct[i * jb->ldc + j];
                ct[i * jb->ldc + j] = eb == 0.0f ? v : (eb == 1.0f ? cv + v : fmaf(cv, eb, v));
              }
          }
        }
      }
    }
  }
  free(bp); free(ap);
}

typedef struct { int is_f32; int use_avx2; orc_d_job d; orc_s_job s; } orc_task;
static void *orc_worker(void *arg) {
  orc_task *t = (orc_task *)arg;
  if (t->is_f32) orc_s_seq(&t->s, t->use_avx2);
  else orc_d_seq(&t->d, t->use_avx2);
  return NULL;
}

int orc_cpu_has_avx2_fma(void) {
#if defined(__x86_64__)
  __builtin_cpu_init();
  return __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma");
#else
  return 0;
#endif
}

/* workers <= 0 selects min(nproc, 8) (outer.rs:145-148). Returns the
 * worker count used. */
static int orc_goto_run(int is_f32, size_t m, size_t n, size_t k, double alpha,
                        const void *a, const void *b, double beta, void *c,
                        int workers) {
  if (workers <= 0) {
    long np = sysconf(_SC_NPROCESSORS_ONLN);
    workers = np > 8 ? 8 : (int)(np > 0 ? np : 4);
  }
  int use_avx2 = orc_cpu_has_avx2_fma();
  size_t mr = is_f32 ? S_MR : D_MR, nr = is_f32 ? S_NR : D_NR;
  if (m == 0 || n == 0) return workers;
  if (k == 0) {
    if (is_f32) orc_gemm_f32(m, n, 0, (float)alpha, NULL, NULL, (float)beta, (float *)c);
    else orc_gemm_f64(m, n, 0, alpha, NULL, NULL, beta, (double *)c);
    return workers;
  }
  size_t mw = 1, nw = 1;
  size_t mx = m > n ? m : n;
  mx = mx > k ? mx : k;
  if (mx >= 256 && m >= mr * 2) orc_pick_grid(m, n, k, (size_t)workers, mr, &mw, &nw);
  size_t m_per = (m + mw - 1) / mw; m_per = ((m_per + mr - 1) / mr) * mr; if (m_per < mr) m_per = mr;
  size_t n_per = (n + nw - 1) / nw; n_per = ((n_per + nr - 1) / nr) * nr; if (n_per < nr) n_per = nr;
  size_t cap = ((m + m_per - 1) / m_per) * ((n + n_per - 1) / n_per);
  orc_task *tasks = (orc_task *)calloc(cap, sizeof(orc_task));
  pthread_t *th = (pthread_t *)calloc(cap, sizeof(pthread_t));
  size_t nt = 0;
  for (size_t ms = 0; ms < m; ms += m_per)
    for (size_t ns = 0; ns < n; ns += n_per) {
      orc_task *t = &tasks[nt];
      t->is_f32 = is_f32; t->use_avx2 = use_avx2;
      size_t msz = m - ms < m_per ? m - ms : m_per, nsz = n - ns < n_per ? n - ns : n_per;
      if (is_f32) {
        orc_s_job j = {msz, nsz, k, (float)alpha, (float)beta, (const float *)a + ms * k,
                       (const float *)b + ns, (float *)c + ms * n + ns, k, n, n};
        t->s = j;
      } else {
        orc_d_job j = {msz, nsz, k, alpha, beta, (const double *)a + ms * k,
                       (const double *)b + ns, (double *)c + ms * n + ns, k, n, n};
        t->d = j;
      }
      ++nt;
    }
  for (size_t i = 0; i < nt; ++i) pthread_create(&th[i], NULL, orc_worker, &tasks[i]);
  for (size_t i = 0; i < nt; ++i) pthread_join(th[i], NULL);
  free(tasks); free(th);
  return workers;
}

int orc_gemm_f64_goto(size_t m, size_t n, size_t k, double alpha,
                      const double *a, const double *b, double beta, double *c,
                      int workers) {
  return orc_goto_run(0, m, n, k, alpha, a, b, beta, c, workers);
}
int orc_gemm_f32_goto(size_t m, size_t n, size_t k, float alpha,
                      const float *a, const float *b, float beta, float *c,
                      int workers) {
  return orc_goto_run(1, m, n, k, alpha, a, b, beta, c, workers);
}

/* Threaded exact integer GEMM (same closed form as orc_gemm_i8_signed,
 * rows split over threads) so that sampled checks at large sizes and the
 * integer CPU baseline finish quickly. */
typedef struct { size_t m0, m1, n, k; const int8_t *a, *b; int32_t *c; } orc_i8_task;
static void *orc_i8_worker(void *arg) {
  orc_i8_task *t = (orc_i8_task *)arg;
  int32_t *row = (int32_t *)malloc(t->n * sizeof(int32_t));
  for (size_t i = t->m0; i < t->m1; ++i) {
    memset(row, 0, t->n * sizeof(int32_t));
    for (size_t p = 0; p < t->k; ++p) {
      int32_t av = t->a[i * t->k + p];
      const int8_t *br = t->b + p * t->n;
      for (size_t j = 0; j < t->n; ++j) row[j] = (int32_t)((uint32_t)row[j] + (uint32_t)(av * (int32_t)br[j]));
    }
    memcpy(t->c + i * t->n, row, t->n * sizeof(int32_t));
  }
  free(row);
  return NULL;
}
int orc_gemm_i8_signed_threaded(size_t m, size_t n, size_t k, const int8_t *a,
                                const int8_t *b, int32_t *c, int workers) {
  if (workers <= 0) {
    long np = sysconf(_SC_NPROCESSORS_ONLN);
    workers = np > 8 ? 8 : (int)(np > 0 ? np : 4);
  }
  orc_i8_task t[64];
  pthread_t th[64];
  if (workers > 64) workers = 64;
  size_t per = (m + workers - 1) / workers;
  int nt = 0;
  for (size_t m0 = 0; m0 < m; m0 += per, ++nt) {
    t[nt].m0 = m0; t[nt].m1 = m0 + per < m ? m0 + per : m; t[nt].n = n; t[nt].k = k;
    t[nt].a = a; t[nt].b = b; t[nt].c = c;
    pthread_create(&th[nt], NULL, orc_i8_worker, &t[nt]);
  }
  for (int i = 0; i < nt; ++i) pthread_join(th[i], NULL);
  return workers;
}

/* ------------------------------------------------------------------ */
/* Triangular solve (trsm.rs)                                           */
/* ------------------------------------------------------------------ */
/* Restates trsm.rs:46-140 (entry points: scale B by alpha when alpha != 1,
 * then recurse), trsm.rs:165-200 / :264-301 (block recursion with NB = 32,
 * m1 = min(round_up(m/2, 32), m - 32)), trsm.rs:205-225 (scalar trailing
 * update B2 -= A21 X1, loops i, p, j) and trsm.rs:227-262 / :304-330 (the
 * unblocked substitution: b[i,j] -= a[i,p] * b[p,j], then *= 1 / a[i,i]).
 * Built with -ffp-contract=off, every step rounds like the reference's
 * (Rust does not contract a*b - c). */
#define ORC_TRSM_NB 32

#define ORC_TRSM(T, SUF)                                                                       \
  static void orc_trsm_scale_##SUF(size_t m, size_t n, T alpha, T *b, size_t ldb) {            \
    for (size_t i = 0; i < m; ++i)                                                             \
      for (size_t j = 0; j < n; ++j) b[i * ldb + j] *= alpha;                                  \
  }                                                                                            \
  static void orc_trsm_update_##SUF(size_t m2, size_t n, size_t k, const T *l, size_t lda,     \
                                    const T *x, size_t ldb, T *b2) {                           \
    for (size_t i = 0; i < m2; ++i)                                                            \
      for (size_t p = 0; p < k; ++p) {                                                         \
        const T lip = l[i * lda + p];                                                          \
        for (size_t j = 0; j < n; ++j) b2[i * ldb + j] -= lip * x[p * ldb + j];                \
      }                                                                                        \
  }                                                                                            \
  static void orc_trsm_lln_unb_##SUF(size_t m, size_t n, const T *a, size_t lda, T *b,         \
                                     size_t ldb, int unit) {                                   \
    for (size_t i = 0; i < m; ++i) {                                                           \
      for (size_t p = 0; p < i; ++p) {                                                         \
        const T aip = a[i * lda + p];                                                          \
        for (size_t j = 0; j < n; ++j) b[i * ldb + j] -= aip * b[p * ldb + j];                 \
      }                                                                                        \
      if (!unit) {                                                                             \
        const T inv = (T)1 / a[i * lda + i];                                                   \
        for (size_t j = 0; j < n; ++j) b[i * ldb + j] *= inv;                                  \
      }                                                                                        \
    }                                                                                          \
  }                                                                                            \
  static void orc_trsm_lun_unb_##SUF(size_t m, size_t n, const T *a, size_t lda, T *b,         \
                                     size_t ldb, int unit) {                                   \
    for (size_t ii = m; ii-- > 0;) {                                                           \
      for (size_t p = ii + 1; p < m; ++p) {                                                    \
        const T aip = a[ii * lda + p];                                                         \
        for (size_t j = 0; j < n; ++j) b[ii * ldb + j] -= aip * b[p * ldb + j];                \
      }                                                                                        \
      if (!unit) {                                                                             \
        const T inv = (T)1 / a[ii * lda + ii];                                                 \
        for (size_t j = 0; j < n; ++j) b[ii * ldb + j] *= inv;                                 \
      }                                                                                        \
    }                                                                                          \
  }                                                                                            \
  static size_t orc_trsm_split_##SUF(size_t m) {                                               \
    size_t h = (m / 2 + ORC_TRSM_NB - 1) / ORC_TRSM_NB * ORC_TRSM_NB;                          \
    return h < m - ORC_TRSM_NB ? h : m - ORC_TRSM_NB;                                          \
  }                                                                                            \
  static void orc_trsm_lln_rec_##SUF(size_t m, size_t n, const T *a, size_t lda, T *b,         \
                                     size_t ldb, int unit) {                                   \
    if (m <= ORC_TRSM_NB) {                                                                    \
      orc_trsm_lln_unb_##SUF(m, n, a, lda, b, ldb, unit);                                      \
      return;                                                                                  \
    }                                                                                          \
    const size_t m1 = orc_trsm_split_##SUF(m), m2 = m - m1;                                    \
    orc_trsm_lln_rec_##SUF(m1, n, a, lda, b, ldb, unit);                                       \
    orc_trsm_update_##SUF(m2, n, m1, a + m1 * lda, lda, b, ldb, b + m1 * ldb);                 \
    orc_trsm_lln_rec_##SUF(m2, n, a + m1 * lda + m1, lda, b + m1 * ldb, ldb, unit);            \
  }                                                                                            \
  static void orc_trsm_lun_rec_##SUF(size_t m, size_t n, const T *a, size_t lda, T *b,         \
                                     size_t ldb, int unit) {                                   \
    if (m <= ORC_TRSM_NB) {                                                                    \
      orc_trsm_lun_unb_##SUF(m, n, a, lda, b, ldb, unit);                                      \
      return;                                                                                  \
    }                                                                                          \
    const size_t m1 = orc_trsm_split_##SUF(m), m2 = m - m1;                                    \
    orc_trsm_lun_rec_##SUF(m2, n, a + m1 * lda + m1, lda, b + m1 * ldb, ldb, unit);            \
    orc_trsm_update_##SUF(m1, n, m2, a + m1, lda, b + m1 * ldb, ldb, b);                       \
    orc_trsm_lun_rec_##SUF(m1, n, a, lda, b, ldb, unit);                                       \
  }                                                                                            \
  void orc_trsm_left_lower_##SUF(size_t m, size_t n, T alpha, const T *a, size_t lda, T *b,    \
                                 size_t ldb, int unit) {                                       \
    if (m == 0 || n == 0) return;                                                              \
    if (alpha != (T)1) orc_trsm_scale_##SUF(m, n, alpha, b, ldb);                              \
    orc_trsm_lln_rec_##SUF(m, n, a, lda, b, ldb, unit);                                        \
  }                                                                                            \
  void orc_trsm_left_upper_##SUF(size_t m, size_t n, T alpha, const T *a, size_t lda, T *b,    \
                                 size_t ldb, int unit) {                                       \
    if (m == 0 || n == 0) return;                                                              \
    if (alpha != (T)1) orc_trsm_scale_##SUF(m, n, alpha, b, ldb);                              \
    orc_trsm_lun_rec_##SUF(m, n, a, lda, b, ldb, unit);                                        \
  }

ORC_TRSM(double, f64)
ORC_TRSM(float, f32)
