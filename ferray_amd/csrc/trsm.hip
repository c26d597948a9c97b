// Triangular solve op(A) X = alpha B (left side, no transpose, lower or upper,
// row-major with row strides lda, ldb; X overwrites B), the GEMM consumer of
// ferray-linalg/src/trsm.rs.
//
// Same block recursion as the reference (trsm.rs:165-200, :264-301): split
// the rows at m1 = min(round_up(m/2, NB), m - NB), solve one half, update the
// other with a GEMM, solve it.  Here the trailing update runs on the MFMA
// DGEMM/SGEMM (C = -1 * A21 X1 + 1 * B2, on the strided blocks where they
// lie, through the row-strided GEMM), and the leaf (m <= NB = 64 rows) is a
// kernel in which each thread solves one column of B by substitution with the
// triangular block in LDS.  The leaf applies the reference's scalar
// operation order (row i: subtract a_ip x_p for p in order, then multiply by
// 1 / a_ii), so a solve that fits one leaf gives the reference's bits.
#include <algorithm>

#include "common.hpp"

namespace fg {

constexpr int TRSM_NB = 64;

template <typename T>
__global__ void trsm_scale_kernel(int m, int n, T alpha, T* __restrict__ b, size_t ldb) {
  const size_t total = (size_t)m * n;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t r = i / n, c = i % n;
    b[r * ldb + c] *= alpha;
  }
}

// Lower leaf with LPC threads per column: thread q of column j owns rows
// q, q + LPC, ..., and the solve runs right-looking, step p finishing x_p on
// its owner, broadcasting it to the others and subtracting L_ip x_p from
// every row i > p.  Each row still receives its updates in increasing p,
// each as a rounded product then a rounded difference, then the scaling by
// 1 / L_ii: the same operations in the same order as the one-thread-per-
// column substitution below, so the same bits, with LPC times the threads
// on the chip.  Rows past m are zero (their L rows and inverse diagonal too)
// and are never written back.
template <typename T, int LPC>
__global__ __launch_bounds__(256) void trsm_leaf_lower_kernel(int m, int n,
                                                               const T* __restrict__ a,
                                                               size_t lda, T* __restrict__ b,
                                                               size_t ldb, int unit_diag) {
  constexpr int NB = TRSM_NB, R = NB / LPC;
  __shared__ T As[NB][NB + 1];
  __shared__ T inv_diag[NB];
  for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) {
    const int r = idx / NB, c = idx % NB;
    As[r][c] = (r < m && c < m) ? a[(size_t)r * lda + c] : T(0);
  }
  __syncthreads();
  if (threadIdx.x < NB) inv_diag[threadIdx.x] = threadIdx.x < m ? T(1) / As[threadIdx.x][threadIdx.x] : T(0);
  __syncthreads();
  const int j = (int)((blockIdx.x * blockDim.x + threadIdx.x) / LPC);
  const int q = threadIdx.x % LPC, lane = threadIdx.x & 63;
  const bool live = j < n;
  T v[R];
#pragma unroll
  for (int t = 0; t < R; ++t) {
    const int i = q + LPC * t;
    v[t] = (live && i < m) ? b[(size_t)i * ldb + j] : T(0);
  }
#pragma unroll
  for (int p = 0; p < NB; ++p) {
    const int owner = p % LPC, tp = p / LPC;
    T xp = v[tp];
    if (!unit_diag) xp *= inv_diag[p];
    if (q == owner) v[tp] = xp;
    xp = __shfl(xp, (lane & ~(LPC - 1)) | owner, 64);
#pragma unroll
    for (int t = tp; t < R; ++t) {
      const int i = q + LPC * t;
      if (i > p) v[t] -= As[i][p] * xp;
    }
  }
#pragma unroll
  for (int t = 0; t < R; ++t) {
    const int i = q + LPC * t;
    if (live && i < m) b[(size_t)i * ldb + j] = v[t];
  }
}

// Leaf solve of an m x m (m <= NB) triangle; thread j owns column j of B.
template <typename T, bool LOWER>
__global__ __launch_bounds__(256) void trsm_leaf_kernel(int m, int n, const T* __restrict__ a,
                                                        size_t lda, T* __restrict__ b, size_t ldb,
                                                        int unit_diag) {
  constexpr int NB = TRSM_NB;
  __shared__ T As[NB][NB + 1];
  __shared__ T inv_diag[NB];
  for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) {
    const int r = idx / NB, c = idx % NB;
    As[r][c] = (r < m && c < m) ? a[(size_t)r * lda + c] : T(0);
  }
  __syncthreads();
  if (threadIdx.x < NB) inv_diag[threadIdx.x] = threadIdx.x < m ? T(1) / As[threadIdx.x][threadIdx.x] : T(0);
  __syncthreads();
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  T x[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) x[i] = i < m ? b[(size_t)i * ldb + j] : T(0);
  if (LOWER) {
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      if (i < m) {
        T v = x[i];
#pragma unroll
        for (int p = 0; p < i; ++p) v -= As[i][p] * x[p];
        if (!unit_diag) v *= inv_diag[i];
        x[i] = v;
      }
    }
  } else {
#pragma unroll
    for (int i = NB - 1; i >= 0; --i) {
      if (i < m) {
        T v = x[i];
#pragma unroll
        for (int p = i + 1; p < NB; ++p)
          if (p < m) v -= As[i][p] * x[p];
        if (!unit_diag) v *= inv_diag[i];
        x[i] = v;
      }
    }
  }
#pragma unroll
  for (int i = 0; i < NB; ++i)
    if (i < m) b[(size_t)i * ldb + j] = x[i];
}

// Strided GEMM on the device (the drivers' row-strided form).
template <typename T>
int gemm_strided(size_t m, size_t n, size_t k, T alpha, const T* a, size_t lda, const T* b,
                 size_t ldb, T beta, T* c, size_t ldc, hipStream_t st);
template <>
int gemm_strided<double>(size_t m, size_t n, size_t k, double alpha, const double* a, size_t lda,
                         const double* b, size_t ldb, double beta, double* c, size_t ldc,
                         hipStream_t st) {
  return ferray_gemm_f64_strided_dev(m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, st);
}
template <>
int gemm_strided<float>(size_t m, size_t n, size_t k, float alpha, const float* a, size_t lda,
                        const float* b, size_t ldb, float beta, float* c, size_t ldc,
                        hipStream_t st) {
  return ferray_gemm_f32_strided_dev(m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, st);
}

template <typename T>
struct TrsmCtx {
  int n;
  size_t lda, ldb;
  bool upper, unit;
  hipStream_t st;
};

static size_t trsm_split(size_t m) {
  const size_t half = (m / 2 + TRSM_NB - 1) / TRSM_NB * TRSM_NB;
  return std::min(half, m - TRSM_NB);
}

// B_t -= A_ts X_s in place (trsm.rs:205-214): A_ts (rt x rs) at a (stride
// lda), X_s (rs x n) and B_t (rt x n) rows of B (stride ldb), read and
// written where they lie by the strided GEMM.
template <typename T>
static int trsm_update(const TrsmCtx<T>& c, size_t rt, size_t rs, const T* a, const T* xs, T* bt) {
  return gemm_strided<T>(rt, (size_t)c.n, rs, T(-1), a, c.lda, xs, c.ldb, T(1), bt, c.ldb, c.st);
}

template <typename T>
static int trsm_rec(const TrsmCtx<T>& c, size_t m, const T* a, T* b) {
  if (m <= (size_t)TRSM_NB) {
    const unsigned blocks = (unsigned)cdiv((size_t)c.n, 256);
    if (c.upper)
      hipLaunchKernelGGL((trsm_leaf_kernel<T, false>), dim3(blocks), dim3(256), 0, c.st, (int)m,
                         c.n, a, c.lda, b, c.ldb, (int)c.unit);
    else if ((size_t)c.n * 4 < (size_t)256 * 256)  // fewer than 256 blocks at 4 per column
      hipLaunchKernelGGL((trsm_leaf_lower_kernel<T, 8>), dim3((unsigned)cdiv((size_t)c.n * 8, 256)),
                         dim3(256), 0, c.st, (int)m, c.n, a, c.lda, b, c.ldb, (int)c.unit);
    else
      hipLaunchKernelGGL((trsm_leaf_lower_kernel<T, 4>), dim3((unsigned)cdiv((size_t)c.n * 4, 256)),
                         dim3(256), 0, c.st, (int)m, c.n, a, c.lda, b, c.ldb, (int)c.unit);
    FG_CHECK(hipGetLastError());
    return FERRAY_OK;
  }
  const size_t m1 = trsm_split(m), m2 = m - m1;
  int s;
  if (!c.upper) {
    // L11 X1 = B1; B2 -= L21 X1; L22 X2 = B2
    if ((s = trsm_rec(c, m1, a, b))) return s;
    if ((s = trsm_update(c, m2, m1, a + m1 * c.lda, b, b + m1 * c.ldb))) return s;
    return trsm_rec(c, m2, a + m1 * c.lda + m1, b + m1 * c.ldb);
  }
  // U22 X2 = B2; B1 -= U12 X2; U11 X1 = B1
  if ((s = trsm_rec(c, m2, a + m1 * c.lda + m1, b + m1 * c.ldb))) return s;
  if ((s = trsm_update(c, m1, m2, a + m1, b + m1 * c.ldb, b))) return s;
  return trsm_rec(c, m1, a, b);
}

template <typename T>
static int trsm_dev(bool upper, size_t m, size_t n, T alpha, const T* a, size_t lda, T* b,
                    size_t ldb, bool unit_diag, hipStream_t st) {
  if (m == 0 || n == 0) return FERRAY_OK;
  if (!a || !b || lda < m || ldb < n) return FERRAY_ERR_INVALID;
  if (m > (1u << 30) || n > (1u << 30)) return FERRAY_ERR_INVALID;
  if (alpha != T(1)) {
    const size_t blocks = std::min<size_t>(cdiv(m * n, 256), 65536);
    hipLaunchKernelGGL((trsm_scale_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, st, (int)m,
                       (int)n, alpha, b, ldb);
    FG_CHECK(hipGetLastError());
  }
  TrsmCtx<T> c{(int)n, lda, ldb, upper, unit_diag, st};
  return trsm_rec(c, m, a, b);
}

}  // namespace fg

using namespace fg;

extern "C" {

int ferray_trsm_dev(int dtype, int upper, size_t m, size_t n, double alpha, const void* a,
                    size_t lda, void* b, size_t ldb, bool unit_diag, void* stream) {
  int s = ensure_device();
  if (s) return s;
  hipStream_t st = resolve_stream(stream);
  if (dtype == FERRAY_DTYPE_F64)
    return trsm_dev<double>(upper != 0, m, n, alpha, static_cast<const double*>(a), lda,
                            static_cast<double*>(b), ldb, unit_diag, st);
  if (dtype == FERRAY_DTYPE_F32)
    return trsm_dev<float>(upper != 0, m, n, (float)alpha, static_cast<const float*>(a), lda,
                           static_cast<float*>(b), ldb, unit_diag, st);
  return FERRAY_ERR_INVALID;
}

}  // extern "C"
