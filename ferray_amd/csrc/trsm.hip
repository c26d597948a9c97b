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
// Diagonal blocks of 65 to 256 rows, where the recursion would otherwise
// issue up to four leaves and three small GEMMs, are solved by one kernel
// (trsm_block_kernel below) that multiplies by the inverses of the 64 x 64
// diagonal blocks, computed once per call into the stream's workspace: a
// blocked solve whose leaves are products on the matrix cores rather than
// substitutions, so its results meet the same error bounds as the
// recursion's GEMM updates rather than repeating the reference's bits.
#include <algorithm>
#include <utility>

#include "common.hpp"
#include "real_tile.hpp"

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

// Inverses of the 64 x 64 diagonal blocks of a triangular matrix (the last
// one may be smaller, padded with zeros): block d at inv + 4096 d, row-major.
// Thread q of column j owns rows q, q + 4, ... of column j of the inverse and
// solves A_dd X = I by substitution (lower: right-looking top down, upper:
// bottom up), x_p passed among the 4 threads of a column by a shuffle.
template <typename T, bool LOWER>
__global__ __launch_bounds__(256) void trsm_diag_inv_kernel(int m, const T* __restrict__ a,
                                                            size_t lda, T* __restrict__ inv,
                                                            int unit_diag) {
  constexpr int NB = TRSM_NB, LPC = 4, R = NB / LPC;
  __shared__ T As[NB][NB + 1];
  __shared__ T inv_diag[NB];
  const int d = blockIdx.x, o = d * NB, h = min(NB, m - o);
  for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) {
    const int r = idx / NB, c = idx % NB;
    const T x = a[(size_t)(o + min(r, h - 1)) * lda + o + min(c, h - 1)];
    As[r][c] = (r < h && c < h) ? x : T(0);
  }
  __syncthreads();
  if (threadIdx.x < NB)
    inv_diag[threadIdx.x] = threadIdx.x < h ? (unit_diag ? T(1) : T(1) / As[threadIdx.x][threadIdx.x]) : T(0);
  __syncthreads();
  const int j = threadIdx.x / LPC, q = threadIdx.x % LPC, lane = threadIdx.x & 63;
  T v[R];
#pragma unroll
  for (int t = 0; t < R; ++t) v[t] = (q + LPC * t == j && j < h) ? T(1) : T(0);
#pragma unroll
  for (int s = 0; s < NB; ++s) {
    const int p = LOWER ? s : NB - 1 - s;
    const int owner = p % LPC, tp = p / LPC;
    T xp = v[tp] * inv_diag[p];
    if (q == owner) v[tp] = xp;
    xp = __shfl(xp, (lane & ~(LPC - 1)) | owner, 64);
#pragma unroll
    for (int t = 0; t < R; ++t) {
      const int i = q + LPC * t;
      if (LOWER ? i > p : i < p) v[t] -= As[i][p] * xp;
    }
  }
  T* out = inv + (size_t)d * NB * NB;
#pragma unroll
  for (int t = 0; t < R; ++t) out[(size_t)(q + LPC * t) * NB + j] = v[t];
}

// Diagonal block of 64 < m <= TRSM_FB rows (starting on a multiple of 64)
// in one kernel.  A workgroup owns TRSM_FW = 16 columns of B, held in LDS
// from the first read to the last write (two workgroups fit a CU), and takes
// the rows in 64-row leaves (lower: top down, upper: bottom up).  Leaf l:
// X_l = inv(A_ll) B_l with the leaf's precomputed inverse, then
// B_c -= A_cl X_l for each 64-row chunk c still open (lower: below the
// leaf, upper: above it), all on the matrix cores: in each phase wave w
// multiplies rows 16w .. 16w+15 of one 64 x 64 block of A (or of the
// inverse) into the leaf's 64 x 16 X.  The wave reads its operands from
// memory straight into MFMA layout, the next phase's while this one
// computes: k is taken in the order where MFMA step s = 2t + b gives lane
// (i, q) column 8t + 2q + b, so each lane's operands are 8 adjacent pairs
// (8 loads of two scalars; a row's 64 columns are read by its 4 lanes).
// Addresses are clamped into the matrix and values masked after arrival.
// Barriers only order the leaf's reads and writes of X (three per leaf).
// B is read and written around the phases rather than before and after
// them (put_rows, store_leaf).
// X rows are stored with bit 0 of the row index flipped where bit 1 is set,
// so that the 4 rows an operand read takes (2 apart) fall in both bank halves.
constexpr int TRSM_FB = 256, TRSM_FW = 16;

// Marks a loaded value as used here, so that the compiler neither sinks the
// load behind a later branch nor past a barrier.
template <typename T>
__device__ __forceinline__ void trsm_pin(T& x) {
  asm volatile("" : "+v"(x));
}

template <typename T, bool LOWER>
__global__ __launch_bounds__(256) void trsm_block_kernel(int m, int n, const T* __restrict__ a,
                                                         size_t lda,
                                                         const T* __restrict__ inv,
                                                         T* __restrict__ b, size_t ldb) {
  constexpr int NB = TRSM_NB, FB = TRSM_FB, FW = TRSM_FW;
  constexpr int PR = 256 / FW;  // rows covered by one pass of the block
  constexpr int KS = NB / 4;    // MFMA steps of a 64-deep product
  using Mf = Mfma16x16x4<T>;
  using acc_t = typename Mf::acc_t;
  typedef T T2 __attribute__((ext_vector_type(2)));
  __shared__ T Xs[FB * FW];
  auto X = [&](int r, int c) -> T& { return Xs[(r ^ ((r >> 1) & 1)) * FW + c]; };
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int j0 = blockIdx.x * FW;
  const int leaves = (m + NB - 1) / NB;
  // row of X holding k-step s of lane group lk, relative to the leaf
  auto krow = [&](int s) { return 8 * (s >> 1) + 2 * lk + (s & 1); };
  // Whether k-step s of this wave's rows of a (triangular) diagonal-block
  // inverse can hold nonzeros: steps 4c .. 4c+3 span its columns 16c ..
  // 16c+15, zero above (lower) or below (upper) the wave's 16 rows; the
  // zero ones are not read.
  auto inv_step = [&](int s) { return LOWER ? (s >> 2) <= wave : (s >> 2) >= wave; };
  // B: thread (c, i) holds rows i + 16 e of column c, e / 4 being the leaf.
  // The first leaf's rows are read first and written to LDS before the
  // first product; the rest are written after it, so their reads overlap
  // it.  Each leaf's rows are stored as soon as its X_l is final.
  const int bc = tid % FW, bj = j0 + bc, bi = tid / FW;
  const int first = LOWER ? 0 : leaves - 1;
  T xin[FB / PR];
#pragma unroll
  for (int g = 0; g < FB / PR; ++g) {
    const int e = LOWER ? g : FB / PR - 1 - g;
    xin[e] = b[(size_t)min(bi + PR * e, m - 1) * ldb + min(bj, n - 1)];
  }
  auto put_rows = [&](bool first_leaf) {
#pragma unroll
    for (int e = 0; e < FB / PR; ++e)
      if ((e / 4 == first) == first_leaf) {
        trsm_pin(xin[e]);
        const int r = bi + PR * e;
        X(r, bc) = r < m ? xin[e] : T(0);
      }
  };
  auto store_leaf = [&](int leaf) {
#pragma unroll
    for (int e = 0; e < FB / PR; ++e)
      if (e / 4 == leaf) {
        const int r = bi + PR * e;
        if (r < m && bj < n) b[(size_t)r * ldb + bj] = X(r, bc);
      }
  };
  put_rows(true);
  const bool pairs = lda % 2 == 0 && reinterpret_cast<uintptr_t>(a) % (2 * sizeof(T)) == 0;
  auto leaf_of = [&](int li) { return LOWER ? li : leaves - 1 - li; };
  auto nch = [&](int leaf) { return LOWER ? leaves - 1 - leaf : leaf; };
  // Row 16 wave + lr of the block of phase (li, f) (f == 0: the leaf's
  // inverse; f >= 1: chunk f - 1 of A), columns 8t + 2 lk + {0, 1}.
  T stage[KS];
  int st_ok_row = 0, st_cols = 0;
  auto fetch = [&](int li, int f) {
    const int leaf = leaf_of(li);
    const int r0 = leaf * NB, h = min(NB, m - r0);
    const int row = 16 * wave + lr;
    if (f == 0) {
      const T* p = inv + (size_t)leaf * NB * NB + (size_t)row * NB + 2 * lk;
#pragma unroll
      for (int t = 0; t < KS / 2; ++t) {
        if (inv_step(2 * t)) {
          const T2 v = *reinterpret_cast<const T2*>(p + 8 * t);
          stage[2 * t] = v.x;
          stage[2 * t + 1] = v.y;
        } else {
          stage[2 * t] = stage[2 * t + 1] = T(0);
        }
      }
      st_ok_row = 1;
      st_cols = NB;
    } else {
      const int cs = LOWER ? r0 + NB * f : NB * (f - 1);
      const T* p = a + (size_t)min(cs + row, m - 1) * lda + r0 + 2 * lk;
      st_ok_row = cs + row < m;
      st_cols = h;
      if (pairs && h == NB) {
#pragma unroll
        for (int t = 0; t < KS / 2; ++t) {
          const T2 v = *reinterpret_cast<const T2*>(p + 8 * t);
          stage[2 * t] = v.x;
          stage[2 * t + 1] = v.y;
        }
      } else {
        const T* q = p - 2 * lk;
#pragma unroll
        for (int s = 0; s < KS; ++s) stage[s] = q[min(krow(s), h - 1)];
      }
    }
  };
  // The staged operands, masked: av[s] = A[row][krow(s)] (zero outside).
  auto operands = [&](T (&av)[KS]) {
#pragma unroll
    for (int s = 0; s < KS; ++s) trsm_pin(stage[s]);
#pragma unroll
    for (int s = 0; s < KS; ++s) av[s] = (st_ok_row && krow(s) < st_cols) ? stage[s] : T(0);
  };
  int pli = 0, pf = 0;  // the phase whose operands are being fetched
  auto fetch_next = [&]() {
    if (pf < nch(leaf_of(pli))) {
      ++pf;
    } else {
      ++pli;
      pf = 0;
    }
    if (pli < leaves) fetch(pli, pf);
  };
  fetch(0, 0);
  for (int li = 0; li < leaves; ++li) {
    const int leaf = leaf_of(li), r0 = leaf * NB, nc = nch(leaf);
    __syncthreads();  // the previous leaf's updates are in X
    {
      T av[KS];
      operands(av);
      fetch_next();
      // X_l = inv(A_ll) B_l: this wave's 16 rows of the leaf; k split in
      // halves, two accumulation chains
      acc_t acc[2] = {};
#pragma unroll
      for (int s = 0; s < KS / 2; ++s)
#pragma unroll
        for (int hk = 0; hk < 2; ++hk) {
          const int ss = s + hk * KS / 2;
          acc[hk] = Mf::run(av[ss], X(r0 + krow(ss), lr), acc[hk]);
        }
      __syncthreads();  // every read of B_l is done
#pragma unroll
      for (int r = 0; r < 4; ++r) X(r0 + 16 * wave + Mf::row(lk, r), lr) = acc[0][r] + acc[1][r];
      if (li == 0) put_rows(false);
      __syncthreads();  // X_l (and after the first leaf, all of B) is in X
      store_leaf(leaf);
    }
    // X_l as the B operands of every update of this leaf, held in registers
    T xb[KS];
    if (nc > 0) {
#pragma unroll
      for (int s = 0; s < KS; ++s) xb[s] = X(r0 + krow(s), lr);
    }
    for (int f = 1; f <= nc; ++f) {
      // B_c -= A_cl X_l: this wave's 16 rows of chunk f - 1
      T av[KS];
      operands(av);
      fetch_next();
      const int row0 = (LOWER ? r0 + NB * f : NB * (f - 1)) + 16 * wave;
      acc_t acc[2];
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[0][r] = X(row0 + Mf::row(lk, r), lr);
      acc[1] = acc_t{};
#pragma unroll
      for (int s = 0; s < KS / 2; ++s)
#pragma unroll
        for (int hk = 0; hk < 2; ++hk) {
          const int ss = s + hk * KS / 2;
          acc[hk] = Mf::run(-av[ss], xb[ss], acc[hk]);
        }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = row0 + Mf::row(lk, r);
        if (row < m) X(row, lr) = acc[0][r] + acc[1][r];
      }
    }
  }
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

// inv: the inverses of the 64 x 64 diagonal blocks of this A, 4096 scalars
// each (the recursion splits on multiples of 64 above 256 rows, so a sub-
// problem's blocks are a run of the whole matrix's).
template <typename T>
static int trsm_rec(const TrsmCtx<T>& c, size_t m, const T* a, T* b, const T* inv) {
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
  if (m <= (size_t)TRSM_FB) {
    const unsigned blocks = (unsigned)cdiv((size_t)c.n, TRSM_FW);
    if (c.upper)
      hipLaunchKernelGGL((trsm_block_kernel<T, false>), dim3(blocks), dim3(256), 0, c.st, (int)m,
                         c.n, a, c.lda, inv, b, c.ldb);
    else
      hipLaunchKernelGGL((trsm_block_kernel<T, true>), dim3(blocks), dim3(256), 0, c.st, (int)m,
                         c.n, a, c.lda, inv, b, c.ldb);
    FG_CHECK(hipGetLastError());
    return FERRAY_OK;
  }
  const size_t m1 = trsm_split(m), m2 = m - m1;
  const T* inv2 = inv + m1 / TRSM_NB * TRSM_NB * TRSM_NB;  // m1 is a multiple of 64 here
  int s;
  if (!c.upper) {
    // L11 X1 = B1; B2 -= L21 X1; L22 X2 = B2
    if ((s = trsm_rec(c, m1, a, b, inv))) return s;
    if ((s = trsm_update(c, m2, m1, a + m1 * c.lda, b, b + m1 * c.ldb))) return s;
    return trsm_rec(c, m2, a + m1 * c.lda + m1, b + m1 * c.ldb, inv2);
  }
  // U22 X2 = B2; B1 -= U12 X2; U11 X1 = B1
  if ((s = trsm_rec(c, m2, a + m1 * c.lda + m1, b + m1 * c.ldb, inv2))) return s;
  if ((s = trsm_update(c, m1, m2, a + m1, b + m1 * c.ldb, b))) return s;
  return trsm_rec(c, m1, a, b, inv);
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
  T* inv = nullptr;
  if (m > (size_t)TRSM_NB) {
    // Inverses of the 64 x 64 diagonal blocks, in workspace slot 0 (the
    // GEMM updates use only the padding slot).
    const size_t nblk = cdiv(m, (size_t)TRSM_NB);
    void* ws = nullptr;
    int s = workspace(st, nblk * TRSM_NB * TRSM_NB * sizeof(T), &ws);
    if (s) return s;
    inv = static_cast<T*>(ws);
    if (upper)
      hipLaunchKernelGGL((trsm_diag_inv_kernel<T, false>), dim3((unsigned)nblk), dim3(256), 0, st,
                         (int)m, a, lda, inv, (int)unit_diag);
    else
      hipLaunchKernelGGL((trsm_diag_inv_kernel<T, true>), dim3((unsigned)nblk), dim3(256), 0, st,
                         (int)m, a, lda, inv, (int)unit_diag);
    FG_CHECK(hipGetLastError());
  }
  return trsm_rec(c, m, a, b, inv);
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
