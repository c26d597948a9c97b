// Split-precision DGEMM on INT8 matrix cores: the integer-modular form of the
// Ozaki scheme (Ozaki, Uchino, Imamura, "Ozaki Scheme II: A GEMM-oriented
// emulation of floating-point matrix multiplication using an integer modular
// technique", 2025).
//
//   1. Scaling (FP64).  Row i of A is scaled by 2^s_i and column j of B by
//      2^t_j so that ||2^s_i a_i||_2 <= 2^g and ||2^t_j b_j||_2 <= 2^g, then
//      rounded to integers A', B'.  With g = floor((log2 M - 2) / 2), where
//      M = m_1 ... m_N is the product of the moduli, Cauchy-Schwarz gives
//      |(A'B')_ij| <= M/4, so A'B' is recovered exactly from its residues.
//   2. Residues (FP64 -> INT8).  For each modulus m_l (pairwise coprime,
//      <= 256) the symmetric residues of A' and B' fit int8.
//   3. N INT8 GEMMs on the INT8 matrix cores (exact: |sum| <= k*2^14 < 2^31
//      for k <= 2^16 per chunk), each reduced mod m_l in the epilogue and
//      stored as one byte; for longer k each further chunk adds its product
//      to the stored residue mod m_l.
//   4. Reconstruction (FP64).  Chinese remaindering with the weights
//      W_l = (M/m_l) * ((M/m_l)^-1 mod m_l) split into 32-bit pieces, so every
//      partial sum is an exact double; the multiple of M is removed exactly,
//      the result is evaluated from balanced 32-bit digits, and scaled back by
//      2^-(s_i + t_j) before alpha/beta.
// Rows or columns holding Inf/NaN produce NaN in the affected outputs.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <type_traits>

#include "mfma_nt.hpp"

namespace fg {

constexpr int OZ_MAX_MODULI = 16;
// k bytes per residue product launch: every i32 sum of a chunk is exact
// (|residue| <= 128, so |acc| <= 2^16 * 2^14 = 2^30).
constexpr size_t OZ_K_CHUNK = (size_t)1 << 16;
// Pairwise coprime moduli, largest first (256 = 2^8, 255 = 3*5*17,
// 253 = 11*23, 247 = 13*19, 217 = 7*31, the rest prime).
static const int kModuli[OZ_MAX_MODULI] = {256, 255, 253, 251, 247, 241, 239, 233,
                                           229, 227, 223, 217, 211, 199, 197, 193};

struct OzConsts {
  int nmod;
  int m[OZ_MAX_MODULI];
  double md[OZ_MAX_MODULI];        // m, as a double
  double inv_md[OZ_MAX_MODULI];    // 1/m, for the residue reduction
  double c17[OZ_MAX_MODULI][2];    // 2^17 mod m, 2^34 mod m
  double w[OZ_MAX_MODULI][4];      // radix-2^R pieces of the CRT weight W_l
  double mp[4];                    // radix-2^R pieces of M
  double inv_M;                    // 1/M (approximate)
  double rad, irad;                // 2^R and 2^-R
};

// Constants for every modulus count, indexed by the count (uploaded once per
// process, before the first split-precision call).  Kernels read them from
// the constant address space, so each one is fetched by a scalar load where
// it is used instead of occupying registers for the whole kernel.
__constant__ OzConsts g_oz[OZ_MAX_MODULI + 1];

// ---------------------------------------------------------------------
// Scaling
// ---------------------------------------------------------------------

// Running (exponent, scaled sum of squares) pair: the 2-norm of the values
// seen so far is 2^E * sqrt(S), with S kept near 1 whatever the magnitudes
// (the scaled-sum-of-squares update of LAPACK's dlassq).
struct Ssq {
  int E;
  double S;
};
__device__ __forceinline__ Ssq ssq_add(Ssq a, double x) {
  if (x == 0.0 || !isfinite(x)) return a;
  const int ex = ilogb(x);
  if (a.E == INT_MIN) return Ssq{ex, ldexp(x, -ex) * ldexp(x, -ex)};
  if (ex > a.E) {
    const double y = ldexp(x, -ex);
    return Ssq{ex, fma(y, y, ldexp(a.S, 2 * (a.E - ex)))};
  }
  const double y = ldexp(x, -a.E);
  return Ssq{a.E, fma(y, y, a.S)};
}
__device__ __forceinline__ Ssq ssq_merge(Ssq a, Ssq b) {
  if (a.E == INT_MIN) return b;
  if (b.E == INT_MIN) return a;
  if (a.E >= b.E) return Ssq{a.E, a.S + ldexp(b.S, 2 * (b.E - a.E))};
  return Ssq{b.E, b.S + ldexp(a.S, 2 * (a.E - b.E))};
}
// Shift that brings a vector of 2-norm 2^E sqrt(S) to at most 2^g.
__device__ __forceinline__ int ssq_shift(Ssq a, int g) {
  if (a.E == INT_MIN) return 0;
  int ex;
  frexp(sqrt(a.S), &ex);  // sqrt(S) <= 2^ex
  return g - (a.E + ex);
}

__device__ __forceinline__ void block_stats(Ssq acc, int nonfinite, int g, int* pe, double* ps,
                                            int* anybad, int& shift, int& bad);

// Row scale of one row of A (s_i = g - e_i with ||a_i||_2 <= 2^e_i) and
// whether it holds a non-finite entry, computed by the whole block; every
// thread gets the results.
__device__ __forceinline__ void row_stats(const double* __restrict__ r, int K, int g, int& shift,
                                          int& bad) {
  __shared__ int pe[256];
  __shared__ double ps[256];
  __shared__ int anybad;
  if (threadIdx.x == 0) anybad = 0;
  __syncthreads();
  // Four independent partial sums so four loads are in flight per thread.
  Ssq acc4[4] = {{INT_MIN, 0.0}, {INT_MIN, 0.0}, {INT_MIN, 0.0}, {INT_MIN, 0.0}};
  int nonfinite = 0;
  int k = threadIdx.x;
  for (; k + 3 * 256 < K; k += 4 * 256) {
    double x[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) x[u] = r[k + u * 256];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      nonfinite |= !isfinite(x[u]);
      acc4[u] = ssq_add(acc4[u], x[u]);
    }
  }
  for (; k < K; k += 256) {
    const double x = r[k];
    nonfinite |= !isfinite(x);
    acc4[0] = ssq_add(acc4[0], x);
  }
  const Ssq acc = ssq_merge(ssq_merge(acc4[0], acc4[1]), ssq_merge(acc4[2], acc4[3]));
  block_stats(acc, nonfinite, g, pe, ps, &anybad, shift, bad);
}

// The block-wide part of row_stats: merges every thread's pair (and
// non-finite flag) in LDS; anybad must be zero and visible to all threads.
__device__ __forceinline__ void block_stats(Ssq acc, int nonfinite, int g, int* pe, double* ps,
                                            int* anybad, int& shift, int& bad) {
  pe[threadIdx.x] = acc.E;
  ps[threadIdx.x] = acc.S;
  __syncthreads();
  if (nonfinite) atomicOr(anybad, 1);
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const Ssq m2 = ssq_merge(Ssq{pe[threadIdx.x], ps[threadIdx.x]},
                               Ssq{pe[threadIdx.x + w], ps[threadIdx.x + w]});
      pe[threadIdx.x] = m2.E;
      ps[threadIdx.x] = m2.S;
    }
    __syncthreads();
  }
  shift = ssq_shift(Ssq{pe[0], ps[0]}, g);
  bad = *anybad;
}

// Column statistics of B (K x N row-major): partial pairs per k-slice
// (gridDim.y slices), combined by oz_col_finish.
__device__ __forceinline__ void oz_col_partial_block(int bx, int by, const double* __restrict__ b,
                                                     int K, int N, int kslice,
                                                     int* __restrict__ pe,
                                                     double* __restrict__ ps,
                                                     int* __restrict__ bad) {
  const int j = bx * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const int k0 = by * kslice, k1 = min(K, k0 + kslice);
  Ssq acc4[4] = {{INT_MIN, 0.0}, {INT_MIN, 0.0}, {INT_MIN, 0.0}, {INT_MIN, 0.0}};
  int nonfinite = 0;
  int k = k0;
  for (; k + 3 < k1; k += 4) {
    double x[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) x[u] = b[(size_t)(k + u) * N + j];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      nonfinite |= !isfinite(x[u]);
      acc4[u] = ssq_add(acc4[u], x[u]);
    }
  }
  for (; k < k1; ++k) {
    const double x = b[(size_t)k * N + j];
    nonfinite |= !isfinite(x);
    acc4[0] = ssq_add(acc4[0], x);
  }
  const Ssq acc = ssq_merge(ssq_merge(acc4[0], acc4[1]), ssq_merge(acc4[2], acc4[3]));
  pe[(size_t)by * N + j] = acc.E;
  ps[(size_t)by * N + j] = acc.S;
  if (nonfinite) atomicOr(bad + j, 1);
}
// Grid (column groups, k slices), launched in y slices from y0.
__global__ void oz_col_partial(const double* __restrict__ b, int K, int N, int kslice,
                               int* __restrict__ pe, double* __restrict__ ps,
                               int* __restrict__ bad, int y0) {
  oz_col_partial_block(blockIdx.x, y0 + (int)blockIdx.y, b, K, N, kslice, pe, ps, bad);
}

__global__ void oz_col_finish(int N, int slices, int g, const int* __restrict__ pe,
                              const double* __restrict__ ps, int* __restrict__ shift) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  Ssq acc{INT_MIN, 0.0};
  for (int t = 0; t < slices; ++t)
    acc = ssq_merge(acc, Ssq{pe[(size_t)t * N + j], ps[(size_t)t * N + j]});
  shift[j] = ssq_shift(acc, g);
}

__global__ void oz_init(int* cbad, int N) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < N) cbad[j] = 0;
}

// ---------------------------------------------------------------------
// Residues
// ---------------------------------------------------------------------

// Symmetric residue of the integer x (|x| < 2^63, held exactly in a double)
// modulo c.m[l], from 16-bit limbs of |x|: s = d0 + d1*(2^16 mod m) +
// d2*(2^32 mod m) + d3*(2^48 mod m) < 2^26, reduced with a float reciprocal
// whose quotient is off by at most one.
// An integer-valued double x, |x| < 2^51, as x = l0 + l1 2^17 + l2 2^34 with
// |l_i| < 2^17, every piece carrying the sign of x (all steps exact).
struct Limbs {
  double l0, l1, l2;
};
__device__ __forceinline__ Limbs limbs_of(double x) {
  const double l2 = trunc(x * 0x1.0p-34);
  const double rem = fma(-l2, 0x1.0p34, x);
  const double l1 = trunc(rem * 0x1.0p-17);
  return Limbs{fma(-l1, 0x1.0p17, rem), l1, l2};
}

// Symmetric residue of x modulo m = c.m[l], in DP arithmetic that is exact
// throughout: s = l0 + l1 (2^17 mod m) + l2 (2^34 mod m) is an integer with
// |s| < 2^26 congruent to x; q = rint(s / m) through the double reciprocal is
// the correctly rounded quotient (s / m is never within 1/512 of a
// half-integer unless it is one, and the reciprocal's error is below 1e-13),
// so r = s - q m lies in [-m/2, m/2]: [-(m-1)/2, (m-1)/2] for odd m, and
// for m = 256 the ends -128 and 128 are the same byte.
__device__ __forceinline__ int sym_residue(const Limbs& L, const OzConsts& c, int l) {
  const double s = fma(L.l2, c.c17[l][1], fma(L.l1, c.c17[l][0], L.l0));
  const double q = rint(s * c.inv_md[l]);
  return (int)fma(-q, c.md[l], s);
}

// Residue bytes of 4 values modulo c.m[l], packed little-endian.
__device__ __forceinline__ unsigned residue_word(const Limbs (&L)[4], const OzConsts& c, int l) {
  const unsigned r0 = (unsigned)sym_residue(L[0], c, l), r1 = (unsigned)sym_residue(L[1], c, l);
  const unsigned r2 = (unsigned)sym_residue(L[2], c, l), r3 = (unsigned)sym_residue(L[3], c, l);
  // byte 0 of each, in order: two byte pairs, then the pairs' low halves
  const unsigned lo = __builtin_amdgcn_perm(r1, r0, 0x0c0c0400u);
  const unsigned hi = __builtin_amdgcn_perm(r3, r2, 0x0c0c0400u);
  return __builtin_amdgcn_perm(hi, lo, 0x05040100u);
}

// Row scale and residue planes of A in one kernel: one block per row of the
// Mp x Kp planes.  The block first reduces the row's 2-norm (the row then
// sits in L2), then turns it into residues, 4 consecutive k per thread, so
// A is read from HBM once.  Rows at or beyond M are zero padding.
template <int NM>
__global__ __launch_bounds__(256) void oz_scale_residues_a(const double* __restrict__ a, int M,
                                                           int K, int Kp, int ldp, int vec,
                                                           int g, int* __restrict__ shift,
                                                           int* __restrict__ bad,
                                                           uint8_t* __restrict__ planes,
                                                           size_t plane_bytes) {
  const OzConsts& c = g_oz[NM];
  const int row = blockIdx.x;
  int sh = 0;
  if (row < M) {
    int bd;
    row_stats(a + (size_t)row * K, K, g, sh, bd);
    if (threadIdx.x == 0) {
      shift[row] = sh;
      bad[row] = bd;
    }
  }
  const double* src_row = a + (size_t)row * K;
  uint8_t* dst_row = planes + (size_t)row * ldp;
  for (int k0 = 4 * threadIdx.x; k0 < Kp; k0 += 4 * 256) {
    double x[4] = {0.0, 0.0, 0.0, 0.0};
    if (row < M) {
      const double* src = src_row + k0;
      if (vec && k0 + 4 <= K) {
        typedef double d2 __attribute__((ext_vector_type(2)));
        const d2 lo = *reinterpret_cast<const d2*>(src);
        const d2 hi = *reinterpret_cast<const d2*>(src + 2);
        x[0] = lo.x;
        x[1] = lo.y;
        x[2] = hi.x;
        x[3] = hi.y;
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (k0 + e < K) x[e] = src[e];
      }
    }
    Limbs L[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) L[e] = limbs_of(rint(ldexp(isfinite(x[e]) ? x[e] : 0.0, sh)));
#pragma unroll
    for (int l = 0; l < NM; ++l)
      *reinterpret_cast<unsigned*>(dst_row + k0 + l * plane_bytes) = residue_word(L, c, l);
  }
}

// The same for one row of at most 8192 elements (Kp <= 8192), reading A from
// memory once: thread t holds k = 4t + 1024 g + e (g < 8, e < 4) of its row
// in registers from the norm to the residues.
template <int NM>
__device__ __forceinline__ void oz_scale_residues_a_row(int row, const double* __restrict__ a,
                                                        int M, int K, int Kp, int ldp,
                                                        int vec, int g, int* __restrict__ shift,
                                                        int* __restrict__ bad,
                                                        uint8_t* __restrict__ planes,
                                                        size_t plane_bytes) {
  constexpr int G = 8;
  const OzConsts& c = g_oz[NM];
  __shared__ int pe[256];
  __shared__ double ps[256];
  __shared__ int anybad;
  if (threadIdx.x == 0) anybad = 0;
  const double* src_row = a + (size_t)row * K;
  double x[G][4];
#pragma unroll
  for (int gi = 0; gi < G; ++gi) {
    const int k0 = 4 * threadIdx.x + 1024 * gi;
#pragma unroll
    for (int e = 0; e < 4; ++e) x[gi][e] = 0.0;
    if (row < M && k0 < Kp) {
      const double* src = src_row + k0;
      if (vec && k0 + 4 <= K) {
        typedef double d2 __attribute__((ext_vector_type(2)));
        const d2 lo = *reinterpret_cast<const d2*>(src);
        const d2 hi = *reinterpret_cast<const d2*>(src + 2);
        x[gi][0] = lo.x;
        x[gi][1] = lo.y;
        x[gi][2] = hi.x;
        x[gi][3] = hi.y;
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (k0 + e < K) x[gi][e] = src[e];
      }
    }
  }
  __syncthreads();  // anybad cleared before any thread sets it
  int sh = 0, bd = 0;
  {
    Ssq acc4[4] = {{INT_MIN, 0.0}, {INT_MIN, 0.0}, {INT_MIN, 0.0}, {INT_MIN, 0.0}};
    int nonfinite = 0;
#pragma unroll
    for (int gi = 0; gi < G; ++gi)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        nonfinite |= !isfinite(x[gi][e]);
        acc4[e] = ssq_add(acc4[e], x[gi][e]);
      }
    const Ssq acc = ssq_merge(ssq_merge(acc4[0], acc4[1]), ssq_merge(acc4[2], acc4[3]));
    block_stats(acc, nonfinite, g, pe, ps, &anybad, sh, bd);
  }
  if (row < M && threadIdx.x == 0) {
    shift[row] = sh;
    bad[row] = bd;
  }
  if (row >= M) sh = 0;
  uint8_t* dst_row = planes + (size_t)row * ldp;
#pragma unroll
  for (int gi = 0; gi < G; ++gi) {
    const int k0 = 4 * threadIdx.x + 1024 * gi;
    if (k0 >= Kp) break;
    Limbs L[4];
#pragma unroll
    for (int e = 0; e < 4; ++e)
      L[e] = limbs_of(rint(ldexp(isfinite(x[gi][e]) ? x[gi][e] : 0.0, sh)));
#pragma unroll
    for (int l = 0; l < NM; ++l)
      *reinterpret_cast<unsigned*>(dst_row + k0 + l * plane_bytes) = residue_word(L, c, l);
  }
}

// B^T planes: plane l is Np x Kp int8 (row n holds column n of B).  A block
// takes 128 k x 32 n of B: coalesced row loads, stored transposed in LDS;
// then each thread turns 4 consecutive k of one n into residues, so a wave's
// 4-byte stores cover two full 128-byte rows of a plane.  Tiles: Np/32 x Kp/128.
template <int NM>
__device__ __forceinline__ void oz_residues_bt_tile(int tn_, int tk_, const double* __restrict__ b,
                                                    int K, int N, int Kp, int ldp,
                                                    const int* __restrict__ shift,
                                                    uint8_t* __restrict__ planes,
                                                    size_t plane_bytes) {
  const OzConsts& c = g_oz[NM];
  __shared__ double tile[32][130];
  const int tid = threadIdx.x;
  const int n0 = tn_ * 32, k0 = tk_ * 128;
  {
    const int nl = tid & 31, n = n0 + nl;
    for (int kr = tid >> 5; kr < 128; kr += 8) {
      const int k = k0 + kr;
      tile[nl][kr] = (k < K && n < N) ? b[(size_t)k * N + n] : 0.0;
    }
  }
  __syncthreads();
  const int tk = tid & 31, tn = tid >> 5;
  // 4 k x 4 n per thread: all 16 limbs first, then one pass per modulus.
  Limbs L[4][4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int nl = tn + 8 * j, n = n0 + nl;
    const int sh = n < N ? shift[n] : 0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const double x = tile[nl][4 * tk + e];
      L[j][e] = limbs_of(rint(ldexp(isfinite(x) ? x : 0.0, sh)));
    }
  }
  uint8_t* dst = planes + (size_t)(n0 + tn) * ldp + k0 + 4 * tk;
#pragma unroll
  for (int l = 0; l < NM; ++l)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      *reinterpret_cast<unsigned*>(dst + (size_t)(8 * j) * ldp + l * plane_bytes) =
          residue_word(L[j], c, l);
}

// A's scale-and-residue pass and the column statistics of B in one launch:
// block b does A row ia(b) or column block ib(b) of B's statistics, the A
// rows spread evenly through the grid, so that the A side's arithmetic and
// the B scan's memory traffic run on the chip at the same time.  Each block
// does exactly what the separate passes did (same values).
template <int NM>
__global__ __launch_bounds__(256) void oz_residues_a_colstats(
    const double* __restrict__ a, int M, int K, int Kp, int ldp, int vec, int g,
    int* __restrict__ shift_a,
    int* __restrict__ bad_a, uint8_t* __restrict__ planes_a, size_t a_plane, int n_a,
    const double* __restrict__ b, int N, int kslice, int n_cx, int n_c, int* __restrict__ pe,
    double* __restrict__ ps, int* __restrict__ bad_b) {
  const long long total = (long long)n_a + n_c, blk = blockIdx.x;
  const long long before = blk * n_a / total, upto = (blk + 1) * n_a / total;
  if (upto > before) {
    oz_scale_residues_a_row<NM>((int)before, a, M, K, Kp, ldp, vec, g, shift_a, bad_a,
                                planes_a, a_plane);
  } else {
    const int t = (int)(blk - upto);
    oz_col_partial_block(t % n_cx, t / n_cx, b, K, N, kslice, pe, ps, bad_b);
  }
}

template <int NM>
__global__ __launch_bounds__(256) void oz_residues_bt(const double* __restrict__ b, int K, int N,
                                                      int Kp, int ldp,
                                                      const int* __restrict__ shift,
                                                      uint8_t* __restrict__ planes,
                                                      size_t plane_bytes, int y0) {
  oz_residues_bt_tile<NM>(blockIdx.x, y0 + (int)blockIdx.y, b, K, N, Kp, ldp, shift, planes,
                          plane_bytes);
}

// GEMM epilogue: residue of the exact i32 product modulo m_plane, as a byte.
// |acc| < 2^31: the product with the rounded reciprocal is within 2^-21 of
// acc/m, so the floor is off by at most one and fixed below.
__device__ __forceinline__ uint8_t residue_byte(int acc, int mm, double inv) {
  const double q = floor((double)acc * inv);
  int v = acc - (int)q * mm;
  if (v < 0) v += mm;
  if (v >= mm) v -= mm;
  return (uint8_t)v;
}

// EpiResidue for a block tile inside C: one plane (the block's), no bounds
// checks (measured 5 % off the 14-moduli product at 8192^3).  ADD: a later
// k chunk, whose product is added to the residue the earlier chunks left
// (|acc| < 2^30, so acc + r stays exact in i32).
template <bool ADD>
struct EpiResidueFast {
  uint8_t* plane;
  int ld, mm;
  double inv;
  __device__ __forceinline__ void operator()(int row, int col, int, int acc) const {
    uint8_t* dst = plane + (size_t)row * ld + col;
    if (ADD) acc += *dst;
    *dst = residue_byte(acc, mm, inv);
  }
};

template <bool ADD>
struct EpiResidueT {
  uint8_t* r;
  size_t plane_bytes;
  int ld, M, N;
  int m[OZ_MAX_MODULI];
  double inv_m[OZ_MAX_MODULI];
  __device__ __forceinline__ void operator()(int row, int col, int z, int acc) const {
    if (row >= M || col >= N) return;
    uint8_t* dst = r + (size_t)z * plane_bytes + (size_t)row * ld + col;
    if (ADD) acc += *dst;
    *dst = residue_byte(acc, m[z], inv_m[z]);
  }
  __device__ __forceinline__ bool interior(int r0, int c0, int rows, int cols) const {
    return r0 + rows <= M && c0 + cols <= N;
  }
  // The NT kernel multiplies plane pair z = blockIdx.y, so every element a
  // block stores belongs to that plane.
  template <class F>
  __device__ __forceinline__ void with_fast(F&& f) const {
    const int z = blockIdx.y;
    f(EpiResidueFast<ADD>{r + (size_t)z * plane_bytes, ld, m[z], inv_m[z]});
  }
};
using EpiResidue = EpiResidueT<false>;

// Reconstruction: C = alpha * 2^-(s_i+t_j) * CRT(r_1..r_N) + beta * C.  Each
// thread finishes 4 consecutive columns of one row (residue planes have a
// row pitch of ldr >= N, a multiple of 8, so one 4-byte load per plane).
// Number of radix pieces the CRT weights are split into: 3 while
// NM * 255 * 2^R stays below 2^52 with 3R >= log2(M) (up to 15 moduli), else
// 4 pieces of 32 bits.
template <int NM>
struct CrtPieces {
  static constexpr int value = NM <= 15 ? 3 : 4;
};

// Finishes the CRT sum from its piece sums S_p = sum_l r_l w_lp (each an
// exact integer below 2^52): removes the multiple of M (exactly, piece by
// piece), normalises the pieces to balanced radix-2^R digits and returns the
// integer value as a double.
template <int P>
__device__ __forceinline__ double crt_finish(const double (&S)[P], const OzConsts& k) {
  double xa = S[P - 1];
#pragma unroll
  for (int p = P - 2; p >= 0; --p) xa = xa * k.rad + S[p];
  const double q = rint(xa * k.inv_M);
  double D[P];
#pragma unroll
  for (int p = 0; p < P; ++p) D[p] = fma(-q, k.mp[p], S[p]);
#pragma unroll
  for (int p = 0; p + 1 < P; ++p) {
    const double t = rint(D[p] * k.irad);
    D[p] = fma(-t, k.rad, D[p]);
    D[p + 1] += t;
  }
  double v = D[P - 1];
#pragma unroll
  for (int p = P - 2; p >= 0; --p) v = v * k.rad + D[p];
  return v;
}

// An integer 0 <= u < 2^32 as a double: its bits under the exponent of 2^52.
__device__ __forceinline__ double u32_to_double(unsigned u) {
  return __hiloint2double(0x43300000, (int)u) - 4503599627370496.0;
}

// Each wave finishes 256 consecutive columns of one row: lane L reconstructs
// columns 4L .. 4L+3 from one 4-byte load per residue plane and stores them
// itself (two 16-byte stores when PAIRS: C's rows 16-byte aligned).  Four
// columns per lane keep the kernel at ~60 registers, so enough waves are
// resident to cover the residue loads; 8 columns per lane with the stores
// regrouped through LDS ran 18-24 % slower at 8192^2 and 14 moduli, with
// identical results.
template <int NM, bool PAIRS>
__global__ __launch_bounds__(256) void oz_reconstruct(const uint8_t* __restrict__ res,
                                                      size_t plane_bytes, int ldr, int M, int N,
                                                      const int* __restrict__ sA,
                                                      const int* __restrict__ tB,
                                                      const int* __restrict__ badA,
                                                      const int* __restrict__ badB, double alpha,
                                                      double beta, double* __restrict__ c,
                                                      OzConsts k) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int segs = (N + 255) / 256;
  const size_t unit = (size_t)blockIdx.x * 4 + w;
  if (unit >= (size_t)M * segs) return;
  const int i = (int)(unit / segs);
  const int j0 = (int)(unit % segs) * 256 + 4 * lane;
  if (j0 >= N) return;
  // Piece sums of the 4 columns, accumulated modulus by modulus.  All
  // residue words are requested at once (the weights are kernel arguments:
  // measured faster than constant-memory or LDS copies here).
  const uint8_t* src = res + (size_t)i * ldr + j0;
  constexpr int P = CrtPieces<NM>::value;
  double S[4][P];
#pragma unroll
  for (int e = 0; e < 4; ++e)
#pragma unroll
    for (int q = 0; q < P; ++q) S[e][q] = 0.0;
  unsigned pk[NM];
#pragma unroll
  for (int l = 0; l < NM; ++l) pk[l] = *reinterpret_cast<const unsigned*>(src + (size_t)l * plane_bytes);
#pragma unroll
  for (int l = 0; l < NM; ++l) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const double r = u32_to_double((pk[l] >> (8 * e)) & 0xffu);
#pragma unroll
      for (int q = 0; q < P; ++q) S[e][q] = fma(r, k.w[l][q], S[e][q]);
    }
  }
  const int si = sA[i], bi = badA[i];
  double v[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int j = min(j0 + e, N - 1);  // columns past N are computed, never stored
    double x = ldexp(crt_finish<P>(S[e], k), -(si + tB[j]));
    if (bi | badB[j]) x = __builtin_nan("");
    v[e] = x * alpha;
  }
  double* row = c + (size_t)i * N + j0;
  if (PAIRS && j0 + 4 <= N) {
    typedef double d2 __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      d2 x = d2{v[2 * h], v[2 * h + 1]};
      if (beta != 0.0) {
        const d2 o = *reinterpret_cast<const d2*>(row + 2 * h);
        x = d2{fma(beta, o[0], x[0]), fma(beta, o[1], x[1])};
      }
      *reinterpret_cast<d2*>(row + 2 * h) = x;
    }
  } else {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if (j0 + e >= N) break;
      double x = v[e];
      if (beta != 0.0) x = fma(beta, row[e], x);
      row[e] = x;
    }
  }
}

// ---------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------

typedef unsigned __int128 u128;

static int make_consts(int nmod, OzConsts& c, int& gamma) {
  if (nmod < 2 || nmod > OZ_MAX_MODULI) return FERRAY_ERR_INVALID;
  std::memset(&c, 0, sizeof(c));
  c.nmod = nmod;
  u128 M = 1;
  for (int l = 0; l < nmod; ++l) M *= (u128)kModuli[l];
  int bits = 0;
  for (u128 t = M; t > 1; t >>= 1) ++bits;  // floor(log2 M)
  gamma = (bits - 2) / 2;
  // CRT radix (see CrtPieces): 3 pieces of R = ceil((bits + 1) / 3) bits up to
  // 15 moduli (R <= 40, so 15 * 255 * 2^R < 2^52), else 4 pieces of 32 bits.
  const int pieces = nmod <= 15 ? 3 : 4;
  const int R = pieces == 3 ? (bits + 1 + 2) / 3 : 32;
  if (pieces * R < bits + 1 || (double)nmod * 255.0 * std::ldexp(1.0, R) >= std::ldexp(1.0, 52))
    return FERRAY_ERR_INVALID;
  const u128 mask = ((u128)1 << R) - 1;
  for (int l = 0; l < nmod; ++l) {
    const int m = kModuli[l];
    c.m[l] = m;
    c.md[l] = (double)m;
    c.inv_md[l] = 1.0 / (double)m;
    for (int t = 0; t < 2; ++t) c.c17[l][t] = (double)(unsigned)(((u128)1 << (17 * (t + 1))) % (u128)m);
    const u128 Ml = M / (u128)m;
    const unsigned mm = (unsigned)(Ml % (u128)m);
    unsigned inv = 0;
    for (unsigned y = 1; y < (unsigned)m; ++y)
      if ((mm * y) % (unsigned)m == 1) {
        inv = y;
        break;
      }
    if (!inv) return FERRAY_ERR_INVALID;
    const u128 W = Ml * (u128)inv;  // < M since inv < m
    for (int p = 0; p < pieces; ++p)
      c.w[l][p] = (double)(unsigned long long)((W >> (R * p)) & mask);
  }
  for (int p = 0; p < pieces; ++p) c.mp[p] = (double)(unsigned long long)((M >> (R * p)) & mask);
  c.inv_M = 1.0 / (double)M;
  c.rad = std::ldexp(1.0, R);
  c.irad = std::ldexp(1.0, -R);
  return FERRAY_OK;
}

static inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Calls f(std::integral_constant<int, nmod>{}) for nmod in [2, 16].
template <class F>
static void with_moduli(int nmod, F&& f) {
  switch (nmod) {
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 3: f(std::integral_constant<int, 3>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 5: f(std::integral_constant<int, 5>{}); break;
    case 6: f(std::integral_constant<int, 6>{}); break;
    case 7: f(std::integral_constant<int, 7>{}); break;
    case 8: f(std::integral_constant<int, 8>{}); break;
    case 9: f(std::integral_constant<int, 9>{}); break;
    case 10: f(std::integral_constant<int, 10>{}); break;
    case 11: f(std::integral_constant<int, 11>{}); break;
    case 12: f(std::integral_constant<int, 12>{}); break;
    case 13: f(std::integral_constant<int, 13>{}); break;
    case 14: f(std::integral_constant<int, 14>{}); break;
    case 15: f(std::integral_constant<int, 15>{}); break;
    default: f(std::integral_constant<int, 16>{}); break;
  }
}

// Uploads the constant tables (once per process).
static int oz_tables() {
  static std::once_flag once;
  static int status = FERRAY_OK;
  std::call_once(once, [] {
    static OzConsts host[OZ_MAX_MODULI + 1];
    std::memset(host, 0, sizeof(host));
    for (int nm = 2; nm <= OZ_MAX_MODULI; ++nm) {
      int g = 0;
      if ((status = make_consts(nm, host[nm], g))) return;
    }
    const hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(g_oz), host, sizeof(host));
    if (e != hipSuccess) status = map_hip_error(e);
  });
  return status;
}

template <class S>
static int ozaki_run(size_t m, size_t n, size_t k, double alpha, const double* a,
                     const double* b, double beta, double* c, int nmod, hipStream_t st) {
  OzConsts cst;
  int gamma = 0;
  int s = make_consts(nmod, cst, gamma);
  if (s) return s;
  if ((s = oz_tables())) return s;
  // Kp: a whole number of GEMM stages and of the 128-k residue tiles.  ldp:
  // the planes' row pitch, 128 bytes past Kp when Kp is a multiple of 1024
  // (rows a power-of-two multiple of 1 KiB apart cost the 256 x 256 main
  // loop 7 % at 8192, measured against pitches of K + 128 .. K + 512).
  const size_t Kp = align_up(k, 128), Mp = align_up(m, S::BM), Np = align_up(n, S::BN);
  const size_t ldp = Kp % 1024 == 0 ? Kp + 128 : Kp;
  const size_t ldr = align_up(n, 8);
  const size_t a_plane = Mp * ldp, b_plane = Np * ldp, r_plane = align_up(m * ldr, 256);
  const int kslice = 256;
  const size_t slices = cdiv(k, kslice);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off = align_up(off + bytes, 256);
    return o;
  };
  const size_t o_pa = take(a_plane * nmod), o_pb = take(b_plane * nmod), o_r = take(r_plane * nmod);
  const size_t o_sa = take(m * 4), o_ba = take(m * 4), o_tb = take(n * 4), o_bb = take(n * 4);
  const size_t o_cm = take(slices * n * 4), o_cs = take(slices * n * 8);
  void* wsv = nullptr;
  if ((s = workspace(st, off, &wsv))) return s;
  uint8_t* ws = reinterpret_cast<uint8_t*>(wsv);
  int* sA = reinterpret_cast<int*>(ws + o_sa);
  int* badA = reinterpret_cast<int*>(ws + o_ba);
  int* tB = reinterpret_cast<int*>(ws + o_tb);
  int* badB = reinterpret_cast<int*>(ws + o_bb);
  int* pe = reinterpret_cast<int*>(ws + o_cm);
  double* ps = reinterpret_cast<double*>(ws + o_cs);

  const unsigned gn = (unsigned)cdiv(n, 256);
  hipLaunchKernelGGL(oz_init, dim3(gn), dim3(256), 0, st, badB, (int)n);
  const int vec_a = (k % 2 == 0) && ((uintptr_t)a % 16 == 0);
  const size_t n_c = (size_t)gn * slices;
  if (Mp + n_c > 0x7fffffff) return FERRAY_ERR_INVALID;
  with_moduli(nmod, [&](auto nm) {
    constexpr int NM = decltype(nm)::value;
    if (Kp <= 8192) {
      // A's pass beside B's column statistics (rows of at most 8192)
      hipLaunchKernelGGL((oz_residues_a_colstats<NM>), dim3((unsigned)(Mp + n_c)), dim3(256), 0,
                         st, a, (int)m, (int)k, (int)Kp, (int)ldp, vec_a, gamma, sA, badA,
                         ws + o_pa, a_plane, (int)Mp, b, (int)n, kslice, (int)gn, (int)n_c, pe,
                         ps, badB);
    } else {
      for_grid_y(slices, [&](int y0, unsigned ny) {
        hipLaunchKernelGGL(oz_col_partial, dim3(gn, ny), dim3(256), 0, st, b, (int)k, (int)n,
                           kslice, pe, ps, badB, y0);
      });
      hipLaunchKernelGGL((oz_scale_residues_a<NM>), dim3((unsigned)Mp), dim3(256), 0, st, a,
                         (int)m, (int)k, (int)Kp, (int)ldp, vec_a, gamma, sA, badA, ws + o_pa,
                         a_plane);
    }
  });
  hipLaunchKernelGGL(oz_col_finish, dim3(gn), dim3(256), 0, st, (int)n, (int)slices, gamma, pe,
                     ps, tB);
  with_moduli(nmod, [&](auto nm) {
    constexpr int NM = decltype(nm)::value;
    for_grid_y(Kp / 128, [&](int y0, unsigned ny) {
      hipLaunchKernelGGL((oz_residues_bt<NM>), dim3((unsigned)(Np / 32), ny), dim3(256), 0, st, b,
                         (int)k, (int)n, (int)Kp, (int)ldp, tB, ws + o_pb, b_plane, y0);
    });
  });
  FG_CHECK(hipGetLastError());

  NtOperands p{};
  p.a = ws + o_pa;
  p.bt = ws + o_pb;
  p.a_plane = a_plane;
  p.b_plane = b_plane;
  p.lda = ldp;
  p.ldb = ldp;
  p.tiles_m = (int)(Mp / S::BM);
  p.tiles_n = (int)(Np / S::BN);
  EpiResidueT<false> e{};
  e.r = ws + o_r;
  e.plane_bytes = r_plane;
  e.ld = (int)ldr;
  e.M = (int)m;
  e.N = (int)n;
  for (int l = 0; l < nmod; ++l) {
    e.m[l] = kModuli[l];
    e.inv_m[l] = 1.0 / (double)kModuli[l];
  }
  // k in chunks of OZ_K_CHUNK: each chunk's i32 sums stay exact (|r| <= 128
  // per residue, so |acc| <= 2^16 * 2^14), and later chunks add their
  // product to the residues mod m, which is the residue of the whole sum.
  for (size_t kb0 = 0; kb0 < Kp; kb0 += OZ_K_CHUNK) {
    NtOperands q = p;
    q.a = p.a + kb0;
    q.bt = p.bt + kb0;
    q.kbytes = (int)std::min<size_t>(OZ_K_CHUNK, Kp - kb0);
    if (kb0 == 0) {
      s = launch_mfma_nt<ELEM_I8, S, false>(q, nmod, e, st);
    } else {
      EpiResidueT<true> ea{};
      ea.r = e.r;
      ea.plane_bytes = e.plane_bytes;
      ea.ld = e.ld;
      ea.M = e.M;
      ea.N = e.N;
      for (int l = 0; l < nmod; ++l) {
        ea.m[l] = e.m[l];
        ea.inv_m[l] = e.inv_m[l];
      }
      s = launch_mfma_nt<ELEM_I8, S, false>(q, nmod, ea, st);
    }
    if (s) return s;
  }

  const size_t units = m * cdiv(n, 256);  // one wave per 256 columns of a row
  if (cdiv(units, 4) > 0x7fffffff) return FERRAY_ERR_INVALID;
  const bool pairs = n % 2 == 0 && (uintptr_t)c % 16 == 0;
  with_moduli(nmod, [&](auto nm) {
    constexpr int NM = decltype(nm)::value;
    if (pairs)
      hipLaunchKernelGGL((oz_reconstruct<NM, true>), dim3((unsigned)cdiv(units, 4)), dim3(256), 0,
                         st, ws + o_r, r_plane, (int)ldr, (int)m, (int)n, sA, tB, badA, badB,
                         alpha, beta, c, cst);
    else
      hipLaunchKernelGGL((oz_reconstruct<NM, false>), dim3((unsigned)cdiv(units, 4)), dim3(256), 0,
                         st, ws + o_r, r_plane, (int)ldr, (int)m, (int)n, sA, tB, badA, badB,
                         alpha, beta, c, cst);
  });
  FG_CHECK(hipGetLastError());
  return FERRAY_OK;
}

// k == 0 / empty handling shared with the native path's semantics.
__global__ void oz_scale_c(double* c, size_t count, double beta) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (size_t)gridDim.x * blockDim.x)
    c[i] = beta == 0.0 ? 0.0 : beta * c[i];
}

static int ozaki_dev(size_t m, size_t n, size_t k, double alpha, const double* a, const double* b,
                     double beta, double* c, int nmod, hipStream_t st) {
  if (m == 0 || n == 0) return FERRAY_OK;
  if (!c || (k && (!a || !b))) return FERRAY_ERR_INVALID;
  if (nmod == 0) nmod = OZ_MAX_MODULI;
  if (nmod < 2 || nmod > OZ_MAX_MODULI) return FERRAY_ERR_INVALID;
  if (k > (1u << 30) || m > (1u << 30) || n > (1u << 30)) return FERRAY_ERR_INVALID;
  if (k == 0) {
    if (beta == 1.0) return FERRAY_OK;
    size_t blocks = cdiv(m * n, 256);
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(oz_scale_c, dim3((unsigned)blocks), dim3(256), 0, st, c, m * n, beta);
    FG_CHECK(hipGetLastError());
    return FERRAY_OK;
  }
  if (m >= 256 && n >= 256) return ozaki_run<NtBig>(m, n, k, alpha, a, b, beta, c, nmod, st);
  return ozaki_run<NtSmall>(m, n, k, alpha, a, b, beta, c, nmod, st);
}

}  // namespace fg

using namespace fg;

extern "C" {

int ferray_dgemm_ozaki_dev(size_t m, size_t n, size_t k, double alpha, const double* a,
                           const double* b, double beta, double* c, int moduli, void* stream) {
  int s = ensure_device();
  if (s) return s;
  return ozaki_dev(m, n, k, alpha, a, b, beta, c, moduli, resolve_stream(stream));
}

int ferray_ozaki_max_moduli(void) { return OZ_MAX_MODULI; }

}  // extern "C"
