// One output tile of a real or complex floating-point GEMM on CDNA4 matrix
// cores, shared by the per-dtype entry points and the grouped kernel.
//
//   f64: v_mfma_f64_16x16x4_f64      f32: v_mfma_f32_16x16x4_f32
//
// A complex product C(m x n) = A(m x k) B(k x n) with interleaved [re, im]
// storage is computed as the real product
//     C_r(m x 2n) = A_r(m x 2k) . B_x(2k x 2n),
//     B_x[2p][2j] = br   B_x[2p][2j+1] = bi
//     B_x[2p+1][2j] = -bi  B_x[2p+1][2j+1] = br
// where A_r is A's memory read as reals and B_x is expanded from B while the
// tile is written to LDS.  C_r is C's interleaved memory, so there is no
// extra HBM pass and the real flop count is exactly 8mnk.
//
// 256 threads = 4 waves (2x2); block tile BM x BN, each wave (BM/2) x (BN/2)
// = TI x TJ MFMA tiles of 16x16; stages of BK k go from HBM through
// registers (16-byte vectors) into double-buffered LDS, one barrier per stage.
// Every output element accumulates the same sequence of MFMA k-steps in
// k order whatever the tile shape, so results do not depend on how a problem
// is tiled or batched.
#pragma once
#include "common.hpp"

namespace fg {

template <typename T>
struct Mfma16x16x4;
// Operand layout of both: lane l supplies A[l%16][l/16] and B[l/16][l%16].
// Result element r of lane l is C[row(l/16, r)][l%16]; the two instructions
// lay the rows out differently (measured on gfx950).
template <>
struct Mfma16x16x4<double> {
  using acc_t = d4;
  static __device__ __forceinline__ acc_t run(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return lk + 4 * r; }
};
template <>
struct Mfma16x16x4<float> {
  using acc_t = f4;
  static __device__ __forceinline__ acc_t run(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
};

template <typename T>
struct RealArgs {
  const T* a;
  const T* b;
  T* c;
  int M, N, K;  // real dimensions (complex: N = 2n, K = 2k)
  int lda, ldb, ldc;  // row strides in scalars of T (complex: twice the element stride)
  T alpha_re, alpha_im, beta_re, beta_im;
  int alpha_one;  // alpha == 1 (+0i): skip the scaling
  int beta_zero;  // beta == 0: C is written without being read
};

constexpr int RT_THREADS = 256;

template <typename T, int BM, int BN, int BK>
struct RealTileShape {
  static constexpr int V = 16 / sizeof(T);  // elements per 16-byte vector
  static constexpr int A_VECS = BM * BK / V;
  // Row pitch of the transposed A tile: padded so that the transposed stores
  // of one LDS phase fall in distinct banks (f64: 16 lanes = 4 rows x 4 k
  // pairs; f32: 32 lanes = 8 rows x 4 k quads), and the two k rows one f32
  // fragment read touches sit on different banks.
  static constexpr int A_PITCH = BM + (sizeof(T) == 8 ? 2 : 18);
  static constexpr int B_VECS = BK * BN / V;          // real B
  static constexpr int C_VECS = BK * BN / (2 * V);    // complex B (pairs expand to 2x2)
  static constexpr size_t lds_bytes = (size_t)2 * BK * (A_PITCH + BN) * sizeof(T);
};

// Computes the BM x BN tile at (m0, n0).  EDGE = false requires the whole
// tile and every k stage in range and 16-byte aligned rows; EDGE = true
// guards every element.  As holds 2*BK*A_PITCH elements, Bs 2*BK*BN (LDS).
template <typename T, bool CPLX, bool EDGE, int BM, int BN, int BK>
__device__ __forceinline__ void real_tile(const RealArgs<T>& p, int m0, int n0, T* As, T* Bs) {
  using MF = Mfma16x16x4<T>;
  using acc_t = typename MF::acc_t;
  using S = RealTileShape<T, BM, BN, BK>;
  constexpr int V = S::V;
  constexpr int TI = BM / 32, TJ = BN / 32;
  constexpr int AV = (S::A_VECS + RT_THREADS - 1) / RT_THREADS;
  constexpr int BVN = CPLX ? S::C_VECS : S::B_VECS;
  constexpr int BV = (BVN + RT_THREADS - 1) / RT_THREADS;
  static_assert(BM % 32 == 0 && BN % 32 == 0 && BK % 4 == 0 && BK % V == 0, "tile shape");
  typedef T vec_t __attribute__((ext_vector_type(V)));
  // As: [2][BK][A_PITCH], Bs: [2][BK][BN]
  constexpr int AP = S::A_PITCH;

  const int M = p.M, N = p.N, K = p.K;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave >> 1) * (BM / 2), wn0 = (wave & 1) * (BN / 2);
  const int lr = lane & 15, lk = lane >> 4;

  // Vector v of the A tile: row v/(BK/V), k = (v%(BK/V))*V .. +V.  Real B:
  // k row v/(BN/V), n = (v%(BN/V))*V .. +V.  Complex B: complex row
  // v/(BN/V), complex cols (v%(BN/V))*V/2 .. +V/2.
  vec_t ra[AV], rb[BV];

  auto load = [&](int k0) {
#pragma unroll
    for (int q = 0; q < AV; ++q) {
      const int v = tid + q * RT_THREADS;
      if (AV * RT_THREADS > S::A_VECS && v >= S::A_VECS) break;
      const int row = v / (BK / V), kc = (v % (BK / V)) * V;
      const int gr = m0 + row;
      if (!EDGE) {
        ra[q] = *reinterpret_cast<const vec_t*>(p.a + (size_t)gr * p.lda + k0 + kc);
      } else {
#pragma unroll
        for (int e = 0; e < V; ++e) {
          const int gk = k0 + kc + e;
          ra[q][e] = (gr < M && gk < K) ? p.a[(size_t)gr * p.lda + gk] : T(0);
        }
      }
    }
#pragma unroll
    for (int q = 0; q < BV; ++q) {
      const int v = tid + q * RT_THREADS;
      if (BV * RT_THREADS > BVN && v >= BVN) break;
      if (!CPLX) {
        const int kr = v / (BN / V), nc = (v % (BN / V)) * V;
        const int gk = k0 + kr;
        if (!EDGE) {
          rb[q] = *reinterpret_cast<const vec_t*>(p.b + (size_t)gk * p.ldb + n0 + nc);
        } else {
#pragma unroll
          for (int e = 0; e < V; ++e) {
            const int gn = n0 + nc + e;
            rb[q][e] = (gk < K && gn < N) ? p.b[(size_t)gk * p.ldb + gn] : T(0);
          }
        }
      } else {
        const int Kc = K >> 1, Nc = N >> 1;
        const int kr = v / (BN / V), cc = (v % (BN / V)) * (V / 2);
        const int kc = (k0 >> 1) + kr, nc = (n0 >> 1) + cc;
        if (!EDGE) {
          rb[q] = *reinterpret_cast<const vec_t*>(p.b + (size_t)kc * p.ldb + (size_t)nc * 2);
        } else {
#pragma unroll
          for (int e = 0; e < V / 2; ++e) {
            const bool ok = kc < Kc && nc + e < Nc;
            const T* src = p.b + (size_t)kc * p.ldb + (size_t)(nc + e) * 2;
            rb[q][2 * e] = ok ? src[0] : T(0);
            rb[q][2 * e + 1] = ok ? src[1] : T(0);
          }
        }
      }
    }
  };
  auto store = [&](int buf) {
    T* as = As + buf * BK * AP;
    T* bs = Bs + buf * BK * BN;
#pragma unroll
    for (int q = 0; q < AV; ++q) {
      const int v = tid + q * RT_THREADS;
      if (AV * RT_THREADS > S::A_VECS && v >= S::A_VECS) break;
      const int row = v / (BK / V), kc = (v % (BK / V)) * V;
#pragma unroll
      for (int e = 0; e < V; ++e) as[(kc + e) * AP + row] = ra[q][e];
    }
#pragma unroll
    for (int q = 0; q < BV; ++q) {
      const int v = tid + q * RT_THREADS;
      if (BV * RT_THREADS > BVN && v >= BVN) break;
      if (!CPLX) {
        const int kr = v / (BN / V), nc = (v % (BN / V)) * V;
        *reinterpret_cast<vec_t*>(&bs[kr * BN + nc]) = rb[q];
      } else {
        const int kr = v / (BN / V), cc = (v % (BN / V)) * (V / 2);
        const int r0 = 2 * kr, c0 = 2 * cc;
#pragma unroll
        for (int e = 0; e < V / 2; ++e) {
          const T br = rb[q][2 * e], bi = rb[q][2 * e + 1];
          bs[r0 * BN + c0 + 2 * e] = br;
          bs[r0 * BN + c0 + 2 * e + 1] = bi;
          bs[(r0 + 1) * BN + c0 + 2 * e] = -bi;
          bs[(r0 + 1) * BN + c0 + 2 * e + 1] = br;
        }
      }
    }
  };

  acc_t acc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  // Stage body; `more` is a constant at each call, so the steady-state
  // stages carry no branches, and the scheduler is given the same pipeline
  // order as the single-problem kernel: the next stage's global loads, the
  // first fragment reads, MFMAs interleaved with the following k-step's
  // reads, and the last k-step's MFMAs interleaved with the LDS writes.
  constexpr int NMF = TI * TJ;
  constexpr int NRD = (TI + TJ) / 2;                              // ds_read2 pairs
  constexpr int NWR = AV * V / 2 + (CPLX ? BV * V : BV);          // LDS write instructions
  constexpr int NVM = AV + BV;
  auto stage = [&](int cur, bool more, int k_next) {
    if (more) load(k_next);
    const T* as = As + cur * BK * AP;
    const T* bs = Bs + cur * BK * BN;
#pragma unroll
    for (int kk = 0; kk < BK / 4; ++kk) {
      T af[TI], bf[TJ];
#pragma unroll
      for (int i = 0; i < TI; ++i) af[i] = as[(kk * 4 + lk) * AP + wm0 + i * 16 + lr];
#pragma unroll
      for (int j = 0; j < TJ; ++j) bf[j] = bs[(kk * 4 + lk) * BN + wn0 + j * 16 + lr];
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) acc[i][j] = MF::run(af[i], bf[j], acc[i][j]);
    }
    if (more) store(cur ^ 1);
    if (more) __builtin_amdgcn_sched_group_barrier(0x020, NVM, 0);
    __builtin_amdgcn_sched_group_barrier(0x100, NRD, 0);
#pragma unroll
    for (int kk = 0; kk < BK / 4 - 1; ++kk) {
#pragma unroll
      for (int q = 0; q < NRD; ++q) {
        __builtin_amdgcn_sched_group_barrier(0x008, NMF / NRD > 0 ? NMF / NRD : 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
      }
    }
    if (more) {
#pragma unroll
      for (int q = 0; q < NWR; ++q) {
        __builtin_amdgcn_sched_group_barrier(0x008, NMF / NWR > 0 ? NMF / NWR : 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);
      }
    } else {
      __builtin_amdgcn_sched_group_barrier(0x008, NMF, 0);
    }
  };

  const int nk = (K + BK - 1) / BK;
  load(0);
  store(0);
  __syncthreads();
  int kt = 0;
  for (; kt + 1 < nk; ++kt) {
    stage(kt & 1, true, (kt + 1) * BK);
    __syncthreads();
  }
  stage(kt & 1, false, 0);

  // Epilogue.  acc[i][j][r] holds C[m0+wm0+16i+row(lk,r)][n0+wn0+16j+lr].
#pragma unroll
  for (int i = 0; i < TI; ++i) {
#pragma unroll
    for (int j = 0; j < TJ; ++j) {
      const int col = n0 + wn0 + j * 16 + lr;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm0 + i * 16 + MF::row(lk, r);
        T v = acc[i][j][r];
        if (!CPLX) {
          if (EDGE && (row >= M || col >= N)) continue;
          T* dst = p.c + (size_t)row * p.ldc + col;
          if (!p.alpha_one) v *= p.alpha_re;
          if (!p.beta_zero) v = fma(p.beta_re, *dst, v);
          *dst = v;
        } else {
          // lanes (2t, 2t+1) hold (re, im) of one complex element
          const T other = __shfl_xor(v, 1);
          const bool odd = lr & 1;
          T re = odd ? other : v, im = odd ? v : other;
          if (EDGE && (row >= M || col >= N)) continue;
          if (!p.alpha_one) {
            const T vr = p.alpha_re * re - p.alpha_im * im;
            const T vi = p.alpha_re * im + p.alpha_im * re;
            re = vr;
            im = vi;
          }
          T* dst = p.c + (size_t)row * p.ldc + (col & ~1);
          T outv;
          if (!p.beta_zero) {
            const T cr = dst[0], ci = dst[1];
            const T sr = p.beta_re * cr - p.beta_im * ci;
            const T si = p.beta_re * ci + p.beta_im * cr;
            outv = odd ? (im + si) : (re + sr);
          } else {
            outv = odd ? im : re;
          }
          dst[odd ? 1 : 0] = outv;
        }
      }
    }
  }
}

// Whether the tile at (m0, n0) can take the unguarded path.
template <typename T, int BM, int BN, int BK>
__device__ __forceinline__ bool real_tile_interior(const RealArgs<T>& p, int m0, int n0) {
  return m0 + BM <= p.M && n0 + BN <= p.N && p.K % BK == 0 && (p.lda * sizeof(T)) % 16 == 0 &&
         (p.ldb * sizeof(T)) % 16 == 0 && ((uintptr_t)p.a % 16) == 0 && ((uintptr_t)p.b % 16) == 0;
}

// Host launcher of the 64 x 64 tile kernel for products too small to fill
// the chip with 128 x 128 tiles (gemm_small.hip).
template <typename T, bool CPLX>
int launch_real_small(const RealArgs<T>& p, hipStream_t st);

}  // namespace fg
