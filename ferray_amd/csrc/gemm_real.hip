// Real and complex floating-point GEMM entry points (DGEMM, SGEMM, ZGEMM,
// CGEMM) on CDNA4 matrix cores.
//
//   f64: v_mfma_f64_16x16x4_f64      f32: v_mfma_f32_16x16x4_f32
//
// One kernel template covers the four.  Its main loop is written out here
// rather than taken from real_tile.hpp: this form is the one measured
// fastest with the current compiler (its register assignment keeps the f64
// accumulator tuples 8-aligned).  Per output element it issues the same MFMA
// k-sequence as real_tile(), which the grouped kernel uses, so both give
// bit-identical results.  A complex
// product C(m x n) = A(m x k) B(k x n) with interleaved [re, im] storage is
// computed as the real product
//     C_r(m x 2n) = A_r(m x 2k) . B_x(2k x 2n),
//     B_x[2p][2j] = br   B_x[2p][2j+1] = bi
//     B_x[2p+1][2j] = -bi  B_x[2p+1][2j+1] = br
// where A_r is A's memory read as reals and B_x is expanded from B while the
// tile is written to LDS.  C_r is C's interleaved memory, so there is no extra
// HBM pass and the real flop count is exactly 8mnk.
//
// Block: 256 threads (4 waves, 2x2), block tile 128x128, BK = 16, each wave
// owns a 64x64 sub-tile = 4x4 MFMA tiles of 16x16.  A and B panels go through
// registers into double-buffered LDS (one barrier per k-step), with an
// explicit instruction schedule for each stage (see the stage body).  Tiles are
// ordered in groups of GROUP_M rows of tiles, and the linear block id is
// remapped so each XCD (blocks are dealt round-robin to the 8 XCDs) works on
// a contiguous range of that order, which keeps the A and B panels a group
// shares inside one XCD's L2.
#include <algorithm>
#include <type_traits>

#include "mfma_nt.hpp"
#include "real_tile.hpp"

namespace fg {

constexpr int RB_M = 128, RB_N = 128, R_THREADS = RT_THREADS, R_GROUP_M = 4;

// Whether a product of `tiles` 128 x 128 tiles runs faster on 64 x 64 tiles:
// below 256 tiles the 128 x 128 grid leaves CUs idle, and up to 2048 tiles a
// last wave of at most 128 of the 512 block slots (2 per CU) is mostly idle
// too.  Measured on MI355X (f64/f32/c64, 256-3072): 64 x 64 is 2.1-2.4x
// faster at 1024^3 and below, +13 % at 1536^3, +10 % at 3072^3, and 13 %
// slower at 2048^3 (256 tiles, one per CU).
static inline bool prefer_small_tiles(size_t tiles) {
  if (tiles < 256) return true;
  const size_t last = tiles % 512;
  return tiles < 2048 && last != 0 && last <= 128;
}
// k-depth of one pipeline stage: 16 for both (f64: 67 KiB of LDS per block,
// two blocks per CU; fewer barriers than 8-deep stages, measured ~0.5 %
// faster at 8192).
template <typename T>
struct StageK {
  static constexpr int value = 16;
};

__device__ __forceinline__ void tile_coords(int tiles_m, int tiles_n, int& tm, int& tn) {
  int nt = tiles_m * tiles_n;
  int bid = blockIdx.x;
  if ((nt & 7) == 0) bid = (bid & 7) * (nt >> 3) + (bid >> 3);
  int per_group = R_GROUP_M * tiles_n;
  int group = bid / per_group;
  int first_m = group * R_GROUP_M;
  int gsz = min(tiles_m - first_m, R_GROUP_M);
  int in_group = bid - group * per_group;
  tm = first_m + in_group % gsz;
  tn = in_group / gsz;
}

// EDGE = false requires M % 128 == 0, N % 128 == 0, K % BK == 0 and 16-byte
// aligned A, B (the host checks); EDGE = true guards every element.
template <typename T, bool CPLX, bool EDGE>
__global__ __launch_bounds__(R_THREADS) __attribute__((amdgpu_waves_per_eu(2, 8)))
void gemm_real_kernel(RealArgs<T> p) {
  using MF = Mfma16x16x4<T>;
  using acc_t = typename MF::acc_t;
  constexpr int BK = StageK<T>::value;
  constexpr int V = 16 / sizeof(T);                  // elements per 16-byte vector
  constexpr int AV = RB_M * BK / V / R_THREADS;      // A vectors per thread
  constexpr int BV = BK * RB_N / V / R_THREADS;      // B vectors per thread (real)
  constexpr int CV = BK * RB_N / (2 * V) / R_THREADS;  // B vectors per thread (complex)
  static_assert(AV >= 1 && BV >= 1 && CV >= 1, "loader shape");
  typedef T vec_t __attribute__((ext_vector_type(V)));
  // A rows are padded so that the transposed stores of one LDS phase (f64:
  // 16 lanes = 4 rows x 4 k pairs; f32: 32 lanes = 8 rows x 4 k quads) fall
  // in distinct banks; the f32 pad also keeps the two k rows that one
  // fragment read touches on different banks.
  constexpr int APAD = sizeof(T) == 8 ? 2 : 18;
  // f64: B rows padded by 8 elements, measured 0.6-0.7 % faster at 8192^3
  // than unpadded (same bits; pads of 2-16 within 0.1 % of each other).
  constexpr int BPAD = sizeof(T) == 8 ? 8 : 0;
  __shared__ T As[2][BK][RB_M + APAD];
  __shared__ T Bs[2][BK][RB_N + BPAD];

  const int M = p.M, N = p.N, K = p.K;
  const int tiles_m = (M + RB_M - 1) / RB_M, tiles_n = (N + RB_N - 1) / RB_N;
  int tm, tn;
  tile_coords(tiles_m, tiles_n, tm, tn);
  const int m0 = tm * RB_M, n0 = tn * RB_N;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave >> 1) * 64, wn0 = (wave & 1) * 64;
  const int lr = lane & 15, lk = lane >> 4;

  // Vector v of the A tile covers row v/(BK/V), k = (v%(BK/V))*V .. +V; vector
  // v of the B tile covers k row v/(BN/V), n = (v%(BN/V))*V .. +V (complex:
  // complex row v/(BN/V), complex cols (v%(BN/V))*V/2 .. +V/2).
  vec_t ra[AV], rb[CPLX ? CV : BV];

  auto load_a = [&](int k0) {
#pragma unroll
    for (int q = 0; q < AV; ++q) {
      const int v = tid + q * R_THREADS, row = v / (BK / V), kc = (v % (BK / V)) * V;
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
  };
  auto load_b = [&](int k0) {
    if (!CPLX) {
#pragma unroll
      for (int q = 0; q < BV; ++q) {
        const int v = tid + q * R_THREADS, kr = v / (RB_N / V), nc = (v % (RB_N / V)) * V;
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
      }
    } else {
      const int Kc = K >> 1, Nc = N >> 1;
#pragma unroll
      for (int q = 0; q < CV; ++q) {
        const int v = tid + q * R_THREADS, kr = v / (RB_N / V), cc = (v % (RB_N / V)) * (V / 2);
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
  auto store_tiles = [&](int buf) {
#pragma unroll
    for (int q = 0; q < AV; ++q) {
      const int v = tid + q * R_THREADS, row = v / (BK / V), kc = (v % (BK / V)) * V;
#pragma unroll
      for (int e = 0; e < V; ++e) As[buf][kc + e][row] = ra[q][e];
    }
    if (!CPLX) {
#pragma unroll
      for (int q = 0; q < BV; ++q) {
        const int v = tid + q * R_THREADS, kr = v / (RB_N / V), nc = (v % (RB_N / V)) * V;
        *reinterpret_cast<vec_t*>(&Bs[buf][kr][nc]) = rb[q];
      }
    } else {
#pragma unroll
      for (int q = 0; q < CV; ++q) {
        const int v = tid + q * R_THREADS, kr = v / (RB_N / V), cc = (v % (RB_N / V)) * (V / 2);
        const int r0 = 2 * kr, c0 = 2 * cc;
#pragma unroll
        for (int e = 0; e < V / 2; ++e) {
          const T br = rb[q][2 * e], bi = rb[q][2 * e + 1];
          Bs[buf][r0][c0 + 2 * e] = br;
          Bs[buf][r0][c0 + 2 * e + 1] = bi;
          Bs[buf][r0 + 1][c0 + 2 * e] = -bi;
          Bs[buf][r0 + 1][c0 + 2 * e + 1] = br;
        }
      }
    }
  };

  acc_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  const int nk = (K + BK - 1) / BK;
  load_a(0);
  load_b(0);
  store_tiles(0);
  __syncthreads();

  // Stage body; `more` is a constant at each call, so the steady-state
  // stages carry no branches.
  auto stage = [&](int cur, bool more, int k_next) {
    if (more) {
      load_a(k_next);
      load_b(k_next);
    }
#pragma unroll
    for (int kk = 0; kk < BK / 4; ++kk) {
      T af[4], bf[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) af[i] = As[cur][kk * 4 + lk][wm0 + i * 16 + lr];
#pragma unroll
      for (int j = 0; j < 4; ++j) bf[j] = Bs[cur][kk * 4 + lk][wn0 + j * 16 + lr];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = MF::run(af[i], bf[j], acc[i][j]);
    }
    if (more) store_tiles(cur ^ 1);
    {
      // Pipeline order for the scheduler (masks: 0x020 VMEM read, 0x100 DS
      // read, 0x200 DS write, 0x008 MFMA): the next stage's global loads,
      // this stage's first fragment reads, MFMAs interleaved with the
      // fragment reads of the following k-step, and the last k-step's MFMAs
      // interleaved with the LDS writes of the next stage, so the stage tail
      // (write, wait, barrier) overlaps matrix work.  Measured at 8192 (f64,
      // f32) and 4096 (complex): 4-10 % faster than the compiler's own
      // order, with identical results (only the instruction order changes).
      // f64 (real and complex) also spreads the global loads over the first
      // k-step's MFMAs (+5 % at 8192); f32 issues them up front (spreading
      // measured 0.4 % slower there).
      constexpr int NRD = 4, NWR = 8, NVM = AV + (CPLX ? CV : BV);
      constexpr bool SPREAD = sizeof(T) == 8;
      if (more && !SPREAD) __builtin_amdgcn_sched_group_barrier(0x020, NVM, 0);
      __builtin_amdgcn_sched_group_barrier(0x100, NRD, 0);
#pragma unroll
      for (int kk = 0; kk < BK / 4 - 1; ++kk) {
#pragma unroll
        for (int q = 0; q < NRD; ++q) {
          if (SPREAD && more && kk == 0) {
            __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
            __builtin_amdgcn_sched_group_barrier(0x020, NVM / NRD, 0);
            __builtin_amdgcn_sched_group_barrier(0x008, 16 / NRD - 2, 0);
          } else {
            __builtin_amdgcn_sched_group_barrier(0x008, 16 / NRD, 0);
          }
          __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
        }
      }
      if (more && SPREAD) {
        // f64: the writes go one per MFMA in the middle of the last k-step
        // (+0.4 % over two per MFMA; f32 measured 0.5 % slower with it).
        __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);
#pragma unroll
        for (int q = 0; q < NWR; ++q) {
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
          __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);
        }
        __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);
      } else if (more) {
#pragma unroll
        for (int q = 0; q < NWR; ++q) {
          __builtin_amdgcn_sched_group_barrier(0x008, 16 / NWR, 0);
          __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);
        }
      } else {
        __builtin_amdgcn_sched_group_barrier(0x008, 16, 0);
      }
    }
  };
  int kt = 0;
  for (; kt + 1 < nk; ++kt) {
    stage(kt & 1, true, (kt + 1) * BK);
    __syncthreads();
  }
  stage(kt & 1, false, 0);

  // Epilogue.  acc[i][j][r] holds C[m0+wm0+16i+row(lk,r)][n0+wn0+16j+lr].
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
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
          T other = __shfl_xor(v, 1);
          const bool odd = lr & 1;
          T re = odd ? other : v, im = odd ? v : other;
          if (EDGE && (row >= M || col >= N)) continue;
          T outv;
          if (!p.alpha_one) {
            T vr = p.alpha_re * re - p.alpha_im * im;
            T vi = p.alpha_re * im + p.alpha_im * re;
            re = vr;
            im = vi;
          }
          T* dst = p.c + (size_t)row * p.ldc + (col & ~1);
          if (!p.beta_zero) {
            T cr = dst[0], ci = dst[1];
            T sr = p.beta_re * cr - p.beta_im * ci;
            T si = p.beta_re * ci + p.beta_im * cr;
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

// ---------------------------------------------------------------------
// SGEMM on 256 x 256 block tiles
// ---------------------------------------------------------------------
// Real f32 products whose tiles all lie inside C: four waves of 128 x 128
// (8 x 8 MFMA tiles, 256 accumulator registers, one wave per SIMD).  A stage
// is 16 k of both panels (32 KiB), copied HBM -> LDS by global_load_lds
// without register staging, two stages in LDS, and the NT kernel's rotated
// pipeline: a stage's last k-step is multiplied after the barrier that opens
// the next stage.  Per output element it issues the same MFMA k-steps in the
// same order as gemm_real_kernel, so the results are bit-identical; four
// times the MFMA work per fragment byte and block.
//   A stage: 256 rows of 16 k (64 bytes); 16-byte chunk c of row r at slot
//            c ^ ((r >> 2) & 3), so the 16 rows x 4 k of a fragment read hit
//            64 distinct banks.
//   B stage: 16 k-rows of 256 columns (1 KiB); 64-byte segment g of k-row r
//            at segment g ^ (r & 3), so the 4 k-rows x 16 columns of a
//            fragment read hit 64 distinct banks.
constexpr int SB_M = 256, SB_N = 256, SB_K = 16;
constexpr int SB_ABYTES = SB_M * SB_K * 4, SB_STAGE = SB_ABYTES + SB_K * SB_N * 4;

__device__ __forceinline__ void sb_tile_coords(int tiles_m, int tiles_n, int& tm, int& tn) {
  const int G = 4;  // one XCD's share of tile rows at 8192
  const int nt = tiles_m * tiles_n;
  int bid = blockIdx.x;
  if ((nt & 7) == 0) bid = (bid & 7) * (nt >> 3) + (bid >> 3);
  const int per_group = G * tiles_n, group = bid / per_group, first_m = group * G;
  const int gsz = min(tiles_m - first_m, G), in_group = bid - group * per_group;
  tm = first_m + in_group % gsz;
  tn = in_group / gsz;
}

template <bool ALPHA_ONE, bool BETA_ZERO>
__global__ __launch_bounds__(256) void sgemm_big_kernel(RealArgs<float> p) {
  using MF = Mfma16x16x4<float>;
  constexpr int TI = 8, TJ = 8, NKK = SB_K / 4;
  constexpr int NMF = TI * TJ, NRD = TI + TJ, NLD = 8;
  extern __shared__ __attribute__((aligned(16))) uint8_t sb_smem[];
  const int tiles_m = p.M / SB_M, tiles_n = p.N / SB_N;
  int tm, tn;
  sb_tile_coords(tiles_m, tiles_n, tm, tn);
  const int m0 = tm * SB_M, n0 = tn * SB_N;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm0 = (wave >> 1) * 128, wn0 = (wave & 1) * 128;
  const int lr = lane & 15, lk = lane >> 4;

  // Copy instruction q of wave w: A rows 16 (4q + w) + lane / 4 (slot lane & 3,
  // whose row has (r >> 2) & 3 = lane >> 4), and B k-row 4q + w (physical
  // segment lane / 4 holds logical segment (lane / 4) ^ w).  Offsets within
  // a stage are 32-bit (the launcher checks that 256 rows fit).
  const float* A = p.a + (size_t)m0 * p.lda;
  const float* B = p.b + n0;
  const unsigned a_off = (unsigned)(((16 * wave + (lane >> 2)) * p.lda + 4 * ((lane & 3) ^ (lane >> 4))) * 4);
  const unsigned a_q = (unsigned)(64 * p.lda * 4);
  const unsigned b_off = (unsigned)((wave * p.ldb + 16 * ((lane >> 2) ^ wave) + 4 * (lane & 3)) * 4);
  const unsigned b_q = (unsigned)(4 * p.ldb * 4);
  auto issue = [&](int buf, int k0) __attribute__((always_inline)) {
    uint8_t* as = sb_smem + buf * SB_STAGE + wave * 1024;
    uint8_t* bs = sb_smem + buf * SB_STAGE + SB_ABYTES + wave * 1024;
    const uint8_t* ga = reinterpret_cast<const uint8_t*>(A + k0);
    const uint8_t* gb = reinterpret_cast<const uint8_t*>(B + (size_t)k0 * p.ldb);
#pragma unroll
    for (int q = 0; q < 4; ++q)
      __builtin_amdgcn_global_load_lds((nt_gvoid*)(ga + (a_off + q * a_q)),
                                       (nt_lvoid*)(as + q * 4096), 16, 0, 0);
#pragma unroll
    for (int q = 0; q < 4; ++q)
      __builtin_amdgcn_global_load_lds((nt_gvoid*)(gb + (b_off + q * b_q)),
                                       (nt_lvoid*)(bs + q * 4096), 16, 0, 0);
  };
  // Fragments of k-step kk: A[wm0 + 16 i + lr][4 kk + lk], B[4 kk + lk][wn0 + 16 j + lr].
  const int a_rd = (wm0 + lr) * 64 + 4 * lk, a_sw = lr >> 2;
  const int b_rd = SB_ABYTES + lk * 1024 + 4 * lr, b_seg = wn0 >> 4;
  auto read_step = [&](const uint8_t* s, int kk, float* fa, float* fb) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < TI; ++i)
      fa[i] = *reinterpret_cast<const float*>(s + a_rd + i * 1024 + ((kk ^ a_sw) << 4));
#pragma unroll
    for (int j = 0; j < TJ; ++j)
      fb[j] = *reinterpret_cast<const float*>(s + b_rd + kk * 4096 + (((b_seg + j) ^ lk) << 6));
  };

  f4 acc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  auto mma = [&](const float* fa, const float* fb) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j) acc[i][j] = MF::run(fa[i], fb[j], acc[i][j]);
  };

  float pa[TI], pb[TJ];  // last k-step of the previous stage
  auto body = [&](auto pend_c, auto more_c, int cur, int k_next) __attribute__((always_inline)) {
    constexpr bool PEND = decltype(pend_c)::value, MORE = decltype(more_c)::value;
    const uint8_t* s = sb_smem + cur * SB_STAGE;
    float fa[NKK - 1][TI], fb[NKK - 1][TJ];
    read_step(s, 0, fa[0], fb[0]);
    if (MORE) issue(cur ^ 1, k_next);
    if (PEND) mma(pa, pb);
#pragma unroll
    for (int kk = 0; kk + 1 < NKK; ++kk) {
      if (kk + 2 < NKK) read_step(s, kk + 1, fa[kk + 1], fb[kk + 1]);
      else read_step(s, kk + 1, pa, pb);
      mma(fa[kk], fb[kk]);
    }
    // Order for the scheduler (masks: 0x010 VMEM, 0x100 DS read, 0x008
    // MFMA), as in the NT kernel.
    __builtin_amdgcn_sched_group_barrier(0x100, NRD, 0);
    if (PEND) {
#pragma unroll
      for (int q = 0; q < NMF; ++q) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        if (MORE && q % (NMF / NLD) == 0) __builtin_amdgcn_sched_group_barrier(0x010, 1, 0);
      }
    } else if (MORE) {
      __builtin_amdgcn_sched_group_barrier(0x010, NLD, 0);
    }
#pragma unroll
    for (int kk = 0; kk + 1 < NKK; ++kk) {
#pragma unroll
      for (int q = 0; q < NRD; ++q) {
        __builtin_amdgcn_sched_group_barrier(0x008, NMF / NRD, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
      }
    }
  };
  auto barrier = [&]() __attribute__((always_inline)) {
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
    __syncthreads();
    __builtin_amdgcn_sched_barrier(0);
  };
  using T_ = std::integral_constant<bool, true>;
  using F_ = std::integral_constant<bool, false>;
  const int nk = p.K / SB_K;
  issue(0, 0);
  __syncthreads();
  if (nk > 1)
    body(F_{}, T_{}, 0, SB_K);
  else
    body(F_{}, F_{}, 0, 0);
  int kt = 1;
  for (; kt + 1 < nk; ++kt) {
    barrier();
    body(T_{}, T_{}, kt & 1, (kt + 1) * SB_K);
  }
  if (kt < nk) {
    barrier();
    body(T_{}, F_{}, kt & 1, 0);
  }
  mma(pa, pb);

  // Epilogue: acc[i][j][r] holds C[m0 + wm0 + 16 i + 4 lk + r][n0 + wn0 + 16 j + lr];
  // every value leaves the accumulators through a VGPR copy.
  float* crow = p.c + (size_t)(m0 + wm0 + 4 * lk) * p.ldc + n0 + wn0 + lr;
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        float* dst = crow + (size_t)(16 * i + MF::row(0, r)) * p.ldc + 16 * j;
        float v = nt_vgpr(acc[i][j][r]);
        if (!ALPHA_ONE) v *= p.alpha_re;
        if (!BETA_ZERO) v = fma(p.beta_re, *dst, v);
        *dst = v;
      }
}

// C = beta * C for k == 0 (outer.rs:214-231, :1249-1266).
template <typename T, bool CPLX>
__global__ void scale_c_kernel(T* c, size_t count, T beta_re, T beta_im) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) {
    if (!CPLX) {
      c[i] = (beta_re == T(0)) ? T(0) : beta_re * c[i];
    } else {
      T* e = c + 2 * i;
      if (beta_re == T(0) && beta_im == T(0)) {
        e[0] = T(0);
        e[1] = T(0);
      } else {
        T cr = e[0], ci = e[1];
        e[0] = beta_re * cr - beta_im * ci;
        e[1] = beta_re * ci + beta_im * cr;
      }
    }
  }
}

// One complex element as stored (re, im); aligned as T, so unaligned
// complex operands are read correctly.
template <typename T>
struct CplxPair {
  T re, im;
};

// Zero-padded copy: dst (rows_p x cols_p, row pitch cols_p) takes src
// (rows x cols, row pitch ld) in its top-left corner and zeros elsewhere.
// Elements are E (the real or complex scalar); x covers columns, y strides
// over rows.
template <typename E>
__global__ void pad_copy_kernel(E* __restrict__ dst, const E* __restrict__ src, size_t rows,
                                size_t cols, size_t ld, size_t rows_p, size_t cols_p) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols_p) return;
  for (size_t r = blockIdx.y; r < rows_p; r += gridDim.y)
    dst[r * cols_p + c] = (r < rows && c < cols) ? src[r * ld + c] : E{};
}

// dst (rows x cols, pitch ld) = top-left corner of src (pitch cols_p).
template <typename E>
__global__ void unpad_copy_kernel(E* __restrict__ dst, const E* __restrict__ src, size_t rows,
                                  size_t cols, size_t ld, size_t cols_p) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  for (size_t r = blockIdx.y; r < rows; r += gridDim.y) dst[r * ld + c] = src[r * cols_p + c];
}

// C = beta * C over an m x n block with row stride ldc (k == 0 with strides).
template <typename T, bool CPLX>
__global__ void scale_c_strided_kernel(T* c, size_t m, size_t n, size_t ldc, T beta_re,
                                       T beta_im) {
  const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  for (size_t r = blockIdx.y; r < m; r += gridDim.y) {
    if (!CPLX) {
      T* e = c + r * ldc + j;
      *e = beta_re == T(0) ? T(0) : beta_re * *e;
    } else {
      T* e = c + 2 * (r * ldc + j);
      if (beta_re == T(0) && beta_im == T(0)) {
        e[0] = T(0);
        e[1] = T(0);
      } else {
        const T cr = e[0], ci = e[1];
        e[0] = beta_re * cr - beta_im * ci;
        e[1] = beta_re * ci + beta_im * cr;
      }
    }
  }
}

static dim3 copy_grid(size_t rows, size_t cols) {
  return dim3((unsigned)cdiv(cols, 256), (unsigned)std::min<size_t>(rows, 65535));
}

// C = alpha A B + beta C with row strides lda >= k, ldb >= n, ldc >= n in
// elements (complex elements for complex); the tight entry points pass
// lda = k, ldb = ldc = n.
template <typename T, bool CPLX>
static int launch_real(size_t m, size_t n, size_t k, T ar, T ai, const T* a, size_t lda,
                       const T* b, size_t ldb, T br, T bi, T* c, size_t ldc, hipStream_t st) {
  if (m == 0 || n == 0) return FERRAY_OK;
  if (!c || (k && (!a || !b))) return FERRAY_ERR_INVALID;
  if (lda < k || ldb < n || ldc < n) return FERRAY_ERR_INVALID;
  const size_t mult = CPLX ? 2 : 1;
  if (m > (1u << 30) || n * mult > (1u << 30) || k * mult > (1u << 30) ||
      lda * mult > (1u << 30) || ldb * mult > (1u << 30) || ldc * mult > (1u << 30))
    return FERRAY_ERR_INVALID;
  if (k == 0) {
    const bool beta_one = br == T(1) && bi == T(0);
    if (beta_one) return FERRAY_OK;
    if (ldc == n) {
      size_t count = m * n;
      size_t blocks = cdiv(count, 256);
      if (blocks > 65536) blocks = 65536;
      hipLaunchKernelGGL((scale_c_kernel<T, CPLX>), dim3((unsigned)blocks), dim3(256), 0, st, c,
                         count, br, bi);
    } else {
      hipLaunchKernelGGL((scale_c_strided_kernel<T, CPLX>), copy_grid(m, n), dim3(256), 0, st, c,
                         m, n, ldc, br, bi);
    }
    FG_CHECK(hipGetLastError());
    return FERRAY_OK;
  }
  RealArgs<T> p;
  p.a = a;
  p.b = b;
  p.c = c;
  p.M = (int)m;
  p.N = (int)(n * mult);
  p.K = (int)(k * mult);
  p.lda = (int)(lda * mult);
  p.ldb = (int)(ldb * mult);
  p.ldc = (int)(ldc * mult);
  p.alpha_re = ar;
  p.alpha_im = ai;
  p.beta_re = br;
  p.beta_im = bi;
  p.alpha_one = (ar == T(1) && ai == T(0));
  p.beta_zero = (br == T(0) && bi == T(0));
  const size_t tiles = cdiv(m, RB_M) * cdiv(n * mult, RB_N);
  if (tiles > 0x7fffffff) return FERRAY_ERR_INVALID;
  // Small or badly quantised grids: 64 x 64 tiles, bit-identical results.
  if (prefer_small_tiles(tiles)) return launch_real_small<T, CPLX>(p, st);
  const bool aligned = ((uintptr_t)a % 16 == 0) && ((uintptr_t)b % 16 == 0) &&
                       (lda * mult * sizeof(T)) % 16 == 0 && (ldb * mult * sizeof(T)) % 16 == 0;
  const bool fast = aligned && m % RB_M == 0 && (n * mult) % RB_N == 0 &&
                    (k * mult) % StageK<T>::value == 0;
  if constexpr (!CPLX && sizeof(T) == 4) {
    // 256 x 256 tiles when every one lies inside C, they fill the chip, and
    // 256 rows of A and 16 rows of B each span less than 2 GiB (the copy
    // offsets within a stage are 32-bit)
    const size_t big = (m / SB_M) * (n / SB_N);
    if (fast && m % SB_M == 0 && n % SB_N == 0 && k % SB_K == 0 && big >= 256 &&
        big <= 0x7fffffff && (size_t)SB_M * lda * 4 < ((size_t)1 << 31) &&
        (size_t)SB_K * ldb * 4 < ((size_t)1 << 31)) {
      static bool attr_set[4] = {false, false, false, false};  // racing first calls are benign
      auto go = [&](void (*kern)(RealArgs<float>), int form) {
        if (!attr_set[form]) {
          FG_CHECK(hipFuncSetAttribute((const void*)kern,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 2 * SB_STAGE));
          attr_set[form] = true;
        }
        hipLaunchKernelGGL(kern, dim3((unsigned)big), dim3(256), 2 * SB_STAGE, st, p);
        FG_CHECK(hipGetLastError());
        return FERRAY_OK;
      };
      if (p.alpha_one)
        return p.beta_zero ? go(sgemm_big_kernel<true, true>, 0)
                           : go(sgemm_big_kernel<true, false>, 1);
      return p.beta_zero ? go(sgemm_big_kernel<false, true>, 2)
                         : go(sgemm_big_kernel<false, false>, 3);
    }
  }
  if (fast) {
    hipLaunchKernelGGL((gemm_real_kernel<T, CPLX, false>), dim3((unsigned)tiles), dim3(R_THREADS),
                       0, st, p);
    FG_CHECK(hipGetLastError());
    return FERRAY_OK;
  }
  // Ragged or unaligned: copy A, B (and C when it is read) into zero-padded
  // buffers of whole tiles, run the unguarded kernel there and copy C's
  // corner back.  The padding adds only zero products after each element's
  // own k-sequence, so the results equal the guarded kernel's bit for bit;
  // the guarded kernel remains for shapes where padding would add more than
  // 15 % work or the workspace cannot be had.
  using E = typename std::conditional<CPLX, CplxPair<T>, T>::type;
  const size_t Mp = cdiv(m, RB_M) * RB_M;
  const size_t Np = cdiv(n * mult, RB_N) * RB_N / mult;
  const size_t Kp = cdiv(k * mult, StageK<T>::value) * StageK<T>::value / mult;
  const double extra = (double)Mp * Np * Kp / ((double)m * n * k);
  void* ws = nullptr;
  if (extra <= 1.15 &&
      workspace(st, (Mp * Kp + Kp * Np + Mp * Np) * sizeof(E) + 3 * 256, &ws, WS_PAD) ==
          FERRAY_OK) {
    E* ap = reinterpret_cast<E*>(ws);
    E* bp = reinterpret_cast<E*>((uint8_t*)ap + cdiv(Mp * Kp * sizeof(E), 256) * 256);
    E* cp = reinterpret_cast<E*>((uint8_t*)bp + cdiv(Kp * Np * sizeof(E), 256) * 256);
    hipLaunchKernelGGL(pad_copy_kernel<E>, copy_grid(Mp, Kp), dim3(256), 0, st, ap,
                       reinterpret_cast<const E*>(a), m, k, lda, Mp, Kp);
    hipLaunchKernelGGL(pad_copy_kernel<E>, copy_grid(Kp, Np), dim3(256), 0, st, bp,
                       reinterpret_cast<const E*>(b), k, n, ldb, Kp, Np);
    if (!p.beta_zero)
      hipLaunchKernelGGL(pad_copy_kernel<E>, copy_grid(Mp, Np), dim3(256), 0, st, cp,
                         reinterpret_cast<const E*>(c), m, n, ldc, Mp, Np);
    FG_CHECK(hipGetLastError());
    RealArgs<T> q = p;
    q.a = reinterpret_cast<const T*>(ap);
    q.b = reinterpret_cast<const T*>(bp);
    q.c = reinterpret_cast<T*>(cp);
    q.M = (int)Mp;
    q.N = (int)(Np * mult);
    q.K = (int)(Kp * mult);
    q.lda = q.K;
    q.ldb = q.N;
    q.ldc = q.N;
    const size_t ptiles = (Mp / RB_M) * (Np * mult / RB_N);
    hipLaunchKernelGGL((gemm_real_kernel<T, CPLX, false>), dim3((unsigned)ptiles),
                       dim3(R_THREADS), 0, st, q);
    hipLaunchKernelGGL(unpad_copy_kernel<E>, copy_grid(m, n), dim3(256), 0, st,
                       reinterpret_cast<E*>(c), cp, m, n, ldc, Np);
    FG_CHECK(hipGetLastError());
    return FERRAY_OK;
  }
  (void)hipGetLastError();  // a failed workspace allocation is not this call's error
  hipLaunchKernelGGL((gemm_real_kernel<T, CPLX, true>), dim3((unsigned)tiles), dim3(R_THREADS),
                     0, st, p);
  FG_CHECK(hipGetLastError());
  return FERRAY_OK;
}

}  // namespace fg

using namespace fg;

extern "C" {

int ferray_gemm_f64_dev(size_t m, size_t n, size_t k, double alpha, const double* a,
                        const double* b, double beta, double* c, void* stream) {
  int s = ensure_device();
  if (s) return s;
  return launch_real<double, false>(m, n, k, alpha, 0.0, a, k, b, n, beta, 0.0, c, n,
                                    resolve_stream(stream));
}

int ferray_gemm_f32_dev(size_t m, size_t n, size_t k, float alpha, const float* a,
                        const float* b, float beta, float* c, void* stream) {
  int s = ensure_device();
  if (s) return s;
  return launch_real<float, false>(m, n, k, alpha, 0.0f, a, k, b, n, beta, 0.0f, c, n,
                                   resolve_stream(stream));
}

int ferray_gemm_c64_dev(size_t m, size_t n, size_t k, double alpha_re, double alpha_im,
                        const double* a, const double* b, double beta_re, double beta_im,
                        double* c, void* stream) {
  int s = ensure_device();
  if (s) return s;
  return launch_real<double, true>(m, n, k, alpha_re, alpha_im, a, k, b, n, beta_re, beta_im, c,
                                   n, resolve_stream(stream));
}

int ferray_gemm_c32_dev(size_t m, size_t n, size_t k, float alpha_re, float alpha_im,
                        const float* a, const float* b, float beta_re, float beta_im, float* c,
                        void* stream) {
  int s = ensure_device();
  if (s) return s;
  return launch_real<float, true>(m, n, k, alpha_re, alpha_im, a, k, b, n, beta_re, beta_im, c,
                                  n, resolve_stream(stream));
}

int ferray_gemm_f64_strided_dev(size_t m, size_t n, size_t k, double alpha, const double* a,
                                size_t lda, const double* b, size_t ldb, double beta, double* c,
                                size_t ldc, void* stream) {
  int s = ensure_device();
  if (s) return s;
  return launch_real<double, false>(m, n, k, alpha, 0.0, a, lda, b, ldb, beta, 0.0, c, ldc,
                                    resolve_stream(stream));
}

int ferray_gemm_f32_strided_dev(size_t m, size_t n, size_t k, float alpha, const float* a,
                                size_t lda, const float* b, size_t ldb, float beta, float* c,
                                size_t ldc, void* stream) {
  int s = ensure_device();
  if (s) return s;
  return launch_real<float, false>(m, n, k, alpha, 0.0f, a, lda, b, ldb, beta, 0.0f, c, ldc,
                                   resolve_stream(stream));
}

int ferray_gemm_c64_strided_dev(size_t m, size_t n, size_t k, double alpha_re, double alpha_im,
                                const double* a, size_t lda, const double* b, size_t ldb,
                                double beta_re, double beta_im, double* c, size_t ldc,
                                void* stream) {
  int s = ensure_device();
  if (s) return s;
  return launch_real<double, true>(m, n, k, alpha_re, alpha_im, a, lda, b, ldb, beta_re, beta_im,
                                   c, ldc, resolve_stream(stream));
}

int ferray_gemm_c32_strided_dev(size_t m, size_t n, size_t k, float alpha_re, float alpha_im,
                                const float* a, size_t lda, const float* b, size_t ldb,
                                float beta_re, float beta_im, float* c, size_t ldc,
                                void* stream) {
  int s = ensure_device();
  if (s) return s;
  return launch_real<float, true>(m, n, k, alpha_re, alpha_im, a, lda, b, ldb, beta_re, beta_im,
                                  c, ldc, resolve_stream(stream));
}

}  // extern "C"
