// The "NT" matrix-core main loop shared by the integer, half-precision and
// split-precision paths: both operands are k-contiguous byte planes (A is
// M x K row-major, B is supplied transposed as N x K), zero padded to whole
// tiles, so the main loop never bounds-checks a load.
//
//   i8:        v_mfma_i32_32x32x32_i8          (32 bytes of k per step, i32 acc)
//   bf16/f16:  v_mfma_f32_32x32x16_{bf16,f16}  (same byte layout, f32 acc)
//   fp8/bf8:   v_mfma_f32_32x32x64_f8f6f4      (64 bytes of k per step, f32 acc;
//              OCP E4M3 / E5M2 bytes)
//
// A k-step's fragment for lane l is row l%32 of the panel: the 16 bytes at
// 16*(l/32) of each 32-byte k group (one group for i8/bf16/f16, two for fp8).
// A and B fragments are read with the same byte map, so the k pairing of the
// two operands is the identity whatever order the matrix core sums in.
//
// Block: WM x WN waves, block tile BM x BN, each wave a (BM/WM) x (BN/WN)
// sub-tile of 32x32 MFMA tiles.  A pipeline stage moves KB bytes of k of both
// panels from HBM through registers into double-buffered LDS (rows padded by
// 16 bytes); one barrier per stage.  The stage body is given to the scheduler
// as an explicit pipeline: this stage's first-step fragment reads, then MFMAs
// interleaved with the fragment reads of the following k-step (the first
// two k-steps also carry the next stage's global loads: measured 3-8 % faster
// than issuing them all up front), and the last step's MFMAs interleaved with
// the LDS writes of the next stage, so the waves of a SIMD keep the matrix
// pipe fed.  Tiles are visited in GROUP_M-row groups and the block id is
// remapped so that each XCD (blocks are dealt round-robin to the 8 XCDs)
// runs a contiguous range of that order, keeping the panels a group shares
// inside one XCD's L2.
//
// The epilogue is a functor called once per output element with
// (row, col, plane, value); it owns all stores and bounds checks.
#pragma once
#include "common.hpp"

namespace fg {

typedef int i16v __attribute__((ext_vector_type(16)));
typedef int i8v __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef unsigned u4v __attribute__((ext_vector_type(4)));

enum MfmaElem { ELEM_I8 = 0, ELEM_BF16 = 1, ELEM_F16 = 2, ELEM_FP8 = 3, ELEM_BF8 = 4 };

struct NtOperands {
  const uint8_t* a;         // plane 0 of A; plane z starts at a + z * a_plane
  const uint8_t* bt;        // plane 0 of B^T
  size_t a_plane, b_plane;  // plane strides in bytes
  size_t lda, ldb;          // row pitch in bytes
  int kbytes;               // bytes of k to reduce, a multiple of the stage depth
  int tiles_m, tiles_n;
  int cross;                // 0: grid row z pairs A plane z with B plane z;
                            // 1: it pairs A plane z / 2 with B plane z % 2
};

// One 32-byte k group per step: a 16-byte fragment per lane.
struct NtFrag16 {
  using frag = i4;
  static constexpr int KSTEP = 32, READS = 1;
  static __device__ __forceinline__ frag read(const uint8_t* p) {
    return *reinterpret_cast<const i4*>(p);
  }
};
// Two 32-byte k groups per step: a 32-byte fragment per lane.
struct NtFrag32 {
  using frag = i8v;
  static constexpr int KSTEP = 64, READS = 2;
  static __device__ __forceinline__ frag read(const uint8_t* p) {
    const i4 lo = *reinterpret_cast<const i4*>(p);
    const i4 hi = *reinterpret_cast<const i4*>(p + 32);
    return i8v{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
  }
};

template <int ELEM>
struct NtAcc;
template <>
struct NtAcc<ELEM_I8> : NtFrag16 {
  using type = i16v;
  static __device__ __forceinline__ type step(i4 a, i4 b, type c) {
    return __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0);
  }
};
template <>
struct NtAcc<ELEM_BF16> : NtFrag16 {
  using type = f16v;
  static __device__ __forceinline__ type step(i4 a, i4 b, type c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf8, a),
                                                   __builtin_bit_cast(bf8, b), c, 0, 0, 0);
  }
};
template <>
struct NtAcc<ELEM_F16> : NtFrag16 {
  using type = f16v;
  static __device__ __forceinline__ type step(i4 a, i4 b, type c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(h8, a),
                                                  __builtin_bit_cast(h8, b), c, 0, 0, 0);
  }
};
// Format operands of the f8f6f4 instruction: 0 = E4M3, 1 = E5M2.  Scale
// operands 0 select the unscaled form.
template <>
struct NtAcc<ELEM_FP8> : NtFrag32 {
  using type = f16v;
  static __device__ __forceinline__ type step(i8v a, i8v b, type c) {
    return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c, 0, 0, 0, 0, 0, 0);
  }
};
template <>
struct NtAcc<ELEM_BF8> : NtFrag32 {
  using type = f16v;
  static __device__ __forceinline__ type step(i8v a, i8v b, type c) {
    return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c, 1, 1, 0, 0, 0, 0);
  }
};

// Shape of one NT configuration.
template <int BM_, int BN_, int WM_, int WN_, int KB_>
struct NtShape {
  static constexpr int BM = BM_, BN = BN_, WM = WM_, WN = WN_, KB = KB_;
  static constexpr int THREADS = WM * WN * 64;
  static constexpr int PITCH = KB + 16;
  static constexpr size_t lds_bytes() { return (size_t)2 * (BM + BN) * PITCH; }
};
using NtBig = NtShape<256, 256, 4, 2, 128>;   // 512 threads, 144 KiB of LDS
using NtSmall = NtShape<128, 128, 2, 2, 64>;  // 256 threads, 40 KiB of LDS

__device__ __forceinline__ void nt_tile(int tiles_m, int tiles_n, int& tm, int& tn) {
  const int G = 8;
  const int nt = tiles_m * tiles_n;
  int bid = blockIdx.x;
  if ((nt & 7) == 0) bid = (bid & 7) * (nt >> 3) + (bid >> 3);
  const int per_group = G * tiles_n, group = bid / per_group, first_m = group * G;
  const int gsz = min(tiles_m - first_m, G), in_group = bid - group * per_group;
  tm = first_m + in_group % gsz;
  tn = in_group / gsz;
}

// Result element r of lane l in a 32x32 tile: row 8*(r/4) + 4*(l/32) + r%4,
// column l%32.  XOR_A flips the sign bit of every A byte (u8 -> u8 - 128).
template <int ELEM, class S, bool XOR_A, class Epi>
__global__ __launch_bounds__(S::THREADS) void mfma_nt_kernel(NtOperands p, Epi epi) {
  using Acc = NtAcc<ELEM>;
  using acc_t = typename Acc::type;
  constexpr int BM = S::BM, BN = S::BN, KB = S::KB, PITCH = S::PITCH, NT = S::THREADS;
  constexpr int TI = BM / S::WM / 32, TJ = BN / S::WN / 32;
  constexpr int CPR = KB / 16;  // 16-byte chunks per row per stage
  constexpr int AV = BM * CPR / NT, BV = BN * CPR / NT;
  constexpr int KSTEP = Acc::KSTEP;
  constexpr int NKK = KB / KSTEP;
  constexpr int NMF = TI * TJ, NRD = (TI + TJ) * Acc::READS, NWR = AV + BV;
  static_assert(KB % KSTEP == 0, "stage depth");
  static_assert(AV >= 1 && BV >= 1 && BM * CPR % NT == 0 && BN * CPR % NT == 0, "loader");
  extern __shared__ __attribute__((aligned(16))) uint8_t nt_smem[];
  uint8_t* As = nt_smem;
  uint8_t* Bs = nt_smem + 2 * BM * PITCH;

  int tm, tn;
  nt_tile(p.tiles_m, p.tiles_n, tm, tn);
  const int z = blockIdx.y;
  const int za = p.cross ? z >> 1 : z, zb = p.cross ? z & 1 : z;
  const uint8_t* A = p.a + (size_t)za * p.a_plane + (size_t)tm * BM * p.lda;
  const uint8_t* B = p.bt + (size_t)zb * p.b_plane + (size_t)tn * BN * p.ldb;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave / S::WN) * (BM / S::WM), wn0 = (wave % S::WN) * (BN / S::WN);
  const int lr = lane & 31, lk = lane >> 5;

  u4v ra[AV], rb[BV];
  auto load = [&](int kb) {
#pragma unroll
    for (int q = 0; q < AV; ++q) {
      const int v = tid + q * NT, row = v / CPR, c = v % CPR;
      ra[q] = *reinterpret_cast<const u4v*>(A + (size_t)row * p.lda + kb + c * 16);
    }
#pragma unroll
    for (int q = 0; q < BV; ++q) {
      const int v = tid + q * NT, row = v / CPR, c = v % CPR;
      rb[q] = *reinterpret_cast<const u4v*>(B + (size_t)row * p.ldb + kb + c * 16);
    }
  };
  auto store = [&](int buf) {
    uint8_t* as = As + buf * BM * PITCH;
    uint8_t* bs = Bs + buf * BN * PITCH;
#pragma unroll
    for (int q = 0; q < AV; ++q) {
      const int v = tid + q * NT, row = v / CPR, c = v % CPR;
      u4v x = ra[q];
      if (XOR_A) x ^= 0x80808080u;
      *reinterpret_cast<u4v*>(as + row * PITCH + c * 16) = x;
    }
#pragma unroll
    for (int q = 0; q < BV; ++q) {
      const int v = tid + q * NT, row = v / CPR, c = v % CPR;
      *reinterpret_cast<u4v*>(bs + row * PITCH + c * 16) = rb[q];
    }
  };

  acc_t acc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) acc[i][j] = acc_t{};

  auto stage = [&](int cur, bool more, int kb_next) {
    const uint8_t* as = As + cur * BM * PITCH;
    const uint8_t* bs = Bs + cur * BN * PITCH;
    if (more) load(kb_next);
    typename Acc::frag af[NKK][TI], bf[NKK][TJ];
#pragma unroll
    for (int kk = 0; kk < NKK; ++kk) {
#pragma unroll
      for (int i = 0; i < TI; ++i)
        af[kk][i] = Acc::read(as + (wm0 + i * 32 + lr) * PITCH + kk * KSTEP + lk * 16);
#pragma unroll
      for (int j = 0; j < TJ; ++j)
        bf[kk][j] = Acc::read(bs + (wn0 + j * 32 + lr) * PITCH + kk * KSTEP + lk * 16);
    }
#pragma unroll
    for (int kk = 0; kk < NKK; ++kk)
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) acc[i][j] = Acc::step(af[kk][i], bf[kk][j], acc[i][j]);
    if (more) store(cur ^ 1);
    // Pipeline order for the scheduler (masks: 0x020 VMEM read, 0x100 DS
    // read, 0x200 DS write, 0x008 MFMA).
    __builtin_amdgcn_sched_group_barrier(0x100, NRD, 0);
#pragma unroll
    for (int kk = 0; kk < NKK - 1; ++kk) {
#pragma unroll
      for (int q = 0; q < NRD; ++q) {
        __builtin_amdgcn_sched_group_barrier(0x008, NMF / NRD > 0 ? NMF / NRD : 1, 0);
        if (more && kk < 2)  // next stage's global loads over the first two k-steps
          __builtin_amdgcn_sched_group_barrier(0x020, (NWR + 2 * NRD - 1) / (2 * NRD), 0);
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

  const int nk = p.kbytes / KB;
  load(0);
  store(0);
  __syncthreads();
  int kt = 0;
  for (; kt + 1 < nk; ++kt) {
    stage(kt & 1, true, (kt + 1) * KB);
    __syncthreads();
  }
  stage(kt & 1, false, 0);

  const int row0 = tm * BM + wm0, col0 = tn * BN + wn0;
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r)
        epi(row0 + i * 32 + 8 * (r >> 2) + 4 * lk + (r & 3), col0 + j * 32 + lr, z, acc[i][j][r]);
}

// Launches the NT main loop over `planes` independent plane pairs.
template <int ELEM, class S, bool XOR_A, class Epi>
inline int launch_mfma_nt(const NtOperands& p, int planes, const Epi& epi, hipStream_t st) {
  auto kern = mfma_nt_kernel<ELEM, S, XOR_A, Epi>;
  static bool attr_set = false;  // every caller sets the same value; a race is benign
  if (!attr_set) {
    FG_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)S::lds_bytes()));
    attr_set = true;
  }
  const size_t tiles = (size_t)p.tiles_m * p.tiles_n;
  if (tiles == 0 || planes <= 0) return FERRAY_OK;
  if (tiles > 0x7fffffff || planes > 65535) return FERRAY_ERR_INVALID;
  if (p.kbytes <= 0 || p.kbytes % S::KB) return FERRAY_ERR_INVALID;
  hipLaunchKernelGGL(kern, dim3((unsigned)tiles, (unsigned)planes), dim3(S::THREADS),
                     S::lds_bytes(), st, p, epi);
  FG_CHECK(hipGetLastError());
  return FERRAY_OK;
}

}  // namespace fg
