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
// The 256 x 256 shape normally runs a second main loop, mfma_nt_dma_kernel
// below: four waves, operands copied straight from HBM into LDS, and a
// rotated pipeline.  launch_mfma_nt picks between the two.
//
// The epilogue is a functor called once per output element with
// (row, col, plane, value); it owns all stores and bounds checks.
#pragma once
#include <type_traits>
#include <utility>

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
  int tiles_m, tiles_n;      // grid row z pairs A plane z with B plane z
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

// Groups of 4 tile rows: with 256 x 256 tiles at 8192 one group is one XCD's
// share; measured against 2, 8 and 16 rows: bf16 8192^3 5-6 % and i8 2 %
// faster than 8, the rest equal.
__device__ __forceinline__ void nt_tile(int tiles_m, int tiles_n, int& tm, int& tn) {
  const int G = 4;
  const int nt = tiles_m * tiles_n;
  int bid = blockIdx.x;
  if ((nt & 7) == 0) bid = (bid & 7) * (nt >> 3) + (bid >> 3);
  const int per_group = G * tiles_n, group = bid / per_group, first_m = group * G;
  const int gsz = min(tiles_m - first_m, G), in_group = bid - group * per_group;
  tm = first_m + in_group % gsz;
  tn = in_group / gsz;
}

// A copy of an accumulator made by a VALU move into an ordinary VGPR.  The
// NT kernels' epilogues take every result through one.  Left to itself the
// compiler stores results straight from the accumulator registers, and that
// form has written one wrong word per block tile and wave (the same constant
// in every tile, one result register) on some MI355X boards, where the
// copied form has not.
template <class T>
__device__ __forceinline__ T nt_vgpr(T x) {
  static_assert(sizeof(T) == 4, "one 32-bit register");
  T y;
  asm("v_mov_b32 %0, %1" : "=v"(y) : "v"(x));
  return y;
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
  const uint8_t* A = p.a + (size_t)z * p.a_plane + (size_t)tm * BM * p.lda;
  const uint8_t* B = p.bt + (size_t)z * p.b_plane + (size_t)tn * BN * p.ldb;
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
        epi(row0 + i * 32 + 8 * (r >> 2) + 4 * lk + (r & 3), col0 + j * 32 + lr, z,
            nt_vgpr(acc[i][j][r]));
}

// ---------------------------------------------------------------------------
// 256 x 256 tiles with the operands loaded straight into LDS.
//
// Four waves (2 x 2), each a 128 x 128 sub-tile, accumulators in AGPRs.  A
// stage is 128 bytes of k of both panels (64 KiB); two LDS buffers.  The
// global -> LDS copies are `global_load_lds_dwordx4`: no register staging and
// no LDS write instructions.  One such instruction writes 1 KiB of LDS at
// consecutive 16-byte slots (lane l at slot l), so the wave fills 8 rows of
// 128 bytes and each lane chooses which 16-byte chunk of its row it fetches:
// row r keeps chunk c at slot c ^ swz(r),
//   swz(r) = ((r >> 1) & 7) ^ ((r & 1) << 2)   (r taken mod 16).
// Chosen by timing ds_read_b128 with both fragment read patterns over
// candidate swizzles on the MI355X: this one reads at the full LDS rate for
// the 16 x 16 pattern (lanes 0-15 on rows 0-15 of one chunk, the other lane
// groups on the next chunks) and the 32 x 32 one, where (r & 7) ^ ((r >> 3) & 1)
// ran the 16 x 16 reads at half rate.
//
// The pipeline is rotated: a stage's last k-step is multiplied after the
// barrier that opens the next stage, interleaved with that stage's copy
// issue, while its first fragment reads are in flight.  The matrix pipe then
// never waits for the barrier or the first LDS reads of a stage.
//
// i8 uses the 16x16x64 MFMA: measured 2-3 % faster here than 32x32x32, and
// integer results are exact either way.  bf16 / f16 use the 16x16x32 MFMA:
// 9 % faster than 32x32x16 at 8192^3 and identical results there (each
// sums its bf16 products into the f32 accumulator the same way).  fp8 /
// bf8 use the NT kernel's 32 x 32 form.  Every kernel here keeps its
// accumulators in registers and stores results from VGPRs (nt_vgpr).
template <int ELEM>
struct NtDmaAcc {  // 32 x 32 tiles: the NT kernel's instruction
  using A = NtAcc<ELEM>;
  using type = typename A::type;
  using frag = typename A::frag;
  static constexpr int TS = 32, KSTEP = A::KSTEP, READS = A::READS;
  static __device__ __forceinline__ type step(frag a, frag b, type c) { return A::step(a, b, c); }
};
template <>
struct NtDmaAcc<ELEM_I8> {
  using type = i4;
  using frag = i4;
  static constexpr int TS = 16, KSTEP = 64, READS = 1;
  static __device__ __forceinline__ type step(i4 a, i4 b, type c) {
    return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
  }
};
template <>
struct NtDmaAcc<ELEM_BF16> {
  using type = f4;
  using frag = i4;
  static constexpr int TS = 16, KSTEP = 64, READS = 1;
  static __device__ __forceinline__ type step(i4 a, i4 b, type c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a),
                                                   __builtin_bit_cast(bf8, b), c, 0, 0, 0);
  }
};
template <>
struct NtDmaAcc<ELEM_F16> {
  using type = f4;
  using frag = i4;
  static constexpr int TS = 16, KSTEP = 64, READS = 1;
  static __device__ __forceinline__ type step(i4 a, i4 b, type c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(h8, a),
                                                  __builtin_bit_cast(h8, b), c, 0, 0, 0);
  }
};

// An epilogue may offer an unchecked form for block tiles inside C:
// `bool interior(row0, col0, rows, cols)` and `with_fast(f)`, which calls f
// with a functor that stores without bounds checks, its run-time flags
// turned into template parameters.
template <class E, class = void>
struct NtFastEpi : std::false_type {};
template <class E>
struct NtFastEpi<E, std::void_t<decltype(std::declval<const E&>().interior(0, 0, 0, 0))>>
    : std::true_type {};

typedef __attribute__((address_space(1))) void nt_gvoid;
typedef __attribute__((address_space(3))) void nt_lvoid;

template <int ELEM, class Epi>
__global__ __launch_bounds__(256) void mfma_nt_dma_kernel(NtOperands p, Epi epi) {
  using Acc = NtDmaAcc<ELEM>;
  using acc_t = typename Acc::type;
  using frag = typename Acc::frag;
  constexpr int BM = 256, BN = 256, KB = 128, TS = Acc::TS, TI = 128 / TS, TJ = 128 / TS;
  constexpr int KSTEP = Acc::KSTEP, NKK = KB / KSTEP, CH = KSTEP / 16, RD = Acc::READS;
  constexpr int STAGE = (BM + BN) * KB;
  constexpr int NMF = TI * TJ, NRD = (TI + TJ) * RD, NLD = 16;
  constexpr int ROUT = TS == 32 ? 16 : 4;  // accumulator registers per tile
  static_assert(NKK >= 2, "the rotated pipeline needs two k-steps per stage");
  extern __shared__ __attribute__((aligned(16))) uint8_t nt_smem[];

  int tm, tn;
  nt_tile(p.tiles_m, p.tiles_n, tm, tn);
  const int z = blockIdx.y;
  const uint8_t* A = p.a + (size_t)z * p.a_plane + (size_t)tm * BM * p.lda;
  const uint8_t* B = p.bt + (size_t)z * p.b_plane + (size_t)tn * BN * p.ldb;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm0 = (wave >> 1) * 128, wn0 = (wave & 1) * 128;
  const int lr = lane & (TS - 1), lk = lane / TS;

  // Copy instruction q of wave w fills LDS rows 8 (4q + w) .. + 7 of each
  // panel: lane l writes slot l & 7 of row l >> 3.  Offsets within the tile
  // are 32-bit (the launcher checks 256 rows fit).
  const int lrow = lane >> 3, lslot = lane & 7;
  // swz of row 8 (4q + w) + lrow, whose low four bits are 8 (w & 1) + lrow
  const int gchunk = lslot ^ ((4 * (wave & 1) + (lrow >> 1)) ^ ((lrow & 1) << 2));
  const unsigned a_off = (unsigned)((wave * 8 + lrow) * p.lda + gchunk * 16);
  const unsigned b_off = (unsigned)((wave * 8 + lrow) * p.ldb + gchunk * 16);
  const unsigned a_q = (unsigned)(32 * p.lda), b_q = (unsigned)(32 * p.ldb);
  auto issue = [&](int buf, int kb) __attribute__((always_inline)) {
    uint8_t* as = nt_smem + buf * STAGE + wave * 1024;
    uint8_t* bs = as + BM * KB;
    const uint8_t* ga = A + kb;
    const uint8_t* gb = B + kb;
#pragma unroll
    for (int q = 0; q < 8; ++q)
      __builtin_amdgcn_global_load_lds((nt_gvoid*)(ga + (a_off + q * a_q)),
                                       (nt_lvoid*)(as + q * 4096), 16, 0, 0);
#pragma unroll
    for (int q = 0; q < 8; ++q)
      __builtin_amdgcn_global_load_lds((nt_gvoid*)(gb + (b_off + q * b_q)),
                                       (nt_lvoid*)(bs + q * 4096), 16, 0, 0);
  };

  // Fragment of k-step kk: lane l reads row l % TS of the tile, 16-byte
  // chunk(s) kk * CH + (TS == 32 ? h * 2 + l / 32 : l / 16) (h = second read
  // of an fp8 fragment), at its swizzled slot.
  const int swz = ((lr >> 1) & 7) ^ ((lr & 1) << 2);
  int xo[NKK][RD];
#pragma unroll
  for (int kk = 0; kk < NKK; ++kk)
#pragma unroll
    for (int h = 0; h < RD; ++h) xo[kk][h] = ((kk * CH + h * 2 + lk) ^ swz) << 4;
  const int arow = (wm0 + lr) * KB, brow = BM * KB + (wn0 + lr) * KB;
  auto read = [&](const uint8_t* base, int kk) __attribute__((always_inline)) {
    if constexpr (RD == 1) {
      return *reinterpret_cast<const i4*>(base + xo[kk][0]);
    } else {
      const i4 lo = *reinterpret_cast<const i4*>(base + xo[kk][0]);
      const i4 hi = *reinterpret_cast<const i4*>(base + xo[kk][RD - 1]);
      return i8v{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
    }
  };
  auto read_step = [&](const uint8_t* s, int kk, frag* fa, frag* fb) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < TI; ++i) fa[i] = read(s + arow + i * TS * KB, kk);
#pragma unroll
    for (int j = 0; j < TJ; ++j) fb[j] = read(s + brow + j * TS * KB, kk);
  };

  acc_t acc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) acc[i][j] = acc_t{};
  auto mma = [&](const frag* fa, const frag* fb) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j) acc[i][j] = Acc::step(fa[i], fb[j], acc[i][j]);
  };

  frag pa[TI], pb[TJ];  // last k-step of the previous stage
  // One stage from LDS buffer `cur`.  PEND: first multiply the previous
  // stage's last k-step.  MORE: copy the next stage (k offset kb_next) into
  // buffer cur ^ 1, whose reads all finished before the opening barrier.
  // The first fragment reads precede the copies in program order: the
  // compiler cannot tell the buffers apart and keeps LDS accesses in order.
  auto body = [&](auto pend_c, auto more_c, int cur, int kb_next) __attribute__((always_inline)) {
    constexpr bool PEND = decltype(pend_c)::value, MORE = decltype(more_c)::value;
    const uint8_t* s = nt_smem + cur * STAGE;
    frag fa[NKK - 1][TI], fb[NKK - 1][TJ];
    read_step(s, 0, fa[0], fb[0]);
    if (MORE) issue(cur ^ 1, kb_next);
    if (PEND) mma(pa, pb);
#pragma unroll
    for (int kk = 0; kk + 1 < NKK; ++kk) {
      if (kk + 2 < NKK) read_step(s, kk + 1, fa[kk + 1], fb[kk + 1]);
      else read_step(s, kk + 1, pa, pb);
      mma(fa[kk], fb[kk]);
    }
    // Order for the scheduler (masks: 0x010 VMEM, 0x100 DS read, 0x008
    // MFMA): the first reads, the pending MFMAs with the copies between
    // them, then each k-step's MFMAs with the next k-step's reads.
    __builtin_amdgcn_sched_group_barrier(0x100, NRD, 0);
    if (PEND) {
#pragma unroll
      for (int q = 0; q < NMF; ++q) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        if (MORE && q % (NMF > NLD ? NMF / NLD : 1) == 0)
          __builtin_amdgcn_sched_group_barrier(0x010, NLD > NMF ? NLD / NMF : 1, 0);
      }
    } else if (MORE) {
      __builtin_amdgcn_sched_group_barrier(0x010, NLD, 0);
    }
#pragma unroll
    for (int kk = 0; kk + 1 < NKK; ++kk) {
#pragma unroll
      for (int q = 0; q < NRD; ++q) {
        __builtin_amdgcn_sched_group_barrier(0x008, NMF / NRD > 0 ? NMF / NRD : 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
      }
    }
  };
  // The barrier between stages: every wave's copies have landed (vmcnt) and
  // its fragment reads have returned (lgkmcnt) before any wave copies into
  // the buffer those reads came from.  The waitcnt builtin keeps the
  // compiler's own count of outstanding reads right.
  auto barrier = [&]() __attribute__((always_inline)) {
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
    __syncthreads();
    __builtin_amdgcn_sched_barrier(0);
  };
  using T = std::integral_constant<bool, true>;
  using F = std::integral_constant<bool, false>;

  const int nk = p.kbytes / KB;
  // The last k-step is multiplied on each path separately: with the two
  // paths joined before it, the register allocator shuffled and spilled
  // accumulators there (the 16 x 16 bf16 / f16 forms).
  auto kloop = [&]() __attribute__((always_inline)) {
    issue(0, 0);
    __syncthreads();
    if (nk > 1) {
      body(F{}, T{}, 0, KB);
      int kt = 1;
      for (; kt + 1 < nk; ++kt) {
        barrier();
        body(T{}, T{}, kt & 1, (kt + 1) * KB);
      }
      barrier();
      body(T{}, F{}, kt & 1, 0);
      mma(pa, pb);
    } else {
      body(F{}, F{}, 0, 0);
      mma(pa, pb);
    }
  };
  kloop();

  // Result element r of lane l: 32 x 32 tiles as in the NT kernel; 16 x 16
  // tiles row 4 (l / 16) + r, column l % 16.
  const int row0 = tm * BM + wm0, col0 = tn * BN + wn0;
  auto store_all = [&](const auto& e) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j)
#pragma unroll
        for (int r = 0; r < ROUT; ++r) {
          const auto v = nt_vgpr(acc[i][j][r]);
          if constexpr (TS == 32)
            e(row0 + i * 32 + 8 * (r >> 2) + 4 * lk + (r & 3), col0 + j * 32 + lr, z, v);
          else
            e(row0 + i * 16 + 4 * lk + r, col0 + j * 16 + lr, z, v);
        }
  };
  // Block tiles wholly inside C take the epilogue's unchecked form when it
  // has one: no per-element bounds or flag branches, so the compiler can
  // issue the epilogue's loads together.
  if constexpr (NtFastEpi<Epi>::value) {
    if (epi.interior(tm * BM, tn * BN, BM, BN))
      epi.with_fast(store_all);
    else
      store_all(epi);
  } else {
    store_all(epi);
  }
}

// Launches the NT main loop over `planes` independent plane pairs.  The
// 256 x 256 shape runs the direct-to-LDS kernel when 256 rows of either
// operand span less than 4 GiB (its in-tile offsets are 32-bit) and A needs
// no sign flip (the copies move bytes unmodified; flipping fragments after
// the reads spills registers).
template <int ELEM, class S, bool XOR_A, class Epi>
inline int launch_mfma_nt(const NtOperands& p, int planes, const Epi& epi, hipStream_t st) {
  const size_t tiles = (size_t)p.tiles_m * p.tiles_n;
  if (tiles == 0 || planes <= 0) return FERRAY_OK;
  if (tiles > 0x7fffffff || planes > 65535) return FERRAY_ERR_INVALID;
  if (p.kbytes <= 0 || p.kbytes % S::KB) return FERRAY_ERR_INVALID;
  const dim3 grid((unsigned)tiles, (unsigned)planes);
  if constexpr (S::BM == 256 && S::BN == 256 && S::KB == 128 && !XOR_A) {
    if (256 * (p.lda > p.ldb ? p.lda : p.ldb) < ((size_t)1 << 32)) {
      auto kern = mfma_nt_dma_kernel<ELEM, Epi>;
      constexpr int lds = 2 * 512 * 128;
      static bool attr_set = false;  // every caller sets the same value; a race is benign
      if (!attr_set) {
        FG_CHECK(hipFuncSetAttribute((const void*)kern,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, lds));
        attr_set = true;
      }
      hipLaunchKernelGGL(kern, grid, dim3(256), lds, st, p, epi);
      FG_CHECK(hipGetLastError());
      return FERRAY_OK;
    }
  }
  auto kern = mfma_nt_kernel<ELEM, S, XOR_A, Epi>;
  static bool attr_set = false;
  if (!attr_set) {
    FG_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)S::lds_bytes()));
    attr_set = true;
  }
  hipLaunchKernelGGL(kern, grid, dim3(S::THREADS), S::lds_bytes(), st, p, epi);
  FG_CHECK(hipGetLastError());
  return FERRAY_OK;
}

}  // namespace fg
