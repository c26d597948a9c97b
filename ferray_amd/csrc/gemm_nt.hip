// Integer and half-precision GEMM entry points on the NT matrix-core main
// loop (mfma_nt.hpp).
//
//   i8 x i8 -> i32    one i8 pass                              (gemm_i8_signed)
//   u8 x i8 -> i32    A's sign bit flipped on load (a - 128), plus 128*colsum(B)
//                                                                      (gemm_i8)
//   i16 x i16 -> i32  the four i8 byte-plane products of a tile in one block
//                                                                     (gemm_i16)
//   quantized_matmul  the u8 x i8 pass with a fused zero-point / scale epilogue
//   bf16/f16 -> f32   one half-precision pass                 (gemm_{bf16,f16}_f32)
//   fp8 -> f32        one pass on the 64-deep f8f6f4 instruction, OCP E4M3 or
//                     E5M2 bytes (no reference entry point; the fp8 member of
//                     the same family)
//
// B is transposed into k-contiguous planes by a prepass (which also forms the
// i16 byte planes and the wrapping column sums); A is read in place when its
// shape allows, otherwise packed the same way (i16 A is always split into its
// two byte planes).  Padding is zero.
//
// Integer arithmetic is exact mod 2^32:
//   u8:  sum a*b = sum (a-128)*b + 128*sum_k b[k][j]
//   i16: a = 256*ah + al' + 128 with ah = a>>8, al' = (a & 255) - 128, so
//        sum a*b = 65536*S(ah,bh) + 256*(S(ah,bl') + S(al',bh)) + S(al',bl')
//                  + 128*rowsum(A) + 128*colsum(B) - 16384*k
// where every S(.,.) is an i8 x i8 MFMA product; all terms wrap in uint32.
// K is reduced in chunks of at most 2^16 so that no i32 accumulator can
// leave the i32 range whatever the operands (|a*b| <= 2^14 per term).
#include <algorithm>
#include <mutex>

#include "mfma_nt.hpp"

namespace fg {

enum NtKind {
  NT_I8S = 0, NT_U8 = 1, NT_I16 = 2, NT_BF16 = 3, NT_F16 = 4, NT_QUANT = 5, NT_E4M3 = 6,
  NT_E5M2 = 7
};

constexpr size_t NT_K_CHUNK = 1u << 16;

// ---------------------------------------------------------------------
// Epilogues
// ---------------------------------------------------------------------

// i32 output of the i8 products: v = acc + 128*colsum[col] (if colsum; u8 A
// was shifted by -128) + (old C if read_c), all wrapping.
template <bool READ_C, bool COLSUM>
struct EpiI32Fast {  // EpiI32 inside C, flags fixed
  int32_t* c;
  int N;
  const int32_t* colsum;
  __device__ __forceinline__ void operator()(int row, int col, int, int acc) const {
    uint32_t v = (uint32_t)acc;
    if (COLSUM) v += 128u * (uint32_t)colsum[col];
    int32_t* dst = c + (size_t)row * N + col;
    if (READ_C) v += (uint32_t)*dst;
    *dst = (int32_t)v;
  }
};

struct EpiI32 {
  int32_t* c;
  int M, N;
  int read_c;
  const int32_t* colsum;   // u8: 128*colsum(B)
  __device__ __forceinline__ void operator()(int row, int col, int, int acc) const {
    if (row >= M || col >= N) return;
    uint32_t v = (uint32_t)acc;
    if (colsum) v += 128u * (uint32_t)colsum[col];
    int32_t* dst = c + (size_t)row * N + col;
    if (read_c) v += (uint32_t)*dst;
    *dst = (int32_t)v;
  }
  __device__ __forceinline__ bool interior(int r0, int c0, int rows, int cols) const {
    return r0 + rows <= M && c0 + cols <= N;
  }
  template <class F>
  __device__ __forceinline__ void with_fast(F&& f) const {
    if (read_c) {
      if (colsum) f(EpiI32Fast<true, true>{c, N, colsum});
      else f(EpiI32Fast<true, false>{c, N, colsum});
    } else {
      if (colsum) f(EpiI32Fast<false, true>{c, N, colsum});
      else f(EpiI32Fast<false, false>{c, N, colsum});
    }
  }
};

// quantized_matmul (mod.rs:524-573): raw = acc + 128*csb (A was shifted by
// -128); corrected = raw - azp*csb - bzp*rsa + k*azp*bzp (wrapping);
// out = (scale_a*scale_b) * (float)corrected, written or added.
template <bool ACC>
struct EpiQuantFast {  // EpiQuant inside C, accumulate fixed
  float* c;
  int N;
  const int32_t* colsum;
  const int32_t* rowsum;
  uint32_t azp, bzp, kab;
  float scale;
  __device__ __forceinline__ void operator()(int row, int col, int, int acc) const {
    const uint32_t cs = (uint32_t)colsum[col];
    const uint32_t v = (uint32_t)acc + 128u * cs - azp * cs - bzp * (uint32_t)rowsum[row] + kab;
    float out = scale * (float)(int32_t)v;
    float* dst = c + (size_t)row * N + col;
    if (ACC) out = *dst + out;
    *dst = out;
  }
};

struct EpiQuant {
  float* c;
  int M, N;
  int accumulate;
  const int32_t* colsum;
  const int32_t* rowsum;
  uint32_t azp, bzp, kab;
  float scale;
  __device__ __forceinline__ void operator()(int row, int col, int, int acc) const {
    if (row >= M || col >= N) return;
    const uint32_t cs = (uint32_t)colsum[col];
    const uint32_t raw = (uint32_t)acc + 128u * cs;
    const uint32_t v = raw - azp * cs - bzp * (uint32_t)rowsum[row] + kab;
    float out = scale * (float)(int32_t)v;
    float* dst = c + (size_t)row * N + col;
    if (accumulate) out = *dst + out;
    *dst = out;
  }
  __device__ __forceinline__ bool interior(int r0, int c0, int rows, int cols) const {
    return r0 + rows <= M && c0 + cols <= N;
  }
  template <class F>
  __device__ __forceinline__ void with_fast(F&& f) const {
    if (accumulate)
      f(EpiQuantFast<true>{c, N, colsum, rowsum, azp, bzp, kab, scale});
    else
      f(EpiQuantFast<false>{c, N, colsum, rowsum, azp, bzp, kab, scale});
  }
};

template <bool ACC>
struct EpiF32Fast {  // EpiF32 inside C, accumulate fixed
  float* c;
  int N;
  __device__ __forceinline__ void operator()(int row, int col, int, float acc) const {
    float* dst = c + (size_t)row * N + col;
    *dst = ACC ? *dst + acc : acc;
  }
};

struct EpiF32 {
  float* c;
  int M, N;
  int accumulate;
  __device__ __forceinline__ void operator()(int row, int col, int, float acc) const {
    if (row >= M || col >= N) return;
    float* dst = c + (size_t)row * N + col;
    *dst = accumulate ? *dst + acc : acc;
  }
  __device__ __forceinline__ bool interior(int r0, int c0, int rows, int cols) const {
    return r0 + rows <= M && c0 + cols <= N;
  }
  template <class F>
  __device__ __forceinline__ void with_fast(F&& f) const {
    if (accumulate)
      f(EpiF32Fast<true>{c, N});
    else
      f(EpiF32Fast<false>{c, N});
  }
};

// ---------------------------------------------------------------------
// Prepasses
// ---------------------------------------------------------------------

// Bt planes from row-major B (K x N) of element E, tile 64 (k) x 64 (n),
// grid (Np / 64, Kp / 64) launched in y slices from y0, zero padded to Np x Kp:
//   MAP 0: byte copy      MAP 1: i16 -> hi byte plane + (lo - 128) plane
//   MAP 2: 16-bit copy
// colsum (optional) accumulates the wrapping column sums of B.
template <typename E, int MAP>
__global__ __launch_bounds__(256) void pack_bt_kernel(const E* __restrict__ b, int K, int N,
                                                      int Kp, int Np, uint8_t* out0,
                                                      uint8_t* out1, int32_t* colsum, int y0) {
  __shared__ E tile[64][64 + 4];
  const int tid = threadIdx.x;
  const int n0 = blockIdx.x * 64, k0 = (y0 + (int)blockIdx.y) * 64;
  {
    const int row = tid >> 2, c0 = (tid & 3) * 16;
    const int k = k0 + row;
    const E* src = b + (size_t)k * N + n0 + c0;
    if (k < K && n0 + c0 + 16 <= N && ((uintptr_t)src % 16) == 0) {
      // 16 in-range elements as 16-byte vector loads
      E v[16];
#pragma unroll
      for (int q = 0; q < (int)(16 * sizeof(E) / 16); ++q)
        *reinterpret_cast<u4v*>(&v[q * 16 / sizeof(E)]) = reinterpret_cast<const u4v*>(src)[q];
#pragma unroll
      for (int e = 0; e < 16; ++e) tile[row][c0 + e] = v[e];
    } else {
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int n = n0 + c0 + e;
        tile[row][c0 + e] = (k < K && n < N) ? b[(size_t)k * N + n] : E(0);
      }
    }
  }
  __syncthreads();
  if (colsum && tid < 64 && n0 + tid < N) {
    uint32_t s = 0;
    for (int r = 0; r < 64; ++r) s += (uint32_t)(int32_t)tile[r][tid];
    atomicAdd(reinterpret_cast<unsigned int*>(colsum + n0 + tid), s);
  }
  const int nl = tid >> 2, kc = (tid & 3) * 16;
  const int n = n0 + nl, kbase = k0 + kc;
  if (n >= Np || kbase >= Kp) return;
  if (MAP == 2) {
    uint16_t v[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = (uint16_t)tile[kc + e][nl];
    uint16_t* dst = reinterpret_cast<uint16_t*>(out0) + (size_t)n * Kp + kbase;
#pragma unroll
    for (int e = 0; e < 16; ++e) dst[e] = v[e];
  } else {
    uint8_t v0[16], v1[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const bool real = (k0 + kc + e < K) && (n < N);
      if (MAP == 0) {
        v0[e] = (uint8_t)tile[kc + e][nl];
      } else {
        const int x = (int)tile[kc + e][nl];
        v0[e] = real ? (uint8_t)(int8_t)(x >> 8) : 0;
        v1[e] = real ? (uint8_t)(int8_t)((x & 255) - 128) : 0;
      }
    }
    *reinterpret_cast<u4v*>(out0 + (size_t)n * Kp + kbase) = __builtin_bit_cast(u4v, v0);
    if (MAP == 1) *reinterpret_cast<u4v*>(out1 + (size_t)n * Kp + kbase) = __builtin_bit_cast(u4v, v1);
  }
}

// B^T of row-major B (K x N) for tile-aligned shapes: ESZ-byte elements
// (1: any 8-bit type, 2: bf16 / f16), K a multiple of 128, N of 128 / ESZ,
// rows 16-byte aligned, no padding.  A block moves a 128 (k) x 128-byte (n)
// tile: 16-byte loads of 128-byte row pieces into LDS (row pitch 33 words),
// then each thread reads one word from each of 16 k rows (4 / ESZ columns)
// and transposes them in registers into 16 k of each of its columns, stored
// as 16-byte pieces that fill whole 128-byte runs of the output rows.
// colsum (bytes only, optional) accumulates the wrapping int8 column sums.
template <int ESZ>
__global__ __launch_bounds__(256) void pack_bt_wide_kernel(const uint8_t* __restrict__ b,
                                                           int row_bytes, int ldo_bytes,
                                                           uint8_t* __restrict__ out,
                                                           int32_t* colsum) {
  constexpr int P = 33;              // LDS row pitch in words
  constexpr int CPW = 4 / ESZ;       // columns per word
  __shared__ uint32_t tile[128 * P];
  const int t = threadIdx.x;
  const size_t nb0 = (size_t)blockIdx.x * 128;  // byte column of the tile in B
  const int k0 = blockIdx.y * 128;
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const int r = p * 32 + (t >> 3), c = t & 7;
    const u4v v = *reinterpret_cast<const u4v*>(b + (size_t)(k0 + r) * row_bytes + nb0 + 16 * c);
    uint32_t* d = tile + r * P + 4 * c;
    d[0] = v[0];
    d[1] = v[1];
    d[2] = v[2];
    d[3] = v[3];
  }
  __syncthreads();
  const int g = t >> 3, c = t & 7;   // word column g, k rows 16c .. 16c + 15
  uint32_t w[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) w[i] = tile[(16 * c + i) * P + g];
  const size_t col0 = nb0 / ESZ + (size_t)CPW * g;  // first output row (column of B)
  uint8_t* dst = out + col0 * ldo_bytes + (size_t)k0 * ESZ + 16 * ESZ * c;
#pragma unroll
  for (int j = 0; j < CPW; ++j) {
    uint32_t o[4 * ESZ];
    if constexpr (ESZ == 1) {
#pragma unroll
      for (int q = 0; q < 4; ++q)
        o[q] = ((w[4 * q] >> (8 * j)) & 0xffu) | (((w[4 * q + 1] >> (8 * j)) & 0xffu) << 8) |
               (((w[4 * q + 2] >> (8 * j)) & 0xffu) << 16) | ((w[4 * q + 3] >> (8 * j)) << 24);
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q)
        o[q] = ((w[2 * q] >> (16 * j)) & 0xffffu) | ((w[2 * q + 1] >> (16 * j)) << 16);
    }
#pragma unroll
    for (int h = 0; h < ESZ; ++h)
      *reinterpret_cast<u4v*>(dst + (size_t)j * ldo_bytes + 16 * h) =
          u4v{o[4 * h], o[4 * h + 1], o[4 * h + 2], o[4 * h + 3]};
  }
  if constexpr (ESZ == 1) {
    if (colsum) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        int s = 0;
#pragma unroll
        for (int i = 0; i < 16; ++i) s += (int)(int8_t)(w[i] >> (8 * j));
        s += __shfl_xor(s, 1);
        s += __shfl_xor(s, 2);
        s += __shfl_xor(s, 4);
        if (c == 0) atomicAdd(reinterpret_cast<unsigned int*>(colsum + col0 + j), (unsigned)s);
      }
    }
  }
}

// Ap planes (Mp x Kp) from row-major A (M x K): MAP 0 and 2 as above; MAP 3
// flips the sign bit of u8 (padding stays 0).
template <typename E, int MAP>
__global__ void pack_a_kernel(const E* __restrict__ a, int M, int K, int Mp, int Kp,
                              uint8_t* out0, uint8_t* out1) {
  const size_t total = (size_t)Mp * Kp;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const int m = (int)(i / Kp), k = (int)(i % Kp);
    const bool real = m < M && k < K;
    const E v = real ? a[(size_t)m * K + k] : E(0);
    static_assert(MAP == 0 || MAP == 2 || MAP == 3, "A maps");
    if (MAP == 0) {
      out0[i] = (uint8_t)v;
    } else if (MAP == 2) {
      reinterpret_cast<uint16_t*>(out0)[i] = (uint16_t)v;
    } else {
      out0[i] = real ? (uint8_t)((uint8_t)v ^ 0x80u) : 0;
    }
  }
}

// u8 -> u8 - 128 (sign bit flipped) over a tile-aligned A, one block per row
// of k16 16-byte words; with `rowsum` also the wrapping sum of the row's u8
// values (the same sums rowsum_kernel gives).
__global__ void flip_sign_kernel(const uint4* __restrict__ a, uint4* __restrict__ out, int k16,
                                 int32_t* __restrict__ rowsum) {
  __shared__ uint32_t part[256];
  const size_t base = (size_t)blockIdx.x * k16;
  uint32_t s = 0;
  for (int i = threadIdx.x; i < k16; i += blockDim.x) {
    uint4 v = a[base + i];
    s = __builtin_amdgcn_udot4(v.x, 0x01010101u, s, false);
    s = __builtin_amdgcn_udot4(v.y, 0x01010101u, s, false);
    s = __builtin_amdgcn_udot4(v.z, 0x01010101u, s, false);
    s = __builtin_amdgcn_udot4(v.w, 0x01010101u, s, false);
    v.x ^= 0x80808080u;
    v.y ^= 0x80808080u;
    v.z ^= 0x80808080u;
    v.w ^= 0x80808080u;
    out[base + i] = v;
  }
  if (!rowsum) return;
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) rowsum[blockIdx.x] = (int32_t)part[0];
}

// Wrapping row sums of a row-major M x K matrix, one block per row.
template <typename E>
__global__ void rowsum_kernel(const E* __restrict__ a, int K, int32_t* out) {
  __shared__ uint32_t part[256];
  const int row = blockIdx.x;
  uint32_t s = 0;
  for (int k = threadIdx.x; k < K; k += blockDim.x) s += (uint32_t)(int32_t)a[(size_t)row * K + k];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[row] = (int32_t)part[0];
}

__global__ void zero_words_kernel(uint32_t* p, size_t count) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

static unsigned grid_for(size_t count) {
  size_t b = cdiv(count, 256);
  if (b > 65536) b = 65536;
  if (b == 0) b = 1;
  return (unsigned)b;
}

static inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

struct Carver {
  size_t off = 0;
  size_t take(size_t bytes) {
    size_t o = off;
    off = align_up(off + bytes, 256);
    return o;
  }
};

// ---------------------------------------------------------------------
// i16 x i16 -> i32: the four byte-plane products in one block
// ---------------------------------------------------------------------
// A 128 x 128 tile of C per block, four waves of 64 x 64.  A stage holds 128
// bytes of k of four planes: A_hi and A_lo' rows of the tile (the A panel,
// LDS rows 0-127 and 128-255) and B_hi and B_lo' columns (the B panel, rows
// 256-383 and 384-511): the same 64 KiB layout, swizzle and direct-to-LDS
// copies as the 256 x 256 kernel, so the LDS bytes per matrix-core
// instruction are the same too.  Each k-step multiplies every A fragment of
// the wave with every B fragment into three register sets:
//   HH += Ah Bh,   X += Ah Bl' + Al' Bh,   LL += Al' Bl'
// (v_mfma_i32_16x16x64_i8, 4 x 16 tiles per set, 192 accumulator registers),
// and the epilogue forms, wrapping in uint32,
//   C = 2^16 HH + 2^8 X + LL + 128 rowsum(A) + 128 colsum(B) - 16384 k (+ C).
// K is reduced in chunks of 2^15 bytes so that no accumulator (X sums two
// products: |X| <= 2 * 2^15 * 2^14 = 2^30) leaves the i32 range.
constexpr size_t I16_K_CHUNK = (size_t)1 << 15;

struct I16Operands {
  const uint8_t* a_hi;
  const uint8_t* a_lo;
  const uint8_t* b_hi;  // B^T planes, n x Kp
  const uint8_t* b_lo;
  size_t lda, ldb;      // row pitch (Kp) in bytes
  int kbytes;           // bytes of k in this chunk, a multiple of 128
  int tiles_m, tiles_n;
  int M, N;
  int32_t* c;
  const int32_t* rowsum;  // first chunk: wrapping row sums of A, else null
  const int32_t* colsum;  // first chunk: wrapping column sums of B, else null
  uint32_t kconst;        // first chunk: -16384 * k, else 0
  int read_c;             // add onto C
};

template <bool READ_C, bool FIRST, bool CHECK>
__device__ __forceinline__ void i16_store(const I16Operands& p, int row, int col, int hh, int x,
                                          int ll) {
  if (CHECK && (row >= p.M || col >= p.N)) return;
  uint32_t v = ((uint32_t)nt_vgpr(hh) << 16) + ((uint32_t)nt_vgpr(x) << 8) + (uint32_t)nt_vgpr(ll);
  if (FIRST) v += 128u * ((uint32_t)p.rowsum[row] + (uint32_t)p.colsum[col]) + p.kconst;
  int32_t* dst = p.c + (size_t)row * p.N + col;
  if (READ_C) v += (uint32_t)*dst;
  *dst = (int32_t)v;
}

__global__ __launch_bounds__(256) void mfma_i16_dma_kernel(I16Operands p) {
  constexpr int KB = 128, STAGE = 512 * KB, TI = 4, TJ = 4, KSTEP = 64, NKK = KB / KSTEP;
  constexpr int NMF = 4 * TI * TJ, NRD = 2 * (TI + TJ), NLD = 16;
  extern __shared__ __attribute__((aligned(16))) uint8_t nt_smem[];

  int tm, tn;
  nt_tile(p.tiles_m, p.tiles_n, tm, tn);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm0 = (wave >> 1) * 64, wn0 = (wave & 1) * 64;
  const int lr = lane & 15, lk = lane >> 4;

  // Copy instruction q of wave w fills LDS rows 8 (4q + w) .. + 7 of each
  // panel (as in the 256 x 256 kernel); q < 4 takes them from the hi plane,
  // q >= 4 from the lo plane, at tile row 8 (4 (q & 3) + w) + lane / 8.
  const int lrow = lane >> 3, lslot = lane & 7;
  const int gchunk = lslot ^ ((4 * (wave & 1) + (lrow >> 1)) ^ ((lrow & 1) << 2));
  const size_t a_t = (size_t)tm * 128 * p.lda, b_t = (size_t)tn * 128 * p.ldb;
  const uint8_t* ga0 = p.a_hi + a_t + (size_t)(wave * 8 + lrow) * p.lda + gchunk * 16;
  const uint8_t* ga1 = p.a_lo + a_t + (size_t)(wave * 8 + lrow) * p.lda + gchunk * 16;
  const uint8_t* gb0 = p.b_hi + b_t + (size_t)(wave * 8 + lrow) * p.ldb + gchunk * 16;
  const uint8_t* gb1 = p.b_lo + b_t + (size_t)(wave * 8 + lrow) * p.ldb + gchunk * 16;
  const size_t a_q = 32 * p.lda, b_q = 32 * p.ldb;
  auto issue = [&](int buf, int kb) __attribute__((always_inline)) {
    uint8_t* as = nt_smem + buf * STAGE + wave * 1024;
    uint8_t* bs = as + 256 * KB;
#pragma unroll
    for (int q = 0; q < 8; ++q)
      __builtin_amdgcn_global_load_lds((nt_gvoid*)((q < 4 ? ga0 : ga1) + kb + (q & 3) * a_q),
                                       (nt_lvoid*)(as + q * 4096), 16, 0, 0);
#pragma unroll
    for (int q = 0; q < 8; ++q)
      __builtin_amdgcn_global_load_lds((nt_gvoid*)((q < 4 ? gb0 : gb1) + kb + (q & 3) * b_q),
                                       (nt_lvoid*)(bs + q * 4096), 16, 0, 0);
  };

  // Fragment of k-step kk: lane l reads row l % 16 of the tile, chunk
  // kk * 4 + l / 16 at its swizzled slot (every fragment's rows start at a
  // multiple of 16, so the row's swizzle is that of l % 16).
  const int swz = ((lr >> 1) & 7) ^ ((lr & 1) << 2);
  int xo[NKK];
#pragma unroll
  for (int kk = 0; kk < NKK; ++kk) xo[kk] = ((kk * 4 + lk) ^ swz) << 4;
  const int ahi = (wm0 + lr) * KB, alo = (128 + wm0 + lr) * KB;
  const int bhi = (256 + wn0 + lr) * KB, blo = (384 + wn0 + lr) * KB;
  // fragment sets: f[0..3] hi, f[4..7] lo
  auto read_step = [&](const uint8_t* s, int kk, i4* fa, i4* fb) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < TI; ++i) {
      fa[i] = *reinterpret_cast<const i4*>(s + ahi + i * 16 * KB + xo[kk]);
      fa[TI + i] = *reinterpret_cast<const i4*>(s + alo + i * 16 * KB + xo[kk]);
    }
#pragma unroll
    for (int j = 0; j < TJ; ++j) {
      fb[j] = *reinterpret_cast<const i4*>(s + bhi + j * 16 * KB + xo[kk]);
      fb[TJ + j] = *reinterpret_cast<const i4*>(s + blo + j * 16 * KB + xo[kk]);
    }
  };

  i4 hh[TI][TJ], xx[TI][TJ], ll[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) hh[i][j] = xx[i][j] = ll[i][j] = i4{};
  // The two updates of X are 32 instructions apart.
  auto mma = [&](const i4* fa, const i4* fb) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j)
        hh[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[i], fb[j], hh[i][j], 0, 0, 0);
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j)
        xx[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[i], fb[TJ + j], xx[i][j], 0, 0, 0);
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j)
        ll[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[TI + i], fb[TJ + j], ll[i][j], 0, 0, 0);
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j)
        xx[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[TI + i], fb[j], xx[i][j], 0, 0, 0);
  };

  i4 pa[2 * TI], pb[2 * TJ];  // last k-step of the previous stage
  // The same rotated stage body as the 256 x 256 kernel: the previous
  // stage's last k-step is multiplied after the opening barrier, with the
  // next stage's copies between its MFMAs.
  auto body = [&](auto pend_c, auto more_c, int cur, int kb_next) __attribute__((always_inline)) {
    constexpr bool PEND = decltype(pend_c)::value, MORE = decltype(more_c)::value;
    const uint8_t* s = nt_smem + cur * STAGE;
    i4 fa[2 * TI], fb[2 * TJ];
    read_step(s, 0, fa, fb);
    if (MORE) issue(cur ^ 1, kb_next);
    if (PEND) mma(pa, pb);
    read_step(s, 1, pa, pb);
    mma(fa, fb);
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
    for (int q = 0; q < NRD; ++q) {
      __builtin_amdgcn_sched_group_barrier(0x008, NMF / NRD, 0);
      __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
    }
  };
  auto barrier = [&]() __attribute__((always_inline)) {
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
    __syncthreads();
    __builtin_amdgcn_sched_barrier(0);
  };
  using T = std::integral_constant<bool, true>;
  using F = std::integral_constant<bool, false>;
  static_assert(NKK == 2, "two k-steps per stage");

  const int nk = p.kbytes / KB;
  issue(0, 0);
  __syncthreads();
  if (nk > 1)
    body(F{}, T{}, 0, KB);
  else
    body(F{}, F{}, 0, 0);
  int kt = 1;
  for (; kt + 1 < nk; ++kt) {
    barrier();
    body(T{}, T{}, kt & 1, (kt + 1) * KB);
  }
  if (kt < nk) {
    barrier();
    body(T{}, F{}, kt & 1, 0);
  }
  mma(pa, pb);

  // Result element r of lane l of a 16 x 16 tile: row 4 (l / 16) + r,
  // column l % 16.
  const int row0 = tm * 128 + wm0, col0 = tn * 128 + wn0;
  auto store_all = [&](auto rc, auto first, auto check) __attribute__((always_inline)) {
    constexpr bool RC = decltype(rc)::value, FI = decltype(first)::value,
                   CK = decltype(check)::value;
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          i16_store<RC, FI, CK>(p, row0 + i * 16 + 4 * lk + r, col0 + j * 16 + lr, hh[i][j][r],
                                xx[i][j][r], ll[i][j][r]);
  };
  const bool inside = tm * 128 + 128 <= p.M && tn * 128 + 128 <= p.N;
  const bool first = p.rowsum != nullptr;
  if (inside) {
    if (p.read_c) {
      if (first) store_all(T{}, T{}, F{}); else store_all(T{}, F{}, F{});
    } else {
      if (first) store_all(F{}, T{}, F{}); else store_all(F{}, F{}, F{});
    }
  } else {
    if (p.read_c) {
      if (first) store_all(T{}, T{}, T{}); else store_all(T{}, F{}, T{});
    } else {
      if (first) store_all(F{}, T{}, T{}); else store_all(F{}, F{}, T{});
    }
  }
}

// i16 rows -> hi byte plane (a >> 8) and lo' plane ((a & 255) - 128), k
// padded with zeros to Kp; one block per row of the Mp-row planes (rows past
// M are zero).  With `rowsum`, also the wrapping sum of each row's values.
// `zero` holds gridDim.x * zper words, all cleared.
__global__ __launch_bounds__(256) void split_i16_rows_kernel(const int16_t* __restrict__ a, int M,
                                                             int K, int Kp, uint8_t* hi,
                                                             uint8_t* lo, int32_t* rowsum,
                                                             uint32_t* zero, int zper) {
  __shared__ uint32_t part[256];
  const int row = blockIdx.x;
  // Side job: block r clears words r*zper .. +zper-1 of `zero` (the column
  // sums that the B pass accumulates next), which saves a launch.
  for (int t = threadIdx.x; t < zper; t += blockDim.x) zero[(size_t)row * zper + t] = 0;
  const int16_t* src = a + (size_t)row * K;
  uint8_t* dh = hi + (size_t)row * Kp;
  uint8_t* dl = lo + (size_t)row * Kp;
  const bool vec = ((uintptr_t)src % 16) == 0;
  uint32_t s = 0;
  for (int k0 = threadIdx.x * 16; k0 < Kp; k0 += blockDim.x * 16) {
    int16_t v[16];
    if (row < M && vec && k0 + 16 <= K) {
      *reinterpret_cast<u4v*>(&v[0]) = *reinterpret_cast<const u4v*>(src + k0);
      *reinterpret_cast<u4v*>(&v[8]) = *reinterpret_cast<const u4v*>(src + k0 + 8);
    } else {
#pragma unroll
      for (int e = 0; e < 16; ++e) v[e] = (row < M && k0 + e < K) ? src[k0 + e] : int16_t(0);
    }
    uint8_t h[16], l[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const bool real = row < M && k0 + e < K;
      const int x = v[e];
      s += (uint32_t)x;
      h[e] = real ? (uint8_t)(int8_t)(x >> 8) : 0;
      l[e] = real ? (uint8_t)(int8_t)((x & 255) - 128) : 0;
    }
    *reinterpret_cast<u4v*>(dh + k0) = __builtin_bit_cast(u4v, h);
    *reinterpret_cast<u4v*>(dl + k0) = __builtin_bit_cast(u4v, l);
  }
  if (!rowsum || row >= M) return;
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) rowsum[row] = (int32_t)part[0];
}

// ---------------------------------------------------------------------
// Driver
// ---------------------------------------------------------------------

// Caller-provided scratch in place of the per-stream workspace (the
// allocation-free entry point): `bytes` of device memory at `ptr`.
struct ExtScratch {
  uint8_t* ptr;
  size_t bytes;
};

template <class S>
static int run_nt_shape(int kind, size_t m, size_t n, size_t k, const void* a, const void* b,
                        void* c, bool accumulate, uint8_t a_zp, float scale_a, int8_t b_zp,
                        float scale_b, hipStream_t st, const ExtScratch* ext) {
  const bool half = (kind == NT_BF16 || kind == NT_F16);
  const bool fp8 = (kind == NT_E4M3 || kind == NT_E5M2);
  const bool is_u8 = (kind == NT_U8 || kind == NT_QUANT);
  const size_t es = half ? 2 : 1;                       // input element bytes
  const size_t pe = half ? 2 : 1;                       // plane element bytes
  const size_t kel = S::KB / pe;                        // k elements per stage
  const size_t Kp = align_up(k, kel);
  const size_t Mp = align_up(m, S::BM), Np = align_up(n, S::BN);
  const bool need_colsum = (kind == NT_U8 || kind == NT_QUANT);
  const bool need_rowsum = kind == NT_QUANT;
  const bool aligned_a = (m % S::BM == 0) && ((k * es) % S::KB == 0) &&
                         ((uintptr_t)a % 16 == 0);
  // Tile-aligned A is read in place, except u8 A on the 256 x 256 shape:
  // there a flipped copy lets the product run the direct-to-LDS kernel,
  // whose copies move bytes unmodified.
  const bool flip_a = aligned_a && is_u8 && S::BM == 256;
  const bool direct_a = aligned_a && !flip_a;

  Carver cv;
  const size_t o_bt0 = cv.take(Np * Kp * pe);
  const size_t o_cs = cv.take(Np * 4);
  const size_t o_rs = cv.take(Mp * 4);
  const size_t o_a0 = direct_a ? 0 : cv.take(flip_a ? m * k : Mp * Kp * pe);
  void* wsv = nullptr;
  if (ext) {
    // Checked before anything is enqueued: an oversized problem leaves C
    // untouched.
    if (cv.off > ext->bytes) return FERRAY_ERR_SCRATCH;
    wsv = ext->ptr;
  } else {
    int s = workspace(st, cv.off, &wsv);
    if (s) return s;
  }
  uint8_t* ws = reinterpret_cast<uint8_t*>(wsv);
  int32_t* csp = need_colsum ? reinterpret_cast<int32_t*>(ws + o_cs) : nullptr;
  int32_t* rsp = need_rowsum ? reinterpret_cast<int32_t*>(ws + o_rs) : nullptr;

  if (need_colsum) {
    hipLaunchKernelGGL(zero_words_kernel, dim3(grid_for(Np)), dim3(256), 0, st,
                       reinterpret_cast<uint32_t*>(csp), Np);
  }
  if (Kp == k && Np == n && k % 128 == 0 && k / 128 <= 65535 && (n * es) % 128 == 0 &&
      (uintptr_t)b % 16 == 0) {
    const dim3 wgrid((unsigned)(n * es / 128), (unsigned)(k / 128));
    if (half)
      hipLaunchKernelGGL(pack_bt_wide_kernel<2>, wgrid, dim3(256), 0, st,
                         reinterpret_cast<const uint8_t*>(b), (int)(n * es), (int)(k * es),
                         ws + o_bt0, nullptr);
    else
      hipLaunchKernelGGL(pack_bt_wide_kernel<1>, wgrid, dim3(256), 0, st,
                         reinterpret_cast<const uint8_t*>(b), (int)n, (int)k, ws + o_bt0, csp);
  } else {
    for_grid_y(cdiv(Kp, 64), [&](int y0, unsigned ny) {
      const dim3 tgrid((unsigned)(Np / 64), ny);
      if (half)
        hipLaunchKernelGGL((pack_bt_kernel<uint16_t, 2>), tgrid, dim3(256), 0, st,
                           reinterpret_cast<const uint16_t*>(b), (int)k, (int)n, (int)Kp, (int)Np,
                           ws + o_bt0, nullptr, nullptr, y0);
      else
        hipLaunchKernelGGL((pack_bt_kernel<int8_t, 0>), tgrid, dim3(256), 0, st,
                           reinterpret_cast<const int8_t*>(b), (int)k, (int)n, (int)Kp, (int)Np,
                           ws + o_bt0, nullptr, csp, y0);
    });
  }
  FG_CHECK(hipGetLastError());

  if (need_rowsum && !flip_a) {
    hipLaunchKernelGGL(rowsum_kernel<uint8_t>, dim3((unsigned)m), dim3(256), 0, st,
                       reinterpret_cast<const uint8_t*>(a), (int)k, rsp);
    FG_CHECK(hipGetLastError());
  }

  NtOperands p{};
  p.bt = ws + o_bt0;
  p.ldb = Kp * pe;
  p.b_plane = 0;
  p.a_plane = 0;
  p.tiles_m = (int)(Mp / S::BM);
  p.tiles_n = (int)(Np / S::BN);
  bool xor_a = false;
  if (direct_a) {
    p.a = reinterpret_cast<const uint8_t*>(a);
    p.lda = k * es;
    xor_a = is_u8;
  } else if (flip_a) {  // flipped copy (and the row sums) in one pass over A
    hipLaunchKernelGGL(flip_sign_kernel, dim3((unsigned)m), dim3(256), 0, st,
                       reinterpret_cast<const uint4*>(a), reinterpret_cast<uint4*>(ws + o_a0),
                       (int)(k / 16), rsp);
    FG_CHECK(hipGetLastError());
    p.a = ws + o_a0;
    p.lda = k;
  } else {
    const unsigned g = grid_for(Mp * Kp);
    if (half)
      hipLaunchKernelGGL((pack_a_kernel<uint16_t, 2>), dim3(g), dim3(256), 0, st,
                         reinterpret_cast<const uint16_t*>(a), (int)m, (int)k, (int)Mp, (int)Kp,
                         ws + o_a0, nullptr);
    else if (is_u8)
      hipLaunchKernelGGL((pack_a_kernel<uint8_t, 3>), dim3(g), dim3(256), 0, st,
                         reinterpret_cast<const uint8_t*>(a), (int)m, (int)k, (int)Mp, (int)Kp,
                         ws + o_a0, nullptr);
    else
      hipLaunchKernelGGL((pack_a_kernel<int8_t, 0>), dim3(g), dim3(256), 0, st,
                         reinterpret_cast<const int8_t*>(a), (int)m, (int)k, (int)Mp, (int)Kp,
                         ws + o_a0, nullptr);
    FG_CHECK(hipGetLastError());
    p.a = ws + o_a0;
    p.lda = Kp * pe;
  }

  const int M = (int)m, N = (int)n;
  // Integer paths reduce K in chunks so no i32 accumulator can overflow.
  const size_t kb_total = Kp * pe;
  const size_t chunk = (half || fp8) ? kb_total : align_up(NT_K_CHUNK, S::KB);
  for (size_t kb0 = 0; kb0 < kb_total; kb0 += chunk) {
    NtOperands q = p;
    q.a = p.a + kb0;
    q.bt = p.bt + kb0;
    q.kbytes = (int)std::min(chunk, kb_total - kb0);
    const bool first = kb0 == 0;
    const int read_c = (!first || accumulate) ? 1 : 0;
    int rc = FERRAY_OK;
    if (half) {
      EpiF32 e{reinterpret_cast<float*>(c), M, N, read_c};
      rc = kind == NT_BF16 ? launch_mfma_nt<ELEM_BF16, S, false>(q, 1, e, st)
                           : launch_mfma_nt<ELEM_F16, S, false>(q, 1, e, st);
    } else if (fp8) {
      EpiF32 e{reinterpret_cast<float*>(c), M, N, read_c};
      rc = kind == NT_E4M3 ? launch_mfma_nt<ELEM_FP8, S, false>(q, 1, e, st)
                           : launch_mfma_nt<ELEM_BF8, S, false>(q, 1, e, st);
    } else if (kind == NT_QUANT) {
      if (first) {
        EpiQuant e{reinterpret_cast<float*>(c), M, N, accumulate ? 1 : 0, csp, rsp,
                   (uint32_t)(int32_t)a_zp, (uint32_t)(int32_t)b_zp,
                   (uint32_t)k * (uint32_t)(int32_t)a_zp * (uint32_t)(int32_t)b_zp,
                   scale_a * scale_b};
        rc = xor_a ? launch_mfma_nt<ELEM_I8, S, true>(q, 1, e, st)
                   : launch_mfma_nt<ELEM_I8, S, false>(q, 1, e, st);
      } else {
        rc = FERRAY_ERR_INVALID;  // unreachable: run_nt routes k > 2^16 to run_quant_long
      }
    } else {
      EpiI32 e{reinterpret_cast<int32_t*>(c), M, N, read_c,
               (kind == NT_U8 && first) ? csp : nullptr};
      rc = xor_a ? launch_mfma_nt<ELEM_I8, S, true>(q, 1, e, st)
                 : launch_mfma_nt<ELEM_I8, S, false>(q, 1, e, st);
    }
    if (rc) return rc;
  }
  return FERRAY_OK;
}

// i16 x i16 -> i32: split A into its two byte planes (with the row sums in
// the same pass), transpose B into two B^T planes (with the column sums),
// then one launch per K chunk of the four-product kernel.
static int run_i16(size_t m, size_t n, size_t k, const int16_t* a, const int16_t* b, int32_t* c,
                   bool accumulate, hipStream_t st, const ExtScratch* ext) {
  const size_t Mp = align_up(m, 128), Np = align_up(n, 128), Kp = align_up(k, 128);
  Carver cv;
  const size_t o_a0 = cv.take(Mp * Kp), o_a1 = cv.take(Mp * Kp);
  const size_t o_b0 = cv.take(Np * Kp), o_b1 = cv.take(Np * Kp);
  // the column sums get Mp * zper words, cleared by A's split (one block per row)
  const size_t zper = cdiv(Np, Mp);
  const size_t o_cs = cv.take(Mp * zper * 4), o_rs = cv.take(Mp * 4);
  void* wsv = nullptr;
  if (ext) {
    if (cv.off > ext->bytes) return FERRAY_ERR_SCRATCH;
    wsv = ext->ptr;
  } else {
    int s = workspace(st, cv.off, &wsv);
    if (s) return s;
  }
  uint8_t* ws = reinterpret_cast<uint8_t*>(wsv);
  int32_t* csp = reinterpret_cast<int32_t*>(ws + o_cs);
  int32_t* rsp = reinterpret_cast<int32_t*>(ws + o_rs);
  // A's split also clears the column sums the B pass then accumulates
  hipLaunchKernelGGL(split_i16_rows_kernel, dim3((unsigned)Mp), dim3(256), 0, st, a, (int)m,
                     (int)k, (int)Kp, ws + o_a0, ws + o_a1, rsp,
                     reinterpret_cast<uint32_t*>(csp), (int)zper);
  for_grid_y(Kp / 64, [&](int y0, unsigned ny) {
    hipLaunchKernelGGL((pack_bt_kernel<int16_t, 1>), dim3((unsigned)(Np / 64), ny), dim3(256), 0,
                       st, b, (int)k, (int)n, (int)Kp, (int)Np, ws + o_b0, ws + o_b1, csp, y0);
  });
  FG_CHECK(hipGetLastError());

  constexpr int lds = 2 * 512 * 128;
  static bool attr_set = false;  // every caller sets the same value; a race is benign
  if (!attr_set) {
    FG_CHECK(hipFuncSetAttribute((const void*)mfma_i16_dma_kernel,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, lds));
    attr_set = true;
  }
  const size_t tiles = (Mp / 128) * (Np / 128);
  if (tiles > 0x7fffffff) return FERRAY_ERR_INVALID;
  for (size_t kb0 = 0; kb0 < Kp; kb0 += I16_K_CHUNK) {
    const bool first = kb0 == 0;
    I16Operands q;
    q.a_hi = ws + o_a0 + kb0;
    q.a_lo = ws + o_a1 + kb0;
    q.b_hi = ws + o_b0 + kb0;
    q.b_lo = ws + o_b1 + kb0;
    q.lda = Kp;
    q.ldb = Kp;
    q.kbytes = (int)std::min(I16_K_CHUNK, Kp - kb0);
    q.tiles_m = (int)(Mp / 128);
    q.tiles_n = (int)(Np / 128);
    q.M = (int)m;
    q.N = (int)n;
    q.c = c;
    q.rowsum = first ? rsp : nullptr;
    q.colsum = first ? csp : nullptr;
    q.kconst = first ? (uint32_t)(0u - 16384u * (uint32_t)k) : 0u;
    q.read_c = (!first || accumulate) ? 1 : 0;
    hipLaunchKernelGGL(mfma_i16_dma_kernel, dim3((unsigned)tiles), dim3(256), lds, st, q);
    FG_CHECK(hipGetLastError());
  }
  return FERRAY_OK;
}

// Wrapping column sums of a row-major K x N int8 matrix.
__global__ void colsum_i8_kernel(const int8_t* __restrict__ b, int K, int N, int32_t* out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  uint32_t s = 0;
  for (int kk = 0; kk < K; ++kk) s += (uint32_t)(int32_t)b[(size_t)kk * N + j];
  out[j] = (int32_t)s;
}

// quantized_matmul epilogue over an exact raw u8 x i8 product.
__global__ void quant_finish_kernel(const int32_t* __restrict__ raw, float* c, int M, int N,
                                    const int32_t* colsum, const int32_t* rowsum, uint32_t azp,
                                    uint32_t bzp, uint32_t kab, float scale, int accumulate) {
  const size_t total = (size_t)M * N;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const int row = (int)(i / N), col = (int)(i % N);
    const uint32_t v = (uint32_t)raw[i] - azp * (uint32_t)colsum[col] -
                       bzp * (uint32_t)rowsum[row] + kab;
    const float out = scale * (float)(int32_t)v;
    c[i] = accumulate ? c[i] + out : out;
  }
}

static int run_nt(int kind, size_t m, size_t n, size_t k, const void* a, const void* b, void* c,
                  bool accumulate, uint8_t a_zp, float scale_a, int8_t b_zp, float scale_b,
                  hipStream_t st, const ExtScratch* ext = nullptr);

// quantized_matmul with k > 2^16: the exact wrapping raw product is formed in
// a stream-ordered temporary by the chunked u8 path, then finished
// elementwise (mod.rs:507-573 in two steps, as the reference does).
static int run_quant_long(size_t m, size_t n, size_t k, const void* a, const void* b, void* c,
                          bool accumulate, uint8_t a_zp, float scale_a, int8_t b_zp,
                          float scale_b, hipStream_t st) {
  void* tmp = nullptr;
  const size_t bytes = m * n * 4 + (m + n) * 4 + 512;
  FG_CHECK(hipMallocAsync(&tmp, bytes, st));
  int32_t* raw = reinterpret_cast<int32_t*>(tmp);
  int32_t* rs = raw + m * n;
  int32_t* cs = rs + m;
  int rc = run_nt(NT_U8, m, n, k, a, b, raw, false, 0, 1.f, 0, 1.f, st);
  if (rc == FERRAY_OK) {
    hipLaunchKernelGGL(rowsum_kernel<uint8_t>, dim3((unsigned)m), dim3(256), 0, st,
                       reinterpret_cast<const uint8_t*>(a), (int)k, rs);
    hipLaunchKernelGGL(colsum_i8_kernel, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, st,
                       reinterpret_cast<const int8_t*>(b), (int)k, (int)n, cs);
    const uint32_t azp = (uint32_t)(int32_t)a_zp, bzp = (uint32_t)(int32_t)b_zp;
    hipLaunchKernelGGL(quant_finish_kernel, dim3(grid_for(m * n)), dim3(256), 0, st, raw,
                       reinterpret_cast<float*>(c), (int)m, (int)n, cs, rs, azp, bzp,
                       (uint32_t)k * azp * bzp, scale_a * scale_b, accumulate ? 1 : 0);
    if (hipGetLastError() != hipSuccess) rc = FERRAY_ERR_HIP;
  }
  if (hipFreeAsync(tmp, st) != hipSuccess && rc == FERRAY_OK) rc = FERRAY_ERR_HIP;
  return rc;
}

static int run_nt(int kind, size_t m, size_t n, size_t k, const void* a, const void* b, void* c,
                  bool accumulate, uint8_t a_zp, float scale_a, int8_t b_zp, float scale_b,
                  hipStream_t st, const ExtScratch* ext) {
  if (m == 0 || n == 0) return FERRAY_OK;
  if (!c || (k && (!a || !b))) return FERRAY_ERR_INVALID;
  if (m > (1u << 30) || n > (1u << 30) || k > (1u << 30)) return FERRAY_ERR_INVALID;
  if (k == 0) {
    if (!accumulate) {
      hipLaunchKernelGGL(zero_words_kernel, dim3(grid_for(m * n)), dim3(256), 0, st,
                         reinterpret_cast<uint32_t*>(c), m * n);
      FG_CHECK(hipGetLastError());
    }
    return FERRAY_OK;
  }
  if (ext && kind == NT_QUANT) return FERRAY_ERR_INVALID;  // the long form allocates
  if (kind == NT_I16)
    return run_i16(m, n, k, reinterpret_cast<const int16_t*>(a),
                   reinterpret_cast<const int16_t*>(b), reinterpret_cast<int32_t*>(c), accumulate,
                   st, ext);
  if (kind == NT_QUANT && k > NT_K_CHUNK) return run_quant_long(m, n, k, a, b, c, accumulate,
                                                                a_zp, scale_a, b_zp, scale_b, st);
  // 256 x 256 tiles when they give at least 128 blocks; below that the
  // 128 x 128 shape fills more CUs (measured: +28 % at 2048^3 and up to 2x
  // at 1024^3; 7 % slower at 3072^3 with 144 big tiles, equal at 4096^3).
  // Both shapes reduce k in the same order, so results are identical.
  if (m >= 256 && n >= 256 && cdiv(m, (size_t)256) * cdiv(n, (size_t)256) >= 128)
    return run_nt_shape<NtBig>(kind, m, n, k, a, b, c, accumulate, a_zp, scale_a, b_zp, scale_b,
                               st, ext);
  return run_nt_shape<NtSmall>(kind, m, n, k, a, b, c, accumulate, a_zp, scale_a, b_zp, scale_b,
                               st, ext);
}

// ---------------------------------------------------------------------
// Fixed scratch of the allocation-free i8 entry point
// ---------------------------------------------------------------------
// One array per device, allocated with the code object when the library is
// loaded and held until the process ends.  In host code `noalloc_scratch`
// names a host-side stand-in; the device address comes from
// hipGetSymbolAddress, once per device, with that device current.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)128 << 20;
static __device__ __align__(256) uint8_t noalloc_scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_noalloc_mu;
static uint8_t* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

static int noalloc_scratch_of_current_device(ExtScratch* out) {
  int dev = -1;
  FG_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return FERRAY_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch));
    if (e != hipSuccess) return map_hip_error(e);
    if (!p) return FERRAY_ERR_HIP;
    g_noalloc_ptr[dev] = reinterpret_cast<uint8_t*>(p);
  }
  out->ptr = g_noalloc_ptr[dev];
  out->bytes = NOALLOC_SCRATCH_BYTES;
  return FERRAY_OK;
}

}  // namespace fg

using namespace fg;

extern "C" {

int ferray_gemm_i8_dev(size_t m, size_t n, size_t k, const uint8_t* a, const int8_t* b,
                       int32_t* c, bool accumulate, void* stream) {
  int s = ensure_device();
  if (s) return s;
  return run_nt(NT_U8, m, n, k, a, b, c, accumulate, 0, 1.f, 0, 1.f, resolve_stream(stream));
}

int ferray_gemm_i8_signed_dev(size_t m, size_t n, size_t k, const int8_t* a, const int8_t* b,
                              int32_t* c, bool accumulate, void* stream) {
  int s = ensure_device();
  if (s) return s;
  return run_nt(NT_I8S, m, n, k, a, b, c, accumulate, 0, 1.f, 0, 1.f, resolve_stream(stream));
}

int ferray_gemm_i8_signed_dev_noalloc(size_t m, size_t n, size_t k, const int8_t* a,
                                      const int8_t* b, int32_t* c, bool accumulate,
                                      void* stream) {
  int s = ensure_device();
  if (s) return s;
  if (!stream) return FERRAY_ERR_INVALID;
  ExtScratch ext;
  if ((s = noalloc_scratch_of_current_device(&ext))) return s;
  return run_nt(NT_I8S, m, n, k, a, b, c, accumulate, 0, 1.f, 0, 1.f,
                reinterpret_cast<hipStream_t>(stream), &ext);
}

int ferray_gemm_i16_dev(size_t m, size_t n, size_t k, const int16_t* a, const int16_t* b,
                        int32_t* c, bool accumulate, void* stream) {
  int s = ensure_device();
  if (s) return s;
  return run_nt(NT_I16, m, n, k, a, b, c, accumulate, 0, 1.f, 0, 1.f, resolve_stream(stream));
}

int ferray_quantized_matmul_dev(size_t m, size_t n, size_t k, const uint8_t* a, uint8_t a_zp,
                                float scale_a, const int8_t* b, int8_t b_zp, float scale_b,
                                float* c, bool accumulate, void* stream) {
  int s = ensure_device();
  if (s) return s;
  return run_nt(NT_QUANT, m, n, k, a, b, c, accumulate, a_zp, scale_a, b_zp, scale_b,
                resolve_stream(stream));
}

int ferray_gemm_bf16_f32_dev(size_t m, size_t n, size_t k, const uint16_t* a, const uint16_t* b,
                             float* c, bool accumulate, void* stream) {
  int s = ensure_device();
  if (s) return s;
  return run_nt(NT_BF16, m, n, k, a, b, c, accumulate, 0, 1.f, 0, 1.f, resolve_stream(stream));
}

int ferray_gemm_f16_f32_dev(size_t m, size_t n, size_t k, const uint16_t* a, const uint16_t* b,
                            float* c, bool accumulate, void* stream) {
  int s = ensure_device();
  if (s) return s;
  return run_nt(NT_F16, m, n, k, a, b, c, accumulate, 0, 1.f, 0, 1.f, resolve_stream(stream));
}

int ferray_gemm_e4m3_f32_dev(size_t m, size_t n, size_t k, const uint8_t* a, const uint8_t* b,
                             float* c, bool accumulate, void* stream) {
  int s = ensure_device();
  if (s) return s;
  return run_nt(NT_E4M3, m, n, k, a, b, c, accumulate, 0, 1.f, 0, 1.f, resolve_stream(stream));
}

int ferray_gemm_e5m2_f32_dev(size_t m, size_t n, size_t k, const uint8_t* a, const uint8_t* b,
                             float* c, bool accumulate, void* stream) {
  int s = ensure_device();
  if (s) return s;
  return run_nt(NT_E5M2, m, n, k, a, b, c, accumulate, 0, 1.f, 0, 1.f, resolve_stream(stream));
}

}  // extern "C"
