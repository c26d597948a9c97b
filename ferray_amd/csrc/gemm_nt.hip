// Integer and half-precision GEMM entry points on the NT matrix-core main
// loop (mfma_nt.hpp).
//
//   i8 x i8 -> i32    one i8 pass                              (gemm_i8_signed)
//   u8 x i8 -> i32    A's sign bit flipped on load (a - 128), plus 128*colsum(B)
//                                                                      (gemm_i8)
//   i16 x i16 -> i32  four i8 byte-plane products in one launch       (gemm_i16)
//   quantized_matmul  the u8 x i8 pass with a fused zero-point / scale epilogue
//   bf16/f16 -> f32   one half-precision pass                 (gemm_{bf16,f16}_f32)
//   fp8 -> f32        one pass on the 64-deep f8f6f4 instruction, OCP E4M3 or
//                     E5M2 bytes (no reference entry point; the fp8 member of
//                     the same family)
//
// B is transposed into k-contiguous planes by a prepass (which also forms the
// i16 byte planes and the wrapping column sums); A is read in place when its
// shape allows, otherwise packed the same way.  Padding is zero.
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

// The four i16 byte-plane products of one K chunk in one launch (grid row z
// = A plane z / 2 x B plane z % 2, shifted left by 16, 8, 8, 0): each adds
// its shifted value into C with an atomic add.  Integer addition mod 2^32 is
// order-independent, so the result is exact whatever order the adds land
// in.  The first chunk's row z = 0 also adds the sum corrections.
struct EpiI16Atomic {
  int32_t* c;
  int M, N;
  const int32_t* colsum;   // 128*colsum(B) term, or null
  const int32_t* rowsum;   // 128*rowsum(A) term, or null
  uint32_t kconst;         // -16384*k, first chunk only
  __device__ __forceinline__ void operator()(int row, int col, int z, int acc) const {
    if (row >= M || col >= N) return;
    const int shift = z == 0 ? 16 : (z == 3 ? 0 : 8);
    uint32_t v = (uint32_t)acc << shift;
    if (z == 0) {
      if (colsum) v += 128u * (uint32_t)colsum[col];
      if (rowsum) v += 128u * (uint32_t)rowsum[row];
      v += kconst;
    }
    atomicAdd(reinterpret_cast<unsigned int*>(c + (size_t)row * N + col), v);
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
// zero padded to Np x Kp:
//   MAP 0: byte copy      MAP 1: i16 -> hi byte plane + (lo - 128) plane
//   MAP 2: 16-bit copy
// colsum (optional) accumulates the wrapping column sums of B.
template <typename E, int MAP>
__global__ __launch_bounds__(256) void pack_bt_kernel(const E* __restrict__ b, int K, int N,
                                                      int Kp, int Np, uint8_t* out0,
                                                      uint8_t* out1, int32_t* colsum) {
  __shared__ E tile[64][64 + 4];
  const int tid = threadIdx.x;
  const int n0 = blockIdx.x * 64, k0 = blockIdx.y * 64;
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

// Ap planes (Mp x Kp) from row-major A (M x K), same maps; MAP 3 flips the
// sign bit of u8 (padding stays 0).
template <typename E, int MAP>
__global__ void pack_a_kernel(const E* __restrict__ a, int M, int K, int Mp, int Kp,
                              uint8_t* out0, uint8_t* out1) {
  const size_t total = (size_t)Mp * Kp;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const int m = (int)(i / Kp), k = (int)(i % Kp);
    const bool real = m < M && k < K;
    const E v = real ? a[(size_t)m * K + k] : E(0);
    if (MAP == 0) {
      out0[i] = (uint8_t)v;
    } else if (MAP == 1) {
      const int x = (int)v;
      out0[i] = real ? (uint8_t)(int8_t)(x >> 8) : 0;
      out1[i] = real ? (uint8_t)(int8_t)((x & 255) - 128) : 0;
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
  const size_t es = (half || kind == NT_I16) ? 2 : 1;  // input element bytes
  const size_t pe = half ? 2 : 1;                       // plane element bytes
  const int planes = (kind == NT_I16) ? 2 : 1;
  const size_t kel = S::KB / pe;                        // k elements per stage
  const size_t Kp = align_up(k, kel);
  const size_t Mp = align_up(m, S::BM), Np = align_up(n, S::BN);
  const bool need_colsum = (kind == NT_U8 || kind == NT_QUANT || kind == NT_I16);
  const bool need_rowsum = (kind == NT_QUANT || kind == NT_I16);
  const bool aligned_a = (kind != NT_I16) && (m % S::BM == 0) && ((k * es) % S::KB == 0) &&
                         ((uintptr_t)a % 16 == 0);
  // Tile-aligned A is read in place, except u8 A on the 256 x 256 shape:
  // there a flipped copy lets the product run the direct-to-LDS kernel,
  // whose copies move bytes unmodified.
  const bool flip_a = aligned_a && is_u8 && S::BM == 256;
  const bool direct_a = aligned_a && !flip_a;

  Carver cv;
  const size_t o_bt0 = cv.take(Np * Kp * pe);
  const size_t o_bt1 = planes == 2 ? cv.take(Np * Kp) : 0;
  const size_t o_cs = cv.take(Np * 4);
  const size_t o_rs = cv.take(Mp * 4);
  const size_t o_a0 = direct_a ? 0 : cv.take(flip_a ? m * k : Mp * Kp * pe);
  const size_t o_a1 = (!direct_a && planes == 2) ? cv.take(Mp * Kp) : 0;
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
  const dim3 tgrid((unsigned)(Np / 64), (unsigned)cdiv(Kp, 64));
  if (kind == NT_I16)
    hipLaunchKernelGGL((pack_bt_kernel<int16_t, 1>), tgrid, dim3(256), 0, st,
                       reinterpret_cast<const int16_t*>(b), (int)k, (int)n, (int)Kp, (int)Np,
                       ws + o_bt0, ws + o_bt1, csp);
  else if (half)
    hipLaunchKernelGGL((pack_bt_kernel<uint16_t, 2>), tgrid, dim3(256), 0, st,
                       reinterpret_cast<const uint16_t*>(b), (int)k, (int)n, (int)Kp, (int)Np,
                       ws + o_bt0, nullptr, nullptr);
  else
    hipLaunchKernelGGL((pack_bt_kernel<int8_t, 0>), tgrid, dim3(256), 0, st,
                       reinterpret_cast<const int8_t*>(b), (int)k, (int)n, (int)Kp, (int)Np,
                       ws + o_bt0, nullptr, csp);
  FG_CHECK(hipGetLastError());

  if (need_rowsum && !flip_a) {
    if (kind == NT_I16)
      hipLaunchKernelGGL(rowsum_kernel<int16_t>, dim3((unsigned)m), dim3(256), 0, st,
                         reinterpret_cast<const int16_t*>(a), (int)k, rsp);
    else
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
    if (kind == NT_I16)
      hipLaunchKernelGGL((pack_a_kernel<int16_t, 1>), dim3(g), dim3(256), 0, st,
                         reinterpret_cast<const int16_t*>(a), (int)m, (int)k, (int)Mp, (int)Kp,
                         ws + o_a0, ws + o_a1);
    else if (half)
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
    } else if (kind == NT_I16) {
      NtOperands r = q;
      r.a_plane = (size_t)((ws + o_a1) - (ws + o_a0));
      r.b_plane = (size_t)((ws + o_bt1) - (ws + o_bt0));
      const int32_t* cs = first ? csp : nullptr;
      const int32_t* rs = first ? rsp : nullptr;
      const uint32_t kc = first ? (uint32_t)(0u - 16384u * (uint32_t)k) : 0u;
      // All four byte-plane products in one launch (hi*hi<<16, hi*lo'<<8,
      // lo'*hi<<8, lo'*lo'), adding into C; C starts at zero unless the
      // call accumulates.
      if (first && !accumulate) {
        hipLaunchKernelGGL(zero_words_kernel, dim3(grid_for(m * n)), dim3(256), 0, st,
                           reinterpret_cast<uint32_t*>(c), m * n);
        FG_CHECK(hipGetLastError());
      }
      r.cross = 1;
      EpiI16Atomic e{reinterpret_cast<int32_t*>(c), M, N, cs, rs, kc};
      rc = launch_mfma_nt<ELEM_I8, S, false>(r, 4, e, st);
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
