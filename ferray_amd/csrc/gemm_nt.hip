// Integer and half-precision GEMM on CDNA4 matrix cores.
//
//   i8 x i8 -> i32    v_mfma_i32_16x16x64_i8    (gemm_i8_signed)
//   u8 x i8 -> i32    same, A shifted by -128 plus 128*colsum(B)  (gemm_i8)
//   i16 x i16 -> i32  four i8 products of byte planes              (gemm_i16)
//   quantized_matmul  u8 x i8 core with a fused f32 epilogue
//   bf16/f16 -> f32   v_mfma_f32_16x16x32_{bf16,f16}
//
// All of these MFMAs take 16 bytes of K per lane, so one "NT" kernel serves
// them: both operands are k-contiguous (A is m x k row-major, B is used as its
// transpose Bt, n x k).  B is transposed (and split / summed where the type
// needs it) by a prepass into library workspace; A is read in place when its
// shape allows, otherwise packed the same way.  Padding is zero, so the main
// kernel never bounds-checks a load.
//
// Integer arithmetic is exact mod 2^32:
//   u8:  sum a*b = sum (a-128)*b + 128*sum_k b[k][j]
//   i16: a = 256*ah + al' + 128 with ah = a>>8, al' = (a & 255) - 128, so
//        sum a*b = 65536*S(ah,bh) + 256*(S(ah,bl') + S(al',bh)) + S(al',bl')
//                  + 128*rowsum(A) + 128*colsum(B) - 16384*k
// where every S(.,.) is an i8 x i8 MFMA product; all terms wrap in uint32.
#include "common.hpp"

namespace fg {

enum NtKind { NT_I8S = 0, NT_U8 = 1, NT_I16 = 2, NT_BF16 = 3, NT_F16 = 4, NT_QUANT = 5 };

constexpr int NT_THREADS = 256;
constexpr int NT_KB = 64;         // bytes of k per pipeline stage
constexpr int NT_ROW = NT_KB + 16;  // LDS row pitch in bytes (16 B pad)

struct NtArgs {
  const uint8_t* a0;  // plane 0 of A (m rows, lda bytes each)
  const uint8_t* a1;  // plane 1 of A (i16 low bytes) or null
  const uint8_t* b0;  // plane 0 of Bt (n rows, ldb bytes each)
  const uint8_t* b1;
  size_t lda, ldb;    // row pitch in bytes
  int kbytes;         // bytes of k to reduce (multiple of NT_KB, zero padded)
  void* c;            // i32 or f32 output, m x n row-major
  int M, N, K;        // real sizes (K in elements)
  int accumulate;
  int xor_a;          // flip the sign bit of every A byte on load (u8 -> u8-128)
  const int32_t* rowsum;  // i16: rowsum(A); quant: rowsum(A) of u8
  const int32_t* colsum;  // u8/quant/i16: colsum(B)
  int32_t azp, bzp;
  float scale;
};

__device__ __forceinline__ void nt_tile_coords(int tiles_m, int tiles_n, int& tm, int& tn) {
  const int G = 8;
  int nt = tiles_m * tiles_n;
  int bid = blockIdx.x;
  if ((nt & 7) == 0) bid = (bid & 7) * (nt >> 3) + (bid >> 3);
  int per_group = G * tiles_n;
  int group = bid / per_group;
  int first_m = group * G;
  int gsz = min(tiles_m - first_m, G);
  int in_group = bid - group * per_group;
  tm = first_m + in_group % gsz;
  tn = in_group / gsz;
}

template <int KIND>
__device__ __forceinline__ f4 mfma_half(uint4 a, uint4 b, f4 c) {
  if constexpr (KIND == NT_BF16)
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a),
                                                   __builtin_bit_cast(bf8, b), c, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(h8, a),
                                                  __builtin_bit_cast(h8, b), c, 0, 0, 0);
}
__device__ __forceinline__ i4 mfma_i8(uint4 a, uint4 b, i4 c) {
  return __builtin_amdgcn_mfma_i32_16x16x64_i8(__builtin_bit_cast(i4, a),
                                               __builtin_bit_cast(i4, b), c, 0, 0, 0);
}

// BM x BN block tile, 4 waves in 2x2, each wave (BM/2) x (BN/2).
template <int KIND, int BM, int BN>
__global__ __launch_bounds__(NT_THREADS) void gemm_nt_kernel(NtArgs p) {
  constexpr bool HALF = (KIND == NT_BF16 || KIND == NT_F16);
  constexpr int PL = (KIND == NT_I16) ? 2 : 1;
  constexpr int TI = BM / 32, TJ = BN / 32;
  constexpr int A_CH = BM * (NT_KB / 16) / NT_THREADS;  // 16-byte chunks per thread
  constexpr int B_CH = BN * (NT_KB / 16) / NT_THREADS;
  static_assert(A_CH >= 1 && B_CH >= 1, "tile too small for the loader");

  __shared__ __attribute__((aligned(16))) uint8_t As[2][PL][BM * NT_ROW];
  __shared__ __attribute__((aligned(16))) uint8_t Bs[2][PL][BN * NT_ROW];

  const int tiles_m = (p.M + BM - 1) / BM, tiles_n = (p.N + BN - 1) / BN;
  int tm, tn;
  nt_tile_coords(tiles_m, tiles_n, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave >> 1) * (BM / 2), wn0 = (wave & 1) * (BN / 2);
  const int lr = lane & 15, lk = lane >> 4;

  uint4 ra[PL][A_CH], rb[PL][B_CH];
  auto load = [&](int kb) {
#pragma unroll
    for (int q = 0; q < A_CH; ++q) {
      int ch = tid + q * NT_THREADS, row = ch >> 2, c16 = ch & 3;
      size_t off = (size_t)(m0 + row) * p.lda + kb + c16 * 16;
      ra[0][q] = *reinterpret_cast<const uint4*>(p.a0 + off);
      if constexpr (PL == 2) ra[1][q] = *reinterpret_cast<const uint4*>(p.a1 + off);
    }
#pragma unroll
    for (int q = 0; q < B_CH; ++q) {
      int ch = tid + q * NT_THREADS, row = ch >> 2, c16 = ch & 3;
      size_t off = (size_t)(n0 + row) * p.ldb + kb + c16 * 16;
      rb[0][q] = *reinterpret_cast<const uint4*>(p.b0 + off);
      if constexpr (PL == 2) rb[1][q] = *reinterpret_cast<const uint4*>(p.b1 + off);
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int q = 0; q < A_CH; ++q) {
      int ch = tid + q * NT_THREADS, row = ch >> 2, c16 = ch & 3;
#pragma unroll
      for (int pl = 0; pl < PL; ++pl) {
        uint4 v = ra[pl][q];
        if (KIND == NT_U8 || KIND == NT_QUANT) {
          if (p.xor_a) {
            v.x ^= 0x80808080u; v.y ^= 0x80808080u; v.z ^= 0x80808080u; v.w ^= 0x80808080u;
          }
        }
        *reinterpret_cast<uint4*>(&As[buf][pl][row * NT_ROW + c16 * 16]) = v;
      }
    }
#pragma unroll
    for (int q = 0; q < B_CH; ++q) {
      int ch = tid + q * NT_THREADS, row = ch >> 2, c16 = ch & 3;
#pragma unroll
      for (int pl = 0; pl < PL; ++pl)
        *reinterpret_cast<uint4*>(&Bs[buf][pl][row * NT_ROW + c16 * 16]) = rb[pl][q];
    }
  };

  using acc_t = typename std::conditional<HALF, f4, i4>::type;
  acc_t acc0[TI][TJ], acc1[TI][TJ], acc2[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) {
      acc0[i][j] = acc_t{0, 0, 0, 0};
      acc1[i][j] = acc_t{0, 0, 0, 0};
      acc2[i][j] = acc_t{0, 0, 0, 0};
    }

  const int nk = p.kbytes / NT_KB;
  load(0);
  store(0);
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    const bool more = kt + 1 < nk;
    if (more) load((kt + 1) * NT_KB);
    uint4 af[PL][TI], bf[PL][TJ];
#pragma unroll
    for (int pl = 0; pl < PL; ++pl) {
#pragma unroll
      for (int i = 0; i < TI; ++i)
        af[pl][i] = *reinterpret_cast<const uint4*>(&As[cur][pl][(wm0 + i * 16 + lr) * NT_ROW + lk * 16]);
#pragma unroll
      for (int j = 0; j < TJ; ++j)
        bf[pl][j] = *reinterpret_cast<const uint4*>(&Bs[cur][pl][(wn0 + j * 16 + lr) * NT_ROW + lk * 16]);
    }
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j) {
        if constexpr (HALF) {
          acc0[i][j] = mfma_half<KIND>(af[0][i], bf[0][j], acc0[i][j]);
        } else if constexpr (KIND == NT_I16) {
          acc0[i][j] = mfma_i8(af[0][i], bf[0][j], acc0[i][j]);   // hi*hi
          acc1[i][j] = mfma_i8(af[0][i], bf[1][j], acc1[i][j]);   // hi*lo'
          acc1[i][j] = mfma_i8(af[1][i], bf[0][j], acc1[i][j]);   // lo'*hi
          acc2[i][j] = mfma_i8(af[1][i], bf[1][j], acc2[i][j]);   // lo'*lo'
        } else {
          acc0[i][j] = mfma_i8(af[0][i], bf[0][j], acc0[i][j]);
        }
      }
    if (more) store(cur ^ 1);
    __syncthreads();
  }

  // Epilogue: acc[i][j][r] is C[m0+wm0+16i+4*lk+r][n0+wn0+16j+lr].
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) {
      const int col = n0 + wn0 + j * 16 + lr;
      if (col >= p.N) continue;
      uint32_t cs = 0;
      if (KIND == NT_U8 || KIND == NT_QUANT || KIND == NT_I16) cs = (uint32_t)p.colsum[col];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm0 + i * 16 + lk * 4 + r;
        if (row >= p.M) continue;
        const size_t idx = (size_t)row * p.N + col;
        if constexpr (HALF) {
          float v = acc0[i][j][r];
          float* dst = reinterpret_cast<float*>(p.c) + idx;
          if (p.accumulate) v = *dst + v;
          *dst = v;
        } else if constexpr (KIND == NT_QUANT) {
          // raw = sum a*b (u8 a); corrected per mod.rs:548-560
          uint32_t raw = (uint32_t)acc0[i][j][r] + 128u * cs;
          uint32_t azp = (uint32_t)p.azp, bzp = (uint32_t)p.bzp;
          uint32_t v = raw - azp * cs - bzp * (uint32_t)p.rowsum[row] +
                       (uint32_t)p.K * azp * bzp;
          float out = p.scale * (float)(int32_t)v;
          float* dst = reinterpret_cast<float*>(p.c) + idx;
          if (p.accumulate) out = *dst + out;
          *dst = out;
        } else {
          uint32_t v;
          if constexpr (KIND == NT_I16) {
            v = ((uint32_t)acc0[i][j][r] << 16) + ((uint32_t)acc1[i][j][r] << 8) +
                (uint32_t)acc2[i][j][r] + 128u * (uint32_t)p.rowsum[row] + 128u * cs -
                16384u * (uint32_t)p.K;
          } else if constexpr (KIND == NT_U8) {
            v = (uint32_t)acc0[i][j][r] + 128u * cs;
          } else {
            v = (uint32_t)acc0[i][j][r];
          }
          int32_t* dst = reinterpret_cast<int32_t*>(p.c) + idx;
          if (p.accumulate) v += (uint32_t)*dst;
          *dst = (int32_t)v;
        }
      }
    }
}

// ---------------------------------------------------------------------
// Prepasses
// ---------------------------------------------------------------------

// Bt[n][k] (n < Np, k < Kp) from row-major B (K x N), zero padded, with the
// per-kind element map; optional wrapping column sums of the original B.
//   kind 0: int8 copy     kind 1: i16 -> hi byte plane + (lo-128) plane
//   kind 2: 16-bit copy
template <typename E, int MAP>
__global__ void pack_bt_kernel(const E* __restrict__ b, int K, int N, int Kp, int Np,
                               uint8_t* out0, uint8_t* out1, int32_t* colsum) {
  __shared__ E tile[32][33];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int n0 = blockIdx.x * 32, k0 = blockIdx.y * 32;
  for (int r = ty; r < 32; r += 8) {
    int k = k0 + r, n = n0 + tx;
    tile[r][tx] = (k < K && n < N) ? b[(size_t)k * N + n] : E(0);
  }
  __syncthreads();
  if (colsum && ty == 0 && n0 + tx < N) {
    uint32_t s = 0;
    for (int r = 0; r < 32; ++r) s += (uint32_t)(int32_t)tile[r][tx];
    atomicAdd(reinterpret_cast<unsigned int*>(colsum + n0 + tx), s);
  }
  for (int r = ty; r < 32; r += 8) {
    int n = n0 + r, k = k0 + tx;
    if (n >= Np || k >= Kp) continue;
    E v = tile[tx][r];
    bool real = (k < K && n < N);
    if (MAP == 0) {
      out0[(size_t)n * Kp + k] = (uint8_t)v;
    } else if (MAP == 1) {
      int x = (int)v;
      out0[(size_t)n * Kp + k] = real ? (uint8_t)(int8_t)(x >> 8) : 0;
      out1[(size_t)n * Kp + k] = real ? (uint8_t)(int8_t)((x & 255) - 128) : 0;
    } else {
      reinterpret_cast<uint16_t*>(out0)[(size_t)n * Kp + k] = (uint16_t)v;
    }
  }
}

// Ap[m][k] (m < Mp, k < Kp) from row-major A (M x K), same maps as above;
// MAP 3: u8 -> u8 ^ 0x80 (padding stays 0).
template <typename E, int MAP>
__global__ void pack_a_kernel(const E* __restrict__ a, int M, int K, int Mp, int Kp,
                              uint8_t* out0, uint8_t* out1) {
  size_t total = (size_t)Mp * Kp;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    int m = (int)(i / Kp), k = (int)(i % Kp);
    bool real = m < M && k < K;
    E v = real ? a[(size_t)m * K + k] : E(0);
    if (MAP == 0) {
      out0[i] = (uint8_t)v;
    } else if (MAP == 1) {
      int x = (int)v;
      out0[i] = real ? (uint8_t)(int8_t)(x >> 8) : 0;
      out1[i] = real ? (uint8_t)(int8_t)((x & 255) - 128) : 0;
    } else if (MAP == 2) {
      reinterpret_cast<uint16_t*>(out0)[i] = (uint16_t)v;
    } else {
      out0[i] = real ? (uint8_t)((uint8_t)v ^ 0x80u) : 0;
    }
  }
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

// Carves aligned sub-buffers out of one workspace allocation.
struct Carver {
  size_t off = 0;
  size_t take(size_t bytes) {
    size_t o = off;
    off = align_up(off + bytes, 256);
    return o;
  }
};

template <int KIND, int BM, int BN>
static void launch_nt(const NtArgs& p, hipStream_t st) {
  size_t tiles = cdiv(p.M, BM) * cdiv(p.N, BN);
  hipLaunchKernelGGL((gemm_nt_kernel<KIND, BM, BN>), dim3((unsigned)tiles), dim3(NT_THREADS), 0,
                     st, p);
}

// One call of the NT family.  `a`, `b` are the caller's device operands.
static int run_nt(int kind, size_t m, size_t n, size_t k, const void* a, const void* b, void* c,
                  bool accumulate, uint8_t a_zp, float scale_a, int8_t b_zp, float scale_b,
                  hipStream_t st) {
  if (m == 0 || n == 0) return FERRAY_OK;
  if (!c || (k && (!a || !b))) return FERRAY_ERR_INVALID;
  if (m > (1u << 30) || n > (1u << 30) || k > (1u << 30)) return FERRAY_ERR_INVALID;
  const bool half = (kind == NT_BF16 || kind == NT_F16);
  if (k == 0) {
    if (!accumulate) {
      hipLaunchKernelGGL(zero_words_kernel, dim3(grid_for(m * n)), dim3(256), 0, st,
                         reinterpret_cast<uint32_t*>(c), m * n);
      FG_CHECK(hipGetLastError());
    }
    return FERRAY_OK;
  }
  const size_t es = half ? 2 : (kind == NT_I16 ? 2 : 1);
  const int BM = (kind == NT_I16) ? 64 : 128, BN = BM;
  const size_t kelems_stage = NT_KB / (half ? 2 : 1);  // elements of k per stage (per plane)
  const size_t Kp = align_up(k, kelems_stage);
  const size_t Mp = align_up(m, BM), Np = align_up(n, BN);
  const size_t pe = half ? 2 : 1;  // plane element bytes
  const bool need_colsum = (kind == NT_U8 || kind == NT_QUANT || kind == NT_I16);
  const bool need_rowsum = (kind == NT_QUANT || kind == NT_I16);
  const bool direct_a = (kind != NT_I16) && (m % BM == 0) && ((k * es) % NT_KB == 0) &&
                        ((uintptr_t)a % 16 == 0);
  const int planes = (kind == NT_I16) ? 2 : 1;

  Carver cv;
  size_t o_bt0 = cv.take(Np * Kp * pe);
  size_t o_bt1 = planes == 2 ? cv.take(Np * Kp) : 0;
  size_t o_cs = cv.take(Np * 4);
  size_t o_rs = cv.take(Mp * 4);
  size_t o_a0 = direct_a ? 0 : cv.take(Mp * Kp * pe);
  size_t o_a1 = (!direct_a && planes == 2) ? cv.take(Mp * Kp) : 0;
  void* wsv = nullptr;
  int s = workspace(st, cv.off, &wsv);
  if (s) return s;
  uint8_t* ws = reinterpret_cast<uint8_t*>(wsv);

  NtArgs p{};
  p.b0 = ws + o_bt0;
  p.b1 = planes == 2 ? ws + o_bt1 : nullptr;
  p.ldb = Kp * pe;
  p.kbytes = (int)(Kp * pe);
  p.c = c;
  p.M = (int)m;
  p.N = (int)n;
  p.K = (int)k;
  p.accumulate = accumulate ? 1 : 0;
  p.colsum = need_colsum ? reinterpret_cast<int32_t*>(ws + o_cs) : nullptr;
  p.rowsum = need_rowsum ? reinterpret_cast<int32_t*>(ws + o_rs) : nullptr;
  p.azp = (int32_t)a_zp;
  p.bzp = (int32_t)b_zp;
  p.scale = scale_a * scale_b;

  if (need_colsum) {
    hipLaunchKernelGGL(zero_words_kernel, dim3(grid_for(Np)), dim3(256), 0, st,
                       reinterpret_cast<uint32_t*>(ws + o_cs), Np);
  }
  dim3 tgrid((unsigned)(Np / 32), (unsigned)(Kp / 32)), tblk(32, 8);
  int32_t* csp = need_colsum ? reinterpret_cast<int32_t*>(ws + o_cs) : nullptr;
  if (kind == NT_I16) {
    hipLaunchKernelGGL((pack_bt_kernel<int16_t, 1>), tgrid, tblk, 0, st,
                       reinterpret_cast<const int16_t*>(b), (int)k, (int)n, (int)Kp, (int)Np,
                       ws + o_bt0, ws + o_bt1, csp);
  } else if (half) {
    hipLaunchKernelGGL((pack_bt_kernel<uint16_t, 2>), tgrid, tblk, 0, st,
                       reinterpret_cast<const uint16_t*>(b), (int)k, (int)n, (int)Kp, (int)Np,
                       ws + o_bt0, nullptr, nullptr);
  } else {
    hipLaunchKernelGGL((pack_bt_kernel<int8_t, 0>), tgrid, tblk, 0, st,
                       reinterpret_cast<const int8_t*>(b), (int)k, (int)n, (int)Kp, (int)Np,
                       ws + o_bt0, nullptr, csp);
  }
  FG_CHECK(hipGetLastError());

  if (need_rowsum) {
    if (kind == NT_I16)
      hipLaunchKernelGGL(rowsum_kernel<int16_t>, dim3((unsigned)m), dim3(256), 0, st,
                         reinterpret_cast<const int16_t*>(a), (int)k,
                         reinterpret_cast<int32_t*>(ws + o_rs));
    else
      hipLaunchKernelGGL(rowsum_kernel<uint8_t>, dim3((unsigned)m), dim3(256), 0, st,
                         reinterpret_cast<const uint8_t*>(a), (int)k,
                         reinterpret_cast<int32_t*>(ws + o_rs));
    FG_CHECK(hipGetLastError());
  }

  const bool is_u8 = (kind == NT_U8 || kind == NT_QUANT);
  if (direct_a) {
    p.a0 = reinterpret_cast<const uint8_t*>(a);
    p.a1 = nullptr;
    p.lda = k * es;
    p.xor_a = is_u8 ? 1 : 0;
  } else {
    unsigned g = grid_for(Mp * Kp);
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
    p.a0 = ws + o_a0;
    p.a1 = planes == 2 ? ws + o_a1 : nullptr;
    p.lda = Kp * pe;
    p.xor_a = 0;
  }

  switch (kind) {
    case NT_I8S: launch_nt<NT_I8S, 128, 128>(p, st); break;
    case NT_U8: launch_nt<NT_U8, 128, 128>(p, st); break;
    case NT_QUANT: launch_nt<NT_QUANT, 128, 128>(p, st); break;
    case NT_I16: launch_nt<NT_I16, 64, 64>(p, st); break;
    case NT_BF16: launch_nt<NT_BF16, 128, 128>(p, st); break;
    case NT_F16: launch_nt<NT_F16, 128, 128>(p, st); break;
    default: return FERRAY_ERR_INVALID;
  }
  FG_CHECK(hipGetLastError());
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

}  // extern "C"
