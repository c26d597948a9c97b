// SGEMM on BF16 matrix cores by operand splitting.
//
// Every f32 operand x is split into bf16 pieces p0 = bf16(x),
// p1 = bf16(x - p0), p2 = bf16(x - p0 - p1); both differences are exact
// (Sterbenz), and x = p0 + p1 + p2 exactly for finite x away from the
// subnormal range.  The product is a sum of piece products, each computed by
// v_mfma_f32_32x32x16_bf16 with f32 accumulation:
//
//   terms = 3 ("bf16x3"): a0 b0 + a0 b1 + a1 b0       (dropped: a1 b1 and the
//        p2 pieces, ~2^-16 |a||b|: about half the bits of f32)
//   terms = 6 ("bf16x6"): a0 b0 + a0 b1 + a1 b0 + a0 b2 + a1 b1 + a2 b0
//        (dropped terms ~2^-24 |a||b|: f32-level accuracy)
//
// The pieces are laid out along k: A' = [A_i0 | A_i1 | ...] (M x T*Kp bf16)
// and B'^T likewise (N x T*Kp, B transposed while split), so the whole sum is
// ONE GEMM of depth T*Kp on the NT matrix-core main loop, with the large
// terms accumulated first.  The epilogue applies alpha and beta as SGEMM
// does (C = alpha*acc + beta*C, beta == 0 writes without reading C).
#include <algorithm>

#include "mfma_nt.hpp"

namespace fg {
namespace {

__device__ __forceinline__ uint16_t bf16_rn(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u)  // Inf stays Inf; NaN stays a (quiet) NaN
    return (uint16_t)((u >> 16) | ((u & 0x7fffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest, ties to even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_value(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ void split3(float x, uint16_t p[3]) {
  p[0] = bf16_rn(x);
  if (!isfinite(x)) {
    p[1] = p[2] = 0;
    return;
  }
  const float r1 = x - bf16_value(p[0]);
  p[1] = bf16_rn(r1);
  p[2] = bf16_rn(r1 - bf16_value(p[1]));
}

// Piece of A and of B in term t.
template <int T>
struct Terms;
template <>
struct Terms<3> {
  static __device__ __forceinline__ int a(int t) { return t == 2 ? 1 : 0; }
  static __device__ __forceinline__ int b(int t) { return t == 1 ? 1 : 0; }
};
template <>
struct Terms<6> {  // (0,0) (0,1) (1,0) (0,2) (1,1) (2,0)
  static __device__ __forceinline__ int a(int t) {
    return t == 2 ? 1 : t == 4 ? 1 : t == 5 ? 2 : 0;
  }
  static __device__ __forceinline__ int b(int t) {
    return t == 1 ? 1 : t == 3 ? 2 : t == 4 ? 1 : 0;
  }
};

// A (M x K f32, row-major) -> A' (Mp x T*Kp bf16), zero outside M x K.
template <int T>
__global__ __launch_bounds__(256) void split_a_kernel(const float* __restrict__ a, int M, int K,
                                                      int Mp, int Kp, uint16_t* __restrict__ out) {
  const size_t total = (size_t)Mp * Kp;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const int m = (int)(i / Kp), k = (int)(i % Kp);
    const float x = (m < M && k < K) ? a[(size_t)m * K + k] : 0.0f;
    uint16_t p[3];
    split3(x, p);
    uint16_t* row = out + (size_t)m * T * Kp;
#pragma unroll
    for (int t = 0; t < T; ++t) row[(size_t)t * Kp + k] = p[Terms<T>::a(t)];
  }
}

// B (K x N f32, row-major) -> B'^T (Np x T*Kp bf16) through a 64 x 64 LDS
// tile; grid (Np / 64, Kp / 64), launched in y slices from y0.
template <int T>
__global__ __launch_bounds__(256) void split_bt_kernel(const float* __restrict__ b, int K, int N,
                                                       int Kp, uint16_t* __restrict__ out,
                                                       int y0) {
  __shared__ float tile[64][65];
  const int tid = threadIdx.x, c = tid & 63;
  const int n0 = blockIdx.x * 64, k0 = (y0 + (int)blockIdx.y) * 64;
  for (int r = tid >> 6; r < 64; r += 4) {
    const int k = k0 + r, n = n0 + c;
    tile[r][c] = (k < K && n < N) ? b[(size_t)k * N + n] : 0.0f;
  }
  __syncthreads();
  for (int r = tid >> 6; r < 64; r += 4) {
    const int n = n0 + r, k = k0 + c;
    uint16_t p[3];
    split3(tile[c][r], p);
    uint16_t* row = out + (size_t)n * T * Kp;
#pragma unroll
    for (int t = 0; t < T; ++t) row[(size_t)t * Kp + k] = p[Terms<T>::b(t)];
  }
}

struct EpiSgemmFast {  // EpiSgemm inside C: no bounds checks
  float* c;
  int N;
  float alpha, beta;
  int alpha_one, beta_zero;
  __device__ __forceinline__ void operator()(int row, int col, int, float acc) const {
    float* dst = c + (size_t)row * N + col;
    float v = alpha_one ? acc : alpha * acc;
    if (!beta_zero) v = fmaf(beta, *dst, v);
    *dst = v;
  }
};

struct EpiSgemm {
  float* c;
  int M, N;
  float alpha, beta;
  int alpha_one, beta_zero;
  __device__ __forceinline__ void operator()(int row, int col, int, float acc) const {
    if (row >= M || col >= N) return;
    float* dst = c + (size_t)row * N + col;
    float v = alpha_one ? acc : alpha * acc;
    if (!beta_zero) v = fmaf(beta, *dst, v);
    *dst = v;
  }
  __device__ __forceinline__ bool interior(int r0, int c0, int rows, int cols) const {
    return r0 + rows <= M && c0 + cols <= N;
  }
  // One unchecked form (flags at run time): four fixed-flag copies of the
  // epilogue made the kernel spill.
  template <class F>
  __device__ __forceinline__ void with_fast(F&& f) const {
    f(EpiSgemmFast{c, N, alpha, beta, alpha_one, beta_zero});
  }
};

inline size_t up(size_t x, size_t a) { return (x + a - 1) / a * a; }
inline unsigned blocks_for(size_t count) {
  size_t b = cdiv(count, 256);
  return (unsigned)std::min<size_t>(std::max<size_t>(b, 1), 65536);
}

template <class S, int T>
int run_split(size_t m, size_t n, size_t k, float alpha, const float* a, const float* b,
              float beta, float* c, hipStream_t st) {
  const size_t Kp = up(k, 64), Mp = up(m, S::BM), Np = up(n, S::BN);
  if ((size_t)T * Kp * 2 > 0x7fffffff || Mp / S::BM * (Np / S::BN) > 0x7fffffff)
    return FERRAY_ERR_INVALID;
  const size_t bytes_a = up(Mp * T * Kp * 2, 256), bytes_b = Np * T * Kp * 2;
  void* wsv = nullptr;
  int s = workspace(st, bytes_a + bytes_b, &wsv);
  if (s) return s;
  uint16_t* pa = reinterpret_cast<uint16_t*>(wsv);
  uint16_t* pb = reinterpret_cast<uint16_t*>(reinterpret_cast<uint8_t*>(wsv) + bytes_a);
  hipLaunchKernelGGL((split_a_kernel<T>), dim3(blocks_for(Mp * Kp)), dim3(256), 0, st, a, (int)m,
                     (int)k, (int)Mp, (int)Kp, pa);
  for_grid_y(Kp / 64, [&](int y0, unsigned ny) {
    hipLaunchKernelGGL((split_bt_kernel<T>), dim3((unsigned)(Np / 64), ny), dim3(256), 0, st, b,
                       (int)k, (int)n, (int)Kp, pb, y0);
  });
  FG_CHECK(hipGetLastError());
  NtOperands p{};
  p.a = reinterpret_cast<const uint8_t*>(pa);
  p.bt = reinterpret_cast<const uint8_t*>(pb);
  p.lda = p.ldb = (size_t)T * Kp * 2;
  p.kbytes = (int)(T * Kp * 2);
  p.tiles_m = (int)(Mp / S::BM);
  p.tiles_n = (int)(Np / S::BN);
  EpiSgemm e{c, (int)m, (int)n, alpha, beta, alpha == 1.0f, beta == 0.0f};
  return launch_mfma_nt<ELEM_BF16, S, false>(p, 1, e, st);
}

}  // namespace
}  // namespace fg

using namespace fg;

extern "C" int ferray_sgemm_bf16x3_dev(size_t m, size_t n, size_t k, float alpha, const float* a,
                                       const float* b, float beta, float* c, int terms,
                                       void* stream) {
  int s = ensure_device();
  if (s) return s;
  if (terms == 0) terms = 3;
  if (terms != 3 && terms != 6) return FERRAY_ERR_INVALID;
  if (m == 0 || n == 0) return FERRAY_OK;
  if (!c || (k && (!a || !b))) return FERRAY_ERR_INVALID;
  if (m > (1u << 30) || n > (1u << 30) || k > (1u << 28)) return FERRAY_ERR_INVALID;
  hipStream_t st = resolve_stream(stream);
  if (k == 0) return ferray_gemm_f32_dev(m, n, 0, alpha, a, b, beta, c, st);
  const bool big = m >= 256 && n >= 256;
  if (terms == 3)
    return big ? run_split<NtBig, 3>(m, n, k, alpha, a, b, beta, c, st)
               : run_split<NtSmall, 3>(m, n, k, alpha, a, b, beta, c, st);
  return big ? run_split<NtBig, 6>(m, n, k, alpha, a, b, beta, c, st)
             : run_split<NtSmall, 6>(m, n, k, alpha, a, b, beta, c, st);
}
