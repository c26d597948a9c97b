// Grouped, persistent GEMM: one launch computes many independent problems of
// uneven shape and mixed dtype (f64, f32, c64, c32, i8 x i8 -> i32).
//
// The host orders the problems by decreasing work (so the largest tiles
// start first), tiles each with the shape a tunable table picks for its dtype
// and size (64x64 or 128x128) and gives each problem the index of its first
// tile in one global numbering.  A persistent grid of up to two blocks per CU
// walks the global tile numbers with a fixed stride and finds a tile's
// problem by binary search over those first-tile indices.  Every tile runs
// the same MFMA k-sequence as the per-dtype entry points (no split of k), so
// a problem's result is bit-identical alone, in any mix, and through its
// single-problem entry point.
//
// Plans own their pinned host and device descriptor buffers, so several
// batches can be in flight on different streams.  A plan's launch is a
// hipGraph built once (descriptor upload -> kernel); per batch only the
// upload size and the grid are updated before the graph is replayed.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "real_tile.hpp"

namespace fg {

typedef unsigned u4g __attribute__((ext_vector_type(4)));
typedef int i16g __attribute__((ext_vector_type(16)));

struct GProblem {
  int dtype, cfg, tiles_n, accumulate;
  int m, n, k, tile_start;  // tile_start: global number of the first tile
  const void* a;
  const void* b;
  void* c;
  double ar, ai, br, bi;
};

constexpr int G_LDS = 40960;
static_assert(RealTileShape<double, 128, 128, 8>::lds_bytes <= G_LDS, "f64 tile LDS");
static_assert(RealTileShape<float, 128, 128, 16>::lds_bytes <= G_LDS, "f32 tile LDS");
static_assert(2 * (128 + 128) * 80 <= G_LDS, "i8 tile LDS");

// ---------------------------------------------------------------------
// i8 x i8 -> i32 tile with row-major B transposed while it is staged:
// 4 k-rows x 4 columns are loaded as dwords and byte-transposed into 4
// k-contiguous dwords.  v_mfma_i32_32x32x32_i8, 4 waves (2x2), 64 bytes of k
// per stage, LDS rows padded to 80 bytes.
// ---------------------------------------------------------------------
template <int BM, int BN, bool EDGE>
__device__ __forceinline__ void i8_tile(const GProblem& P, int m0, int n0, uint8_t* smem) {
  constexpr int KB = 64, PITCH = 80;
  constexpr int TI = BM / 64, TJ = BN / 64;
  constexpr int AU = BM * 4, BU = 16 * (BN / 4);  // A 16-byte chunks, B 4x4 units per stage
  constexpr int AQ = (AU + 255) / 256, BQ = (BU + 255) / 256;
  uint8_t* As = smem;                    // [2][BM][PITCH]
  uint8_t* Bs = smem + 2 * BM * PITCH;   // [2][BN][PITCH]
  const int8_t* A = reinterpret_cast<const int8_t*>(P.a);
  const int8_t* B = reinterpret_cast<const int8_t*>(P.b);
  const int M = P.m, N = P.n, K = P.k;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave >> 1) * (BM / 2), wn0 = (wave & 1) * (BN / 2);
  const int lr = lane & 31, lk = lane >> 5;
  u4g ra[AQ];
  unsigned rb[BQ][4];
  auto load = [&](int k0) {
#pragma unroll
    for (int q = 0; q < AQ; ++q) {
      const int u = tid + q * 256;
      if (AQ * 256 > AU && u >= AU) break;
      const int row = u >> 2, c = (u & 3) * 16;
      const int gr = m0 + row;
      if (!EDGE) {
        ra[q] = *reinterpret_cast<const u4g*>(A + (size_t)gr * K + k0 + c);
      } else {
        uint8_t v[16];
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          const int gk = k0 + c + e;
          v[e] = (gr < M && gk < K) ? (uint8_t)A[(size_t)gr * K + gk] : 0;
        }
        ra[q] = __builtin_bit_cast(u4g, v);
      }
    }
#pragma unroll
    for (int q = 0; q < BQ; ++q) {
      const int u = tid + q * 256;
      if (BQ * 256 > BU && u >= BU) break;
      const int kq = u / (BN / 4), nq = u % (BN / 4);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gk = k0 + 4 * kq + r, gn = n0 + 4 * nq;
        if (!EDGE) {
          rb[q][r] = *reinterpret_cast<const unsigned*>(B + (size_t)gk * N + gn);
        } else {
          unsigned w = 0;
#pragma unroll
          for (int e = 0; e < 4; ++e)
            if (gk < K && gn + e < N) w |= (unsigned)(uint8_t)B[(size_t)gk * N + gn + e] << (8 * e);
          rb[q][r] = w;
        }
      }
    }
  };
  auto store = [&](int buf) {
    uint8_t* as = As + buf * BM * PITCH;
    uint8_t* bs = Bs + buf * BN * PITCH;
#pragma unroll
    for (int q = 0; q < AQ; ++q) {
      const int u = tid + q * 256;
      if (AQ * 256 > AU && u >= AU) break;
      *reinterpret_cast<u4g*>(as + (u >> 2) * PITCH + (u & 3) * 16) = ra[q];
    }
#pragma unroll
    for (int q = 0; q < BQ; ++q) {
      const int u = tid + q * 256;
      if (BQ * 256 > BU && u >= BU) break;
      const int kq = u / (BN / 4), nq = u % (BN / 4);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const unsigned w = ((rb[q][0] >> (8 * j)) & 0xffu) | (((rb[q][1] >> (8 * j)) & 0xffu) << 8) |
                           (((rb[q][2] >> (8 * j)) & 0xffu) << 16) |
                           (((rb[q][3] >> (8 * j)) & 0xffu) << 24);
        *reinterpret_cast<unsigned*>(bs + (4 * nq + j) * PITCH + 4 * kq) = w;
      }
    }
  };
  i16g acc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) acc[i][j] = i16g{};
  const int nk = (K + KB - 1) / KB;
  load(0);
  store(0);
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    const bool more = kt + 1 < nk;
    if (more) load((kt + 1) * KB);
    const uint8_t* as = As + cur * BM * PITCH;
    const uint8_t* bs = Bs + cur * BN * PITCH;
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
      i4 af[TI], bf[TJ];
#pragma unroll
      for (int i = 0; i < TI; ++i)
        af[i] = *reinterpret_cast<const i4*>(as + (wm0 + i * 32 + lr) * PITCH + kk * 32 + lk * 16);
#pragma unroll
      for (int j = 0; j < TJ; ++j)
        bf[j] = *reinterpret_cast<const i4*>(bs + (wn0 + j * 32 + lr) * PITCH + kk * 32 + lk * 16);
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[i], bf[j], acc[i][j], 0, 0, 0);
    }
    if (more) store(cur ^ 1);
    __syncthreads();
  }
  int32_t* C = reinterpret_cast<int32_t*>(P.c);
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = m0 + wm0 + i * 32 + 8 * (r >> 2) + 4 * lk + (r & 3);
        const int col = n0 + wn0 + j * 32 + lr;
        if (EDGE && (row >= M || col >= N)) continue;
        int32_t* dst = C + (size_t)row * N + col;
        uint32_t v = (uint32_t)acc[i][j][r];
        if (P.accumulate) v += (uint32_t)*dst;
        *dst = (int32_t)v;
      }
}

template <int BM, int BN>
__device__ __noinline__ void i8_tile_dispatch(const GProblem& P, int m0, int n0, uint8_t* smem) {
  const bool interior = m0 + BM <= P.m && n0 + BN <= P.n && P.k % 64 == 0 && P.n % 4 == 0 &&
                        ((uintptr_t)P.a % 16) == 0 && ((uintptr_t)P.b % 4) == 0;
  if (interior) i8_tile<BM, BN, false>(P, m0, n0, smem);
  else i8_tile<BM, BN, true>(P, m0, n0, smem);
}

template <typename T, bool CPLX, int BM, int BN, int BK>
__device__ __noinline__ void real_dispatch(const GProblem& P, int m0, int n0, uint8_t* smem) {
  RealArgs<T> r;
  r.a = reinterpret_cast<const T*>(P.a);
  r.b = reinterpret_cast<const T*>(P.b);
  r.c = reinterpret_cast<T*>(P.c);
  r.M = P.m;
  r.N = CPLX ? 2 * P.n : P.n;
  r.K = CPLX ? 2 * P.k : P.k;
  r.lda = r.K;
  r.ldb = r.N;
  r.ldc = r.N;
  r.alpha_re = (T)P.ar;
  r.alpha_im = (T)P.ai;
  r.beta_re = (T)P.br;
  r.beta_im = (T)P.bi;
  r.alpha_one = (r.alpha_re == T(1) && r.alpha_im == T(0));
  r.beta_zero = (r.beta_re == T(0) && r.beta_im == T(0));
  T* As = reinterpret_cast<T*>(smem);
  T* Bs = As + 2 * BK * RealTileShape<T, BM, BN, BK>::A_PITCH;
  if (real_tile_interior<T, BM, BN, BK>(r, m0, n0))
    real_tile<T, CPLX, false, BM, BN, BK>(r, m0, n0, As, Bs);
  else
    real_tile<T, CPLX, true, BM, BN, BK>(r, m0, n0, As, Bs);
}

// Descriptor block: GHeader, then GProblem[nprob] ordered by tile_start.
struct GHeader {
  int nprob, ntiles;
  int pad[62];
};
static_assert(sizeof(GHeader) == 256, "header size");

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 8)))
void grouped_kernel(const uint8_t* __restrict__ desc) {
  __shared__ __attribute__((aligned(16))) uint8_t smem[G_LDS];
  const GHeader* h = reinterpret_cast<const GHeader*>(desc);
  const GProblem* probs = reinterpret_cast<const GProblem*>(desc + sizeof(GHeader));
  const int np = h->nprob, ntiles = h->ntiles;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    int lo = 0, hi = np - 1;  // last problem whose first tile is <= t
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (probs[mid].tile_start <= t) lo = mid;
      else hi = mid - 1;
    }
    const GProblem P = probs[lo];
    const int tile = t - P.tile_start;
    const int tm = tile / P.tiles_n, tn = tile % P.tiles_n;
    const int bm = P.cfg ? 128 : 64;
    const int m0 = tm * bm, n0 = tn * bm;
    switch (P.dtype * 2 + P.cfg) {
      case 0: real_dispatch<double, false, 64, 64, 8>(P, m0, n0, smem); break;
      case 1: real_dispatch<double, false, 128, 128, 8>(P, m0, n0, smem); break;
      case 2: real_dispatch<float, false, 64, 64, 16>(P, m0, n0, smem); break;
      case 3: real_dispatch<float, false, 128, 128, 16>(P, m0, n0, smem); break;
      case 4: real_dispatch<double, true, 64, 64, 8>(P, m0, n0, smem); break;
      case 5: real_dispatch<double, true, 128, 128, 8>(P, m0, n0, smem); break;
      case 6: real_dispatch<float, true, 64, 64, 16>(P, m0, n0, smem); break;
      case 7: real_dispatch<float, true, 128, 128, 16>(P, m0, n0, smem); break;
      case 8: i8_tile_dispatch<64, 64>(P, m0, n0, smem); break;
      case 9: i8_tile_dispatch<128, 128>(P, m0, n0, smem); break;
      default: break;
    }
    __syncthreads();  // the next tile reuses LDS
  }
}

// ---------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------

// Tile-shape table: cfg (0 = 64x64, 1 = 128x128) per dtype, size bucket
// b = clamp(ceil(log2(max(m, n_real))), 4, 11) - 4, small-side class
// s = [min(m, n_real) < 64, < 128, >= 128] and depth class
// q = [k_real < 128, < 512, >= 512] (k_real = 2k for complex).  Defaults:
// 128x128 from max side 256 up; set by ferray_set_tile_table (the autotuner's
// output).
constexpr int G_DTYPES = 5, G_BUCKETS = 8, G_SMALL = 3, G_KCLS = 3;
constexpr int G_CELLS = G_BUCKETS * G_SMALL * G_KCLS;
static std::mutex g_table_mu;
static int g_table[G_DTYPES][G_BUCKETS][G_SMALL][G_KCLS];
static bool g_table_init = [] {
  for (int d = 0; d < G_DTYPES; ++d)
    for (int b = 0; b < G_BUCKETS; ++b)
      for (int s = 0; s < G_SMALL; ++s)
        for (int q = 0; q < G_KCLS; ++q) g_table[d][b][s][q] = b >= 4 ? 1 : 0;
  return true;
}();

static int bucket_of(size_t m, size_t n_real) {
  size_t d = std::max(m, n_real);
  int b = 0;
  while (((size_t)1 << b) < d) ++b;
  b = std::min(std::max(b, 4), 11);
  return b - 4;
}
static int small_of(size_t m, size_t n_real) {
  const size_t d = std::min(m, n_real);
  return d < 64 ? 0 : (d < 128 ? 1 : 2);
}
static int kcls_of(size_t k_real) { return k_real < 128 ? 0 : (k_real < 512 ? 1 : 2); }

int grouped_cfg(int dtype, size_t m, size_t n, size_t k) {
  const bool cplx = dtype == FERRAY_DTYPE_C64 || dtype == FERRAY_DTYPE_C32;
  const size_t n_real = cplx ? 2 * n : n, k_real = cplx ? 2 * k : k;
  std::lock_guard<std::mutex> lk(g_table_mu);
  return g_table[dtype][bucket_of(m, n_real)][small_of(m, n_real)][kcls_of(k_real)];
}

static size_t problem_tiles(int dtype, int cfg, size_t m, size_t n, int* tiles_n) {
  const size_t bm = cfg ? 128 : 64;
  const size_t n_real = (dtype == FERRAY_DTYPE_C64 || dtype == FERRAY_DTYPE_C32) ? 2 * n : n;
  *tiles_n = (int)cdiv(n_real, bm);
  return cdiv(m, bm) * cdiv(n_real, bm);
}

static int check_problem(const ferray_gemm_problem& p) {
  if (p.dtype < 0 || p.dtype >= G_DTYPES) return FERRAY_ERR_INVALID;
  if (p.m > (1u << 30) || p.n > (1u << 29) || p.k > (1u << 29)) return FERRAY_ERR_INVALID;
  if (p.dtype == FERRAY_DTYPE_I8 && p.k > (1u << 16)) return FERRAY_ERR_UNSUPPORTED;
  if (p.m && p.n && !p.c) return FERRAY_ERR_INVALID;
  if (p.m && p.n && p.k && (!p.a || !p.b)) return FERRAY_ERR_INVALID;
  return FERRAY_OK;
}

// Problems with k == 0 keep the single-entry semantics exactly (C = beta*C,
// or zero / untouched for integers).
static int run_k0(const ferray_gemm_problem& p, hipStream_t st) {
  switch (p.dtype) {
    case FERRAY_DTYPE_F64:
      return ferray_gemm_f64_dev(p.m, p.n, 0, p.alpha_re, nullptr, nullptr, p.beta_re,
                                 (double*)p.c, st);
    case FERRAY_DTYPE_F32:
      return ferray_gemm_f32_dev(p.m, p.n, 0, (float)p.alpha_re, nullptr, nullptr,
                                 (float)p.beta_re, (float*)p.c, st);
    case FERRAY_DTYPE_C64:
      return ferray_gemm_c64_dev(p.m, p.n, 0, p.alpha_re, p.alpha_im, nullptr, nullptr,
                                 p.beta_re, p.beta_im, (double*)p.c, st);
    case FERRAY_DTYPE_C32:
      return ferray_gemm_c32_dev(p.m, p.n, 0, (float)p.alpha_re, (float)p.alpha_im, nullptr,
                                 nullptr, (float)p.beta_re, (float)p.beta_im, (float*)p.c, st);
    default:
      return ferray_gemm_i8_signed_dev(p.m, p.n, 0, nullptr, nullptr, (int32_t*)p.c,
                                       p.accumulate != 0, st);
  }
}

}  // namespace fg

// A plan: pinned host staging + device copy of the descriptor block for up
// to `max_problems` problems, and the graph of its launch.
struct ferray_grouped_plan {
  size_t max_problems = 0, max_tiles = 0;
  uint8_t* host = nullptr;  // GHeader + GProblem[max_problems]
  uint8_t* dev = nullptr;
  size_t bytes = 0;
  hipGraph_t graph = nullptr;
  hipGraphNode_t copy_node = nullptr, kernel_node = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t done = nullptr;
  int max_grid = 0;
  bool pending = false;
};

namespace fg {

static size_t plan_bytes(size_t probs) { return sizeof(GHeader) + probs * sizeof(GProblem); }

static int grid_blocks() {
  static int blocks = 0;
  if (!blocks) {
    hipDeviceProp_t prop;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
      return 512;
    blocks = prop.multiProcessorCount * 2;
  }
  return blocks;
}

// Fills the plan's host staging for `count` problems; returns the number of
// tiles through *ntiles_out and the number of descriptor bytes to upload
// through *bytes_out.
static int build_desc(ferray_grouped_plan* plan, const ferray_gemm_problem* probs, size_t count,
                      size_t* ntiles_out, size_t* bytes_out) {
  std::vector<std::pair<size_t, size_t>> order;  // (work, index)
  order.reserve(count);
  for (size_t i = 0; i < count; ++i) {
    const ferray_gemm_problem& p = probs[i];
    if (!p.m || !p.n || !p.k) continue;
    const size_t f = (p.dtype == FERRAY_DTYPE_C64 || p.dtype == FERRAY_DTYPE_C32) ? 4 : 1;
    order.push_back({p.m * p.n * p.k * f, i});
  }
  if (order.size() > plan->max_problems) return FERRAY_ERR_INVALID;
  std::stable_sort(order.begin(), order.end(),
                   [](const auto& x, const auto& y) { return x.first > y.first; });
  GProblem* gp = reinterpret_cast<GProblem*>(plan->host + sizeof(GHeader));
  size_t total = 0;
  for (size_t j = 0; j < order.size(); ++j) {
    const ferray_gemm_problem& p = probs[order[j].second];
    GProblem& g = gp[j];
    g.dtype = p.dtype;
    g.cfg = grouped_cfg(p.dtype, p.m, p.n, p.k);
    const size_t nt = problem_tiles(p.dtype, g.cfg, p.m, p.n, &g.tiles_n);
    if (total + nt > plan->max_tiles || total + nt > 0x7fffffff) return FERRAY_ERR_INVALID;
    g.tile_start = (int)total;
    total += nt;
    g.accumulate = p.accumulate;
    g.m = (int)p.m;
    g.n = (int)p.n;
    g.k = (int)p.k;
    g.a = p.a;
    g.b = p.b;
    g.c = p.c;
    g.ar = p.alpha_re;
    g.ai = p.alpha_im;
    g.br = p.beta_re;
    g.bi = p.beta_im;
  }
  GHeader* h = reinterpret_cast<GHeader*>(plan->host);
  h->nprob = (int)order.size();
  h->ntiles = (int)total;
  *ntiles_out = total;
  *bytes_out = plan_bytes(order.size());
  return FERRAY_OK;
}

// The launch graph: descriptor upload (pinned host -> device), then the
// persistent kernel.  Built once per plan; launches update the upload size
// and the grid.
static int plan_graph(ferray_grouped_plan* plan) {
  FG_CHECK(hipGraphCreate(&plan->graph, 0));
  FG_CHECK(hipGraphAddMemcpyNode1D(&plan->copy_node, plan->graph, nullptr, 0, plan->dev,
                                   plan->host, plan->bytes, hipMemcpyHostToDevice));
  hipKernelNodeParams kp = {};
  void* args[] = {&plan->dev};
  kp.func = reinterpret_cast<void*>(grouped_kernel);
  kp.gridDim = dim3(plan->max_grid);
  kp.blockDim = dim3(256);
  kp.sharedMemBytes = 0;
  kp.kernelParams = args;
  kp.extra = nullptr;
  FG_CHECK(hipGraphAddKernelNode(&plan->kernel_node, plan->graph, &plan->copy_node, 1, &kp));
  FG_CHECK(hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0));
  return FERRAY_OK;
}

static int plan_update(ferray_grouped_plan* plan, size_t bytes, int grid) {
  FG_CHECK(hipGraphExecMemcpyNodeSetParams1D(plan->exec, plan->copy_node, plan->dev, plan->host,
                                             bytes, hipMemcpyHostToDevice));
  hipKernelNodeParams kp = {};
  void* args[] = {&plan->dev};
  kp.func = reinterpret_cast<void*>(grouped_kernel);
  kp.gridDim = dim3(grid);
  kp.blockDim = dim3(256);
  kp.sharedMemBytes = 0;
  kp.kernelParams = args;
  kp.extra = nullptr;
  FG_CHECK(hipGraphExecKernelNodeSetParams(plan->exec, plan->kernel_node, &kp));
  return FERRAY_OK;
}

}  // namespace fg

using namespace fg;

extern "C" {

int ferray_set_tile_table(const int* table, size_t len) {
  const bool legacy = len == (size_t)G_DTYPES * G_BUCKETS;
  if (!table || (!legacy && len != (size_t)G_DTYPES * G_CELLS)) return FERRAY_ERR_INVALID;
  for (size_t i = 0; i < len; ++i)
    if (table[i] != 0 && table[i] != 1) return FERRAY_ERR_INVALID;
  std::lock_guard<std::mutex> lk(g_table_mu);
  for (int d = 0; d < G_DTYPES; ++d)
    for (int b = 0; b < G_BUCKETS; ++b)
      for (int s = 0; s < G_SMALL; ++s)
        for (int q = 0; q < G_KCLS; ++q)
          g_table[d][b][s][q] = legacy ? table[d * G_BUCKETS + b]
                                       : table[((d * G_BUCKETS + b) * G_SMALL + s) * G_KCLS + q];
  return FERRAY_OK;
}

int ferray_get_tile_table(int* table, size_t len) {
  const bool legacy = len == (size_t)G_DTYPES * G_BUCKETS;
  if (!table || (!legacy && len != (size_t)G_DTYPES * G_CELLS)) return FERRAY_ERR_INVALID;
  std::lock_guard<std::mutex> lk(g_table_mu);
  for (int d = 0; d < G_DTYPES; ++d)
    for (int b = 0; b < G_BUCKETS; ++b) {
      if (legacy) {
        table[d * G_BUCKETS + b] = g_table[d][b][G_SMALL - 1][G_KCLS - 1];
        continue;
      }
      for (int s = 0; s < G_SMALL; ++s)
        for (int q = 0; q < G_KCLS; ++q)
          table[((d * G_BUCKETS + b) * G_SMALL + s) * G_KCLS + q] = g_table[d][b][s][q];
    }
  return FERRAY_OK;
}

int ferray_grouped_plan_create(size_t max_problems, size_t max_tiles, ferray_grouped_plan** out) {
  int s = ensure_device();
  if (s) return s;
  if (!out || !max_problems || !max_tiles) return FERRAY_ERR_INVALID;
  *out = nullptr;
  auto* p = new (std::nothrow) ferray_grouped_plan();
  if (!p) return FERRAY_ERR_OOM;
  p->max_problems = max_problems;
  p->max_tiles = max_tiles;
  p->bytes = plan_bytes(max_problems);
  p->max_grid = grid_blocks();
  hipError_t e = hipHostMalloc((void**)&p->host, p->bytes, hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc((void**)&p->dev, p->bytes);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&p->done, hipEventDisableTiming);
  if (e != hipSuccess) {
    ferray_grouped_plan_destroy(p);
    return map_hip_error(e);
  }
  std::memset(p->host, 0, p->bytes);
  if ((s = plan_graph(p))) {
    ferray_grouped_plan_destroy(p);
    return s;
  }
  *out = p;
  return FERRAY_OK;
}

int ferray_grouped_plan_destroy(ferray_grouped_plan* p) {
  if (!p) return FERRAY_OK;
  if (p->pending && p->done) (void)hipEventSynchronize(p->done);
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->done) (void)hipEventDestroy(p->done);
  if (p->dev) (void)hipFree(p->dev);
  if (p->host) (void)hipHostFree(p->host);
  delete p;
  return FERRAY_OK;
}

int ferray_grouped_plan_launch(ferray_grouped_plan* p, const ferray_gemm_problem* problems,
                               size_t count, void* stream) {
  int s = ensure_device();
  if (s) return s;
  if (!p || (count && !problems)) return FERRAY_ERR_INVALID;
  hipStream_t st = resolve_stream(stream);
  for (size_t i = 0; i < count; ++i)
    if ((s = check_problem(problems[i]))) return s;
  // The staging area may still feed the previous replay.
  if (p->pending) FG_CHECK(hipEventSynchronize(p->done));
  p->pending = false;
  for (size_t i = 0; i < count; ++i)
    if (problems[i].m && problems[i].n && !problems[i].k)
      if ((s = run_k0(problems[i], st))) return s;
  size_t ntiles = 0, bytes = 0;
  if ((s = build_desc(p, problems, count, &ntiles, &bytes))) return s;
  if (ntiles == 0) return FERRAY_OK;
  if ((s = plan_update(p, bytes, (int)std::min<size_t>(ntiles, p->max_grid)))) return s;
  FG_CHECK(hipGraphLaunch(p->exec, st));
  FG_CHECK(hipEventRecord(p->done, st));
  p->pending = true;
  return FERRAY_OK;
}

int ferray_grouped_plan_sync(ferray_grouped_plan* p) {
  if (!p) return FERRAY_ERR_INVALID;
  if (p->pending) FG_CHECK(hipEventSynchronize(p->done));
  p->pending = false;
  return FERRAY_OK;
}

// One-shot grouped call through a per-thread plan that grows on demand.
int ferray_gemm_grouped_dev(const ferray_gemm_problem* problems, size_t count, void* stream) {
  int s = ensure_device();
  if (s) return s;
  if (count == 0) return FERRAY_OK;
  if (!problems) return FERRAY_ERR_INVALID;
  size_t tiles = 0;
  for (size_t i = 0; i < count; ++i) {
    if ((s = check_problem(problems[i]))) return s;
    const ferray_gemm_problem& p = problems[i];
    int tn;
    if (p.m && p.n && p.k) tiles += problem_tiles(p.dtype, grouped_cfg(p.dtype, p.m, p.n, p.k), p.m, p.n, &tn);
  }
  // The per-thread plan lives until the thread ends; it is not released in a
  // destructor because thread teardown can run after the HIP runtime's own.
  struct Holder {
    ferray_grouped_plan* plan = nullptr;
  };
  thread_local Holder h;
  if (!h.plan || h.plan->max_problems < count || h.plan->max_tiles < tiles) {
    ferray_grouped_plan_destroy(h.plan);
    h.plan = nullptr;
    if ((s = ferray_grouped_plan_create(std::max(count, (size_t)64), std::max(tiles, (size_t)4096),
                                        &h.plan)))
      return s;
  }
  return ferray_grouped_plan_launch(h.plan, problems, count, stream);
}

int ferray_matmul_batched_dev(int dtype, size_t batch, size_t a_batch, size_t b_batch, size_t m,
                              size_t k, size_t n, const void* a, const void* b, void* c,
                              void* stream) {
  if (dtype < FERRAY_DTYPE_F64 || dtype > FERRAY_DTYPE_I8) return FERRAY_ERR_INVALID;
  if (batch == 0 || m == 0 || n == 0) return FERRAY_OK;
  if (a_batch == 0 || b_batch == 0 || !c || (k && (!a || !b))) return FERRAY_ERR_INVALID;
  const size_t e = (dtype == FERRAY_DTYPE_F64 || dtype == FERRAY_DTYPE_C32) ? 8
                   : dtype == FERRAY_DTYPE_F32                              ? 4
                   : dtype == FERRAY_DTYPE_C64                              ? 16
                                                                            : 1;
  const size_t ce = dtype == FERRAY_DTYPE_I8 ? 4 : e;
  std::vector<ferray_gemm_problem> probs(batch);
  for (size_t i = 0; i < batch; ++i) {
    ferray_gemm_problem& p = probs[i];
    p.dtype = dtype;
    p.accumulate = 0;
    p.m = m;
    p.n = n;
    p.k = k;
    p.a = a ? static_cast<const uint8_t*>(a) + (i % a_batch) * m * k * e : nullptr;
    p.b = b ? static_cast<const uint8_t*>(b) + (i % b_batch) * k * n * e : nullptr;
    p.c = static_cast<uint8_t*>(c) + i * m * n * ce;
    p.alpha_re = 1.0;
    p.alpha_im = 0.0;
    p.beta_re = 0.0;
    p.beta_im = 0.0;
  }
  return ferray_gemm_grouped_dev(probs.data(), batch, stream);
}

}  // extern "C"
