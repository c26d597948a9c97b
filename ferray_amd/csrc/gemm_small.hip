// Real and complex GEMMs too small to fill the chip with 128 x 128 tiles.
//
// A product with fewer 128 x 128 tiles than the chip has block slots leaves
// CUs idle (1000^3 is 64 tiles for 256 CUs).  These run on 64 x 64 tiles
// (four times the blocks), and products with fewer than 192 of those on
// 32 x 32 tiles (one 16 x 16 MFMA tile per wave), with the tile routine of
// the grouped kernel (real_tile.hpp).  That routine accumulates every output
// element over the same MFMA k-sequence as the 128 x 128 kernel, so the
// results are bit-identical whichever path a size takes.
#include "real_tile.hpp"

namespace fg {

constexpr int SM_GROUP_M = 8;

// k-depth of one stage: f64 8 (33 KiB of LDS per block), f32 16.
template <typename T>
struct SmallStageK {
  static constexpr int value = sizeof(T) == 8 ? 8 : 16;
};

template <typename T, bool CPLX, int SM_TILE>
__global__ __launch_bounds__(RT_THREADS) void gemm_small_kernel(RealArgs<T> p) {
  constexpr int BK = SmallStageK<T>::value;
  using S = RealTileShape<T, SM_TILE, SM_TILE, BK>;
  __shared__ __attribute__((aligned(16))) T smem[S::lds_bytes / sizeof(T)];
  const int tiles_m = (p.M + SM_TILE - 1) / SM_TILE, tiles_n = (p.N + SM_TILE - 1) / SM_TILE;
  // Tile order as the 128 x 128 kernel: groups of SM_GROUP_M tile rows, and
  // each XCD (blocks are dealt round-robin to the 8 XCDs) takes a contiguous
  // range of that order.
  const int nt = tiles_m * tiles_n;
  int bid = blockIdx.x;
  if ((nt & 7) == 0) bid = (bid & 7) * (nt >> 3) + (bid >> 3);
  const int per_group = SM_GROUP_M * tiles_n, group = bid / per_group;
  const int first_m = group * SM_GROUP_M, gsz = min(tiles_m - first_m, SM_GROUP_M);
  const int in_group = bid - group * per_group;
  const int m0 = (first_m + in_group % gsz) * SM_TILE, n0 = (in_group / gsz) * SM_TILE;
  T* As = smem;
  T* Bs = smem + 2 * BK * S::A_PITCH;
  if (real_tile_interior<T, SM_TILE, SM_TILE, BK>(p, m0, n0))
    real_tile<T, CPLX, false, SM_TILE, SM_TILE, BK>(p, m0, n0, As, Bs);
  else
    real_tile<T, CPLX, true, SM_TILE, SM_TILE, BK>(p, m0, n0, As, Bs);
}

template <typename T, bool CPLX>
int launch_real_small(const RealArgs<T>& p, hipStream_t st) {
  const size_t tiles = cdiv((size_t)p.M, 64) * cdiv((size_t)p.N, 64);
  if (tiles == 0) return FERRAY_OK;
  if (tiles > 0x7fffffff) return FERRAY_ERR_INVALID;
  // 32 x 32 tiles while 64 x 64 tiles would leave a quarter of the 256 CUs
  // idle (measured, f64: 512^3 27.9 -> 16.0 us, 768^3 46.7 -> 39.5 us,
  // 1024 x 512 x 1024 60.5 -> 39.1 us, c64 512^3 87 -> 39 us; at 900^3, 225
  // tiles, the 64 x 64 form stays faster: 68.8 against 83.4 us)
  if (tiles < 192) {
    const size_t t32 = cdiv((size_t)p.M, 32) * cdiv((size_t)p.N, 32);
    hipLaunchKernelGGL((gemm_small_kernel<T, CPLX, 32>), dim3((unsigned)t32), dim3(RT_THREADS), 0,
                       st, p);
  } else {
    hipLaunchKernelGGL((gemm_small_kernel<T, CPLX, 64>), dim3((unsigned)tiles), dim3(RT_THREADS),
                       0, st, p);
  }
  FG_CHECK(hipGetLastError());
  return FERRAY_OK;
}

template int launch_real_small<double, false>(const RealArgs<double>&, hipStream_t);
template int launch_real_small<double, true>(const RealArgs<double>&, hipStream_t);
template int launch_real_small<float, false>(const RealArgs<float>&, hipStream_t);
template int launch_real_small<float, true>(const RealArgs<float>&, hipStream_t);

}  // namespace fg
