// rz_tile_inst.hip -- the k_tile / k_tile_heavy instantiations of one output
// dtype and their launchers.  Compiled once per dtype with
// -DRZ_TILE_T=<C type> -DRZ_TILE_NAME=<u8|u16|...|f64> (rusterize_amd/build.py).
#include <algorithm>
#include <cstdlib>

#include "rz_launch.h"
#include "rz_tile_t.h"

#if !defined(RZ_TILE_T) || !defined(RZ_TILE_NAME)
#error "compile with -DRZ_TILE_T=<type> -DRZ_TILE_NAME=<name>"
#endif

namespace rzk {
namespace {

using T = RZ_TILE_T;

// Persistent grid for one tile-kernel instantiation: enough workgroups to
// fill every CU at the kernel's occupancy (no more than the units need), a
// multiple of the 8 XCDs.
template <typename K>
uint32_t tile_grid(K kernel, uint32_t& cached, uint64_t n_tiles) {
  if (!cached) {
    int dev = 0, cus = 256, per_cu = 1;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0) != hipSuccess || per_cu < 1)
      per_cu = 1;
    cached = (uint32_t)(cus * per_cu);
  }
  const uint32_t g = (uint32_t)std::min<uint64_t>(cached, std::max<uint64_t>(n_tiles, 1));
  return std::max<uint32_t>(NUM_XCD, (g + NUM_XCD - 1) / NUM_XCD * NUM_XCD);
}

// Tiles per queue grab: the tuned value, else 4 once every wave has many
// tiles (measured best of 1, 2, 4 on the 40,000^2 workload), else 1.
// RUSTERIZE_CHUNK overrides both.
inline uint32_t queue_chunk(uint64_t n_tiles, uint32_t grid, uint32_t tuned) {
  if (const char* e = std::getenv("RUSTERIZE_CHUNK")) {
    const int v = std::atoi(e);
    if (v >= 1 && v <= 4096) return (uint32_t)v;
  }
  if (tuned) return tuned;
  return n_tiles >= (uint64_t)grid * 16 ? 4u : 1u;
}

template <int FUN>
void launch_tile(TileArgs a, bool lines, hipStream_t s) {
  const uint64_t tiles = (uint64_t)(a.t_hi - a.t_lo);
  if (lines) {
    static uint32_t cached = 0;
    const uint32_t g = tile_grid(k_tile_lp<T, FUN>, cached, tiles);
    a.chunk = queue_chunk(tiles, g, a.chunk);
    hipLaunchKernelGGL((k_tile_lp<T, FUN>), dim3(g), dim3(256), 0, s, a);
  } else {
    static uint32_t cached = 0;
    const uint32_t g = tile_grid(k_tile<T, FUN>, cached, tiles);
    a.chunk = queue_chunk(tiles, g, a.chunk);
    hipLaunchKernelGGL((k_tile<T, FUN>), dim3(g), dim3(256), 0, s, a);
  }
}

// k_tile_heavy: a persistent grid (the work counter hands out units)
template <int FUN>
void launch_heavy(const TileArgs& a, hipStream_t s) {
  static uint32_t cached = 0;
  if (!cached) {
    int dev = 0, cus = 256, per_cu = 1;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_tile_heavy<T, FUN>, 256, 0) != hipSuccess ||
        per_cu < 1)
      per_cu = 1;
    cached = (uint32_t)(cus * per_cu);
  }
  hipLaunchKernelGGL((k_tile_heavy<T, FUN>), dim3(cached), dim3(256), 0, s, a);
}

}  // namespace

#define RZ_CAT2(a, b) a##b
#define RZ_CAT(a, b) RZ_CAT2(a, b)

void RZ_CAT(launch_tile_, RZ_TILE_NAME)(int fun, const TileArgs& a, bool lines, hipStream_t s) {
  switch (fun) {
    case 0: launch_tile<0>(a, lines, s); break;
    case 1: launch_tile<1>(a, lines, s); break;
    case 2: launch_tile<2>(a, lines, s); break;
    case 3: launch_tile<3>(a, lines, s); break;
    case 4: launch_tile<4>(a, lines, s); break;
    case 5: launch_tile<5>(a, lines, s); break;
    default: launch_tile<6>(a, lines, s); break;
  }
}

void RZ_CAT(launch_heavy_, RZ_TILE_NAME)(int fun, const TileArgs& a, hipStream_t s) {
  switch (fun) {
    case 0: launch_heavy<0>(a, s); break;
    case 1: launch_heavy<1>(a, s); break;
    case 2: launch_heavy<2>(a, s); break;
    case 3: launch_heavy<3>(a, s); break;
    case 4: launch_heavy<4>(a, s); break;
    case 5: launch_heavy<5>(a, s); break;
    default: launch_heavy<6>(a, s); break;
  }
}

}  // namespace rzk
