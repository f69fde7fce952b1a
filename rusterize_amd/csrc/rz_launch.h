// rz_launch.h -- host launchers of the tile kernels, one pair per output
// dtype.  Each is defined in its own translation unit (rz_tile_inst.hip
// compiled once per dtype), so the k_tile / k_tile_heavy instantiations of
// the ten dtypes build in parallel.
#pragma once

#include <hip/hip_runtime.h>

#include "rz_kernels.h"

namespace rzk {

#define RZ_TILE_LAUNCHERS(NAME)                                                      \
  void launch_tile_##NAME(int fun, const TileArgs& a, bool lines, hipStream_t s);    \
  void launch_heavy_##NAME(int fun, const TileArgs& a, hipStream_t s);
RZ_TILE_LAUNCHERS(u8)
RZ_TILE_LAUNCHERS(u16)
RZ_TILE_LAUNCHERS(u32)
RZ_TILE_LAUNCHERS(u64)
RZ_TILE_LAUNCHERS(i8)
RZ_TILE_LAUNCHERS(i16)
RZ_TILE_LAUNCHERS(i32)
RZ_TILE_LAUNCHERS(i64)
RZ_TILE_LAUNCHERS(f32)
RZ_TILE_LAUNCHERS(f64)
#undef RZ_TILE_LAUNCHERS

}  // namespace rzk
