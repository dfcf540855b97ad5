// Workgroup renumbering for the 8 XCDs of gfx950.
#pragma once

#include <hip/hip_runtime.h>

namespace cxp {

// The hardware hands workgroup i of a launch to XCD i % 8, and each XCD has
// its own L2.  Tiled kernels whose neighbouring tiles share halo data take
// their tile number from xcd_block(blockIdx.x, gridDim.x): a bijection of
// [0, nb) under which XCD x runs the contiguous tile range
// [x q + min(x, r), ...) (q = nb / 8, r = nb % 8), so neighbours' halos
// meet in one L2.
__device__ __forceinline__ int xcd_block(int i, int nb) {
  const int q = nb / 8, r = nb % 8, x = i % 8;
  return x * q + (x < r ? x : r) + i / 8;
}

} // namespace cxp
