// The persistent batched fill: one launch walks a device-resident list of
// separable tiles (z-marching, sep_tile) and thin tiles (whole-block,
// thin_tile) of many patches and components.
#pragma once

#include "kernels.hpp"
#include "thin.hpp"

namespace cxp {

// The host splits the list into one contiguous range per workgroup,
// [wg_start[b], wg_start[b+1]):
//  * (patch, component) tile groups are assigned to XCDs balancing estimated
//    cost; workgroups are dispatched round-robin over the 8 XCDs, so block b
//    runs on XCD b % 8, and neighbouring tiles of a group share coarse halo
//    in that XCD's L2;
//  * each XCD's tiles go, longest first, to the least loaded of its
//    workgroups, so no workgroup ends much later than the others;
//  * each workgroup's range interleaves thin and separable tiles in
//    proportion to their cost, so both kinds are in flight at all times.
// Dynamic LDS: the larger of the separable routine's buffer and the largest
// thin block of the list.
// SU >= 0: every separable task of the list uses stencil SU in all three
// directions (a list of one VVV or CCC operator), and the marching routine
// is compiled for it without dispatch.
template <int SU>
__global__ void __launch_bounds__(CXP_NT, 4)
    k_fill_batched(const SepTask *__restrict__ tasks,
                   const SepTile *__restrict__ tiles,
                   const int *__restrict__ wg_start) {
  extern __shared__ double fill_smem[];
  const int n0 = wg_start[blockIdx.x], n1 = wg_start[blockIdx.x + 1];
  for (int n = n0; n < n1; ++n) {
    const SepTile tl = tiles[n];
    const SepTask t = tasks[tl.task];
    if (tl.cfg < 0)
      thin_tile(t, tl.comp, tl.x0, tl.y0, fill_smem);
    else
      sep_tile<SU>(make_geom(tl.cfg), t, tl.comp, tl.x0, tl.y0, tl.z0, tl.z1,
                   fill_smem);
    __syncthreads();  // the next tile rewrites the LDS buffers
  }
}

// Calls f(kernel) with the k_fill_batched instance for uniform stencil su
// (or the dispatching one for su < 0).
template <class F> inline auto with_fill_kernel(int su, F &&f) {
  switch (su) {
#define CXP_FK(S)                                                              \
  case S:                                                                      \
    return f(k_fill_batched<S>);
    CXP_LINEAR_SIDS(CXP_FK)
#undef CXP_FK
  }
  return f(k_fill_batched<-1>);
}

} // namespace cxp
