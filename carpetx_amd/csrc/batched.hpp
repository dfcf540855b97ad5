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
//
// Dynamic tail: the static ranges hold most of the work, but the SIMDs
// issue the oldest resident wave first, so workgroups dispatched later on a
// compute unit progress more slowly than any fixed estimate says.  The
// smallest tiles of each XCD class x therefore sit in a queue, tiles
// [q.start[x], q.start[x + 1]), which the class's workgroups claim q.chunk
// at a time through counter q.ctr[32 x] once their own range is done (the
// next claim is issued while the current chunk runs).  The
// last workgroup to finish (counter q.ctr[32 q.nq]) zeroes the counters for
// the next launch, so launches of one list must not overlap: the host orders
// a launch on another stream after the list's previous launch.
struct FillQueue {
  int *ctr;  // null: no queue
  int start[9];
  int nq;
  int chunk;
};

template <int SU>
__device__ __forceinline__ void fill_one(const SepTask *__restrict__ tasks,
                                         const SepTile *__restrict__ tiles,
                                         int n, double *smem) {
  const SepTile tl = tiles[n];
  const SepTask t = tasks[tl.task];
  if (tl.cfg < 0) {
    thin_tile(t, tl.comp, tl.x0, tl.y0, smem);
  } else if (tl.cfg < 16) {
    sep_tile<SU, 1>(make_geom(tl.cfg), t, tl.comp, tl.x0, tl.y0, tl.z0,
                    tl.z1, smem);
  } else if constexpr (SU >= 0) {
    // two row pairs per thread: lists of one stencil only (the planner
    // keeps mixed lists to one row pair, whose registers the dispatching
    // kernel could not hold for both)
    sep_tile<SU, 2>(make_geom(tl.cfg - 16, 2), t, tl.comp, tl.x0, tl.y0,
                    tl.z0, tl.z1, smem);
  }
  __syncthreads();  // the next tile rewrites the LDS buffers
}

template <int SU>
__global__ void __launch_bounds__(CXP_NT, 4)
    k_fill_batched(const SepTask *__restrict__ tasks,
                   const SepTile *__restrict__ tiles,
                   const int *__restrict__ wg_start, FillQueue q) {
  extern __shared__ double fill_smem[];
  __shared__ int claim;
  const int n0 = wg_start[blockIdx.x], n1 = wg_start[blockIdx.x + 1];
  for (int n = n0; n < n1; ++n)
    fill_one<SU>(tasks, tiles, n, fill_smem);
  if (q.ctr == nullptr)
    return;
  const int x = blockIdx.x % q.nq;
  const int p0 = q.start[x], p1 = q.start[x + 1];
  // Thread 0 claims the next chunk before the current one runs, so the
  // atomic's round trip overlaps the tile instead of preceding it; the last
  // claim of each workgroup goes unused.
  int next = 0;
  if (threadIdx.x == 0)
    next = __hip_atomic_fetch_add(q.ctr + 32 * x, 1, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
  for (;;) {
    if (threadIdx.x == 0)
      claim = p0 + q.chunk * next;
    __syncthreads();
    // every later write of claim follows at least one tile's barrier
    const int c0 = claim;
    if (c0 >= p1)
      break;
    if (threadIdx.x == 0)
      next = __hip_atomic_fetch_add(q.ctr + 32 * x, 1, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
    const int c1 = min(c0 + q.chunk, p1);
    for (int n = c0; n < c1; ++n)
      fill_one<SU>(tasks, tiles, n, fill_smem);
  }
  if (threadIdx.x == 0 &&
      __hip_atomic_fetch_add(q.ctr + 32 * q.nq, 1, __ATOMIC_ACQ_REL,
                             __HIP_MEMORY_SCOPE_AGENT) ==
          (int)gridDim.x - 1) {
    for (int i = 0; i <= q.nq; ++i)
      __hip_atomic_store(q.ctr + 32 * i, 0, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
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
