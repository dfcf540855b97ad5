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

// Several regions of one (coarse, fine) FAB pair in one launch, without a
// work list: up to CXP_MAX_REGIONS regions, region i owning blocks
// [start[i], start[i + 1]) of grid.x (thin tiles, or separable tiles of
// width 2^lx[i] and one row pair per thread with z-chunks of zchunk[i]
// planes); grid.y = component.  The regions must not overlap: each tile
// writes its own points.
#define CXP_MAX_REGIONS 8
struct RegionLaunch {
  SepTask t[CXP_MAX_REGIONS];
  int start[CXP_MAX_REGIONS + 1];
  int ntx[CXP_MAX_REGIONS], nty[CXP_MAX_REGIONS];
  int lx[CXP_MAX_REGIONS];  // -1: thin tiles (16 x 16, the whole z range)
  int zchunk[CXP_MAX_REGIONS];
  int n;
};

__global__ void __launch_bounds__(CXP_NT) k_sep_regions(RegionLaunch r) {
  extern __shared__ double reg_smem[];
  const int b = xcd_block(blockIdx.x, gridDim.x);
  int i = 0;
  while (i + 1 < r.n && b >= r.start[i + 1])
    ++i;
  const SepTask &t = r.t[i];
  const int local = b - r.start[i];
  const int ntx = r.ntx[i];
  if (r.lx[i] < 0) {
    const int x0 = (t.rlo[0] & ~1) + 16 * (local % ntx);
    const int y0 = (t.rlo[1] & ~1) + 16 * (local / ntx);
    thin_tile(t, blockIdx.y, x0, y0, reg_smem);
    return;
  }
  const Geom g = make_geom(r.lx[i], 1);
  const int nty = r.nty[i], zc = r.zchunk[i];
  const int x0 = (t.rlo[0] & ~1) + g.tx * (local % ntx);
  const int y0 = (t.rlo[1] & ~1) + g.ty * ((local / ntx) % nty);
  const int z0 = t.rlo[2] + zc * (local / (ntx * nty));
  const int z1 = min(t.rhi[2], z0 + zc - 1);
  sep_tile<-1, 1>(g, t, blockIdx.y, x0, y0, z0, z1, reg_smem);
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
