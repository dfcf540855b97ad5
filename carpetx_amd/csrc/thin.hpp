// Separable prolongation of patches at most CXP_THIN_NZ fine planes thick
// in z (the z-slabs of a ghost shell).
//
// The z-marching tile routine spends most of its plane iterations on such a
// patch filling and draining its register window: a 3-plane slab needs the
// span of the z stencil plus two coarse planes for three fine planes of
// output.  Here a workgroup of 128 threads owns a 16 x 16 fine (x, y) tile
// and the patch's whole z range, and loads the tile's coarse block once:
//
//   coarse block (CZ planes x CY rows x CX)          -> LDS rows
//   x-pass, one row per thread, in place              -> LDS rows
//   y-pass, one (fine x, fine y pair) per thread, all coarse planes
//                                                      -> registers
//   z-pass from those registers, all fine z            -> fine FAB
//
// Every pass applies the reference 1-D operator in the order x, y, z, as the
// z-marching routine does, so results are bit-identical to it.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace cxp {

#define CXP_THIN_NZ 4   // largest fine z extent of a thin patch
// LDS pitch of a block row (doubles).  The x-pass gives each thread a row,
// so a wave's lanes are P doubles apart: an odd P spreads them over every
// bank pair, where an even one folds them onto a few (P = 18 measured 5.8%
// faster than 16 on the uneven level for that reason).
#define CXP_THIN_P 17
#define CXP_THIN_KMAX 11  // coarse planes of the block: span_z + 3 <= 11

// Static extents of the coarse block for stencil windows of span s = whi-wlo:
// 8 + s coarse x and y, s + 3 coarse planes (4 fine planes reach 3).
__host__ __device__ inline int thin_rows(int span_y) { return 8 + span_y; }
__host__ __device__ inline int thin_planes(int span_z) { return span_z + 3; }

// LDS doubles for a task's stencil windows: one row of CXP_THIN_P doubles
// per (coarse plane, coarse y) of the block.
__host__ __device__ inline int thin_lds_doubles(int span_y, int span_z) {
  return thin_planes(span_z) * thin_rows(span_y) * CXP_THIN_P;
}

__host__ __device__ inline bool thin_eligible(const SepTask &t) {
  if (t.rhi[2] - t.rlo[2] + 1 > CXP_THIN_NZ)
    return false;
  for (int d = 0; d < 3; ++d)
    if (t.whi[d] - t.wlo[d] > 8)
      return false;
  return true;
}

// x-pass of one block row, in place: the row's 8 + span coarse values are
// replaced by the 16 fine-x values of the tile.
template <int S>
__device__ __forceinline__ void thin_xrow(double *__restrict__ row) {
  constexpr int wlo = sid_wlo(S);
  constexpr int n = 8 + sid_whi(S) - sid_wlo(S);
  double c[n];
#pragma unroll
  for (int i = 0; i < n; ++i)
    c[i] = row[i];
#pragma unroll
  for (int fx = 0; fx < 16; fx += 2) {
    const auto acc = [&](int o) { return c[(fx >> 1) - wlo + o]; };
    // two 8-byte stores: with an odd pitch a row is only 8-byte aligned
    row[fx] = Op1<S>::apply(acc, 0);
    row[fx + 1] = Op1<S>::apply(acc, 1);
  }
}

// y-pass of one fine x and the fine y pair of coarse row j (block-local
// j = fyp - wlo), for every coarse plane k < CZ of the block: y[k][off].
template <int S>
__device__ __forceinline__ void thin_ypair(const double *__restrict__ xs,
                                           int cyn, int fyp, int CZ,
                                           double (&y)[CXP_THIN_KMAX][2]) {
  constexpr int wlo = sid_wlo(S);
  constexpr int lo = sid_wlo(S), hi = sid_whi(S);
#pragma unroll
  for (int k = 0; k < CXP_THIN_KMAX; ++k) {
    if (k < CZ) {
      const double *col = xs + (k * cyn + fyp - wlo) * CXP_THIN_P;
      double c[hi - lo + 1];
#pragma unroll
      for (int o = lo; o <= hi; ++o)
        c[o - lo] = col[o * CXP_THIN_P];
      const auto acc = [&](int o) { return c[o - lo]; };
      y[k][0] = Op1<S>::apply(acc, 0);
      y[k][1] = Op1<S>::apply(acc, 1);
    }
  }
}

// z-pass of one fine (x, y) column: the nzt <= 4 fine planes from z0.  With
// P = z0 & 1, fine plane z0 + dz reads coarse planes (dz + P) / 2 - wlo + o
// of the block at offset (P + dz) & 1.
template <int S, int P>
__device__ __forceinline__ void thin_zcol(const double (&y)[CXP_THIN_KMAX][2],
                                          int h, int nzt, double (&out)[4]) {
  constexpr int wlo = sid_wlo(S);
#pragma unroll
  for (int dz = 0; dz < 4; ++dz) {
    if (dz < nzt)
      out[dz] = Op1<S>::apply(
          [&](int o) { return y[((dz + P) >> 1) - wlo + o][h]; }, (P + dz) & 1);
  }
}

// One thin tile.  All threads of the block call it; smem holds
// thin_lds_doubles(span_y, span_z) doubles.
__device__ __forceinline__ void thin_tile(const SepTask &t, long long comp,
                                          int x0, int y0, double *smem) {
  const int tid = threadIdx.x;
  const int CYn = thin_rows(t.whi[1] - t.wlo[1]);
  const int z0 = t.rlo[2], z1 = t.rhi[2];
  const int nzt = z1 - z0 + 1;
  const int xe = min(x0 + 15, t.rhi[0]);
  const int ye = min(y0 + 15, t.rhi[1]);
  const int ci0 = (x0 >> 1) + t.wlo[0];
  const int cj0 = (y0 >> 1) + t.wlo[1];
  const int kf = (z0 >> 1) + t.wlo[2];
  const int CX = (xe >> 1) + t.whi[0] - ci0 + 1;   // columns present
  const int CY = (ye >> 1) + t.whi[1] - cj0 + 1;   // rows present
  const int CZ = (z1 >> 1) + t.whi[2] - kf + 1;    // planes present

  DFab crse = t.crse;
  crse.p += comp * t.ccs;
  DFab fine = t.fine;
  fine.p += comp * t.fcs;
  const __amdgpu_buffer_rsrc_t rc = make_rsrc(crse.p, t.cbytes);
  const __amdgpu_buffer_rsrc_t rf = make_rsrc(fine.p, t.fbytes);
  const int nrows = CZ * CYn;

  // coarse block -> LDS rows k CYn + cy: lane lx = coarse x, and thread row
  // group ly takes rows cy = ly and ly + 8 of every plane k < CZ (CYn <= 16),
  // so a lane's two row offsets and masks hold for all planes and the plane
  // step k sz is a uniform offset.  Absent columns and rows read as 0.
  {
    const int lx = tid & 15, ly = tid >> 4;
    const bool xok = lx < CX;
    const bool hok1 = ly + 8 < CYn;
    const unsigned sy8 = 8u * (unsigned)crse.sy, sz8 = 8u * (unsigned)crse.sz;
    const unsigned base = 8u * (unsigned)(ci0 + lx - crse.lo[0]) +
                          (unsigned)(cj0 + ly - crse.lo[1]) * sy8 +
                          (unsigned)(kf - crse.lo[2]) * sz8;
    const int vo0 = (xok && ly < CY) ? (int)base : CXP_OOB;
    const int vo1 = (xok && ly + 8 < CY) ? (int)(base + 8u * sy8) : CXP_OOB;
    double v[CXP_THIN_KMAX][2];
#pragma unroll
    for (int k = 0; k < CXP_THIN_KMAX; ++k) {
      if (k < CZ) {
        v[k][0] = bld_s(rc, vo0, (int)(k * sz8));
        if (CYn > 8)
          v[k][1] = bld_s(rc, vo1, (int)(k * sz8));
      }
    }
    double *row0 = smem + ly * CXP_THIN_P + lx;
#pragma unroll
    for (int k = 0; k < CXP_THIN_KMAX; ++k) {
      if (k < CZ) {
        row0[k * CYn * CXP_THIN_P] = v[k][0];
        if (hok1)
          row0[(k * CYn + 8) * CXP_THIN_P] = v[k][1];
      }
    }
  }
  __syncthreads();
  // x-pass, one row per thread, in place
  for (int rho = tid; rho < nrows; rho += CXP_NT) {
    switch (t.sid[0]) {
#define CXP_TX(S)                                                              \
  case S:                                                                      \
    thin_xrow<S>(smem + rho * CXP_THIN_P);                                     \
    break;
      CXP_LINEAR_SIDS(CXP_TX)
#undef CXP_TX
    }
  }
  __syncthreads();
  // y-pass into registers, then z-pass: fine x = x0 + fx, fine y pair
  // y0 + 2 fyp + {0, 1}
  const int fx = tid & 15, fyp = tid >> 4;
  double y[CXP_THIN_KMAX][2];
  switch (t.sid[1]) {
#define CXP_TY(S)                                                              \
  case S:                                                                      \
    thin_ypair<S>(smem + fx, CYn, fyp, CZ, y);                                 \
    break;
    CXP_LINEAR_SIDS(CXP_TY)
#undef CXP_TY
  }
  const int gx = x0 + fx;
  const bool xin = gx >= t.rlo[0] && gx <= t.rhi[0];
  const int par = z0 & 1;
  const unsigned fsz8 = 8u * (unsigned)fine.sz;
  const unsigned fz0 = (unsigned)(z0 - fine.lo[2]) * fsz8;
  unsigned fo = 8u * (unsigned)(gx - fine.lo[0]) +
                (unsigned)(y0 + 2 * fyp - fine.lo[1]) * (8u * (unsigned)fine.sy);
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int gy = y0 + 2 * fyp + h;
    const bool own = xin && gy >= t.rlo[1] && gy <= t.rhi[1];
    double out[4] = {0.0, 0.0, 0.0, 0.0};
    switch (t.sid[2] * 2 + par) {
#define CXP_TZ(S)                                                              \
  case 2 * S:                                                                  \
    thin_zcol<S, 0>(y, h, nzt, out);                                           \
    break;                                                                     \
  case 2 * S + 1:                                                              \
    thin_zcol<S, 1>(y, h, nzt, out);                                           \
    break;
      CXP_LINEAR_SIDS(CXP_TZ)
#undef CXP_TZ
    }
#pragma unroll
    for (int dz = 0; dz < 4; ++dz) {
      if (dz < nzt)
        bst(rf, own ? (int)(fo + fz0 + dz * fsz8) : CXP_OOB, out[dz]);
    }
    fo += 8u * (unsigned)fine.sy;
  }
}

// Per-box launch: grid.x = 16 x 16 tiles (x fastest), grid.y = component.
__global__ void __launch_bounds__(CXP_NT)
    k_thin_box(SepTask t, int ntx) {
  extern __shared__ double thin_smem[];
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int x0 = (t.rlo[0] & ~1) + 16 * (b % ntx);
  const int y0 = (t.rlo[1] & ~1) + 16 * (b / ntx);
  thin_tile(t, blockIdx.y, x0, y0, thin_smem);
}

} // namespace cxp
