// Adjoint (transpose) of the rf=2 prolongation: crse_bar += P^T fine_bar.
//
// The reference has no adjoints; these kernels are written from the forward
// operators.  All of them gather over coarse points, so no atomics are
// needed:
//  * linear operators (POLY, HERMITE, CONS): P = Pz (x) Py (x) Px, so
//    P^T = Px^T (x) Py^T (x) Pz^T and the three transposes commute.  A
//    workgroup owns a coarse (x, y) tile, marches through fine z-planes,
//    applies x^T and y^T inside each plane (LDS), and accumulates z^T into a
//    9-slot register window of coarse planes that is flushed and rotated
//    once per coarse plane.
//  * ENO and linear-fallback operators (prolongate_3d_rf2_impl.hxx
//    :1417-1667, :1836-2148): each fine point's value is the tensor stencil
//    of one branch (an ENO shift triple, and the high-order or the linear
//    stencil), chosen from the data and constant between decision surfaces,
//    so the Jacobian at crse_in is that branch's stencil.  One pass records
//    every fine point's branch (the forward's own decision code); a second,
//    one thread per coarse point, gathers over the fine points whose
//    stencils can reach it.
//  * MINMOD operators: the Jacobian is a chain of per-pass branches
//    (minmod_adjoint.hpp); one pass records each MINMOD application's branch
//    by columns, a second gathers (tiled for CCC, per coarse point else).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "../../include/cxp.h"
#include "adj_zfirst.hpp"
#include "kernels.hpp"
#include "minmod_launch.hpp"
#include "ops.hpp"

namespace cxp {

// The weight tables of the linear operators (kAdjWc) and the z-first
// adjoint routines are in adj_zfirst.hpp.

// --------------------------------------------------------------------------
// Linear adjoint
//
// Workgroup = 256 threads owning a 16 x 16 coarse (x, y) output tile and a
// chunk of coarse z-planes.  It marches through the fine z-planes that feed
// the chunk.  Per fine plane:
//   fine window (2(16+span_x) x 2(16+span_y)) -> LDS, split by x parity,
//       loaded two fine planes ahead into registers with raw buffer loads
//       (points outside the fine region read as 0, so no branches);
//   x^T: fine x -> coarse x for every window row      -> LDS
//   y^T: fine y -> coarse y, one output per thread   -> register
//   z^T: accumulate into a register window of coarse planes, one plane
//        completed (and added to crse_bar) per pair of fine planes.
// Each transpose is compiled per stencil with its nonzero weights only.

template <int S> constexpr int adj_span() { return sid_whi(S) - sid_wlo(S); }

// sum_q w[S][OFF][q] * src[(b - q) * stride], zero weights skipped
template <int S, int OFF, int Q> struct AdjTaps {
  __device__ static __forceinline__ void add(double &s, const double *src,
                                             int b, int stride) {
    constexpr double w = kAdjWc.w[S][OFF][Q];
    if constexpr (w != 0.0)
      s += w * src[(b - Q) * stride];
    AdjTaps<S, OFF, Q + 1>::add(s, src, b, stride);
  }
};
template <int S, int OFF> struct AdjTaps<S, OFF, 9> {
  __device__ static __forceinline__ void add(double &, const double *, int,
                                             int) {}
};

// x^T of one window row: Fe/Fo hold the row's even/odd fine points; output
// coarse local cx = sum_{off,q} w[off][q] * F_off[cx + span - q].
template <int S>
__device__ __forceinline__ double adj_x(const double *Fe, const double *Fo,
                                        int cx) {
  constexpr int span = adj_span<S>();
  double s = 0.0;
  AdjTaps<S, 0, 0>::add(s, Fe, cx + span, 1);
  AdjTaps<S, 1, 0>::add(s, Fo, cx + span, 1);
  return s;
}

// y^T: Q rows are fine window rows (stride 16); row of (cy, q, off) is
// 2 (cy + span - q) + off.
template <int S>
__device__ __forceinline__ double adj_y(const double *Q, int cx, int cy) {
  constexpr int span = adj_span<S>();
  double s = 0.0;
  AdjTaps<S, 0, 0>::add(s, Q + cx, cy + span, 32);
  AdjTaps<S, 1, 0>::add(s, Q + cx + 16, cy + span, 32);
  return s;
}

template <int S, int OFF, int Q> struct AdjZ {
  __device__ static __forceinline__ void add(double (&acc)[9], double v) {
    constexpr double w = kAdjWc.w[S][OFF][Q];
    if constexpr (w != 0.0)
      acc[Q] += w * v;
    AdjZ<S, OFF, Q + 1>::add(acc, v);
  }
};
template <int S, int OFF> struct AdjZ<S, OFF, 9> {
  __device__ static __forceinline__ void add(double (&)[9], double) {}
};

#define CXP_ADJ_LINEAR_SIDS(X)                                                 \
  X(SID_VPOLY1) X(SID_VPOLY3) X(SID_VPOLY5) X(SID_VPOLY7) X(SID_CPOLY0)        \
  X(SID_CPOLY1) X(SID_CPOLY2) X(SID_CPOLY3) X(SID_CPOLY4) X(SID_CPOLY5)        \
  X(SID_CPOLY6) X(SID_CPOLY7) X(SID_VCONS0) X(SID_VCONS1) X(SID_CCONS0)        \
  X(SID_CCONS1) X(SID_CCONS2) X(SID_CCONS3) X(SID_CCONS4) X(SID_CCONS5)        \
  X(SID_CCONS6) X(SID_CCONS7)

struct AdjTask {
  DFab fbar;  // fine adjoint (input)
  DFab cbar;  // coarse adjoint (accumulated)
  int rlo[3], rhi[3];
  int sid[3];
  int wlo[3], whi[3];
  int olo[3], ohi[3];  // coarse output range
  long long fcs, ccs;
  int fbytes, cbytes;  // bytes of one component of each FAB
};

#define CXP_ADJ_T 16                 // coarse tile edge (x and y)
#define CXP_ADJ_WMAX 48              // 2 (16 + max span 8)
#define CXP_ADJ_R 9                  // ceil(48 * 48 / 256) window loads
#define CXP_ADJ_HALF (CXP_ADJ_WMAX / 2)

// SU >= 0: every direction uses stencil SU (VVV / CCC operators), compiled
// without runtime dispatch; SU < 0: per-direction runtime dispatch.
template <int SU>
__global__ void __launch_bounds__(256)
    k_adj_linear(AdjTask t, int ntx, int nty, int kchunk) {
  __shared__ double Fs[2 * CXP_ADJ_WMAX * CXP_ADJ_HALF]; // [parity][row][x/2]
  __shared__ double Qs[CXP_ADJ_WMAX * CXP_ADJ_T];        // [row][cx]
  const int tid = threadIdx.x;
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int tx = b % ntx, ty = (b / ntx) % nty, tz = b / (ntx * nty);
  const long long comp = blockIdx.y;
  const int ci0 = t.olo[0] + tx * CXP_ADJ_T;
  const int cj0 = t.olo[1] + ty * CXP_ADJ_T;
  const int kca = t.olo[2] + tz * kchunk;
  const int kcb = min(t.ohi[2], kca + kchunk - 1);
  const int spx = t.whi[0] - t.wlo[0], spy = t.whi[1] - t.wlo[1];
  const int W = 2 * (CXP_ADJ_T + spx), H = 2 * (CXP_ADJ_T + spy);
  const int hw = W >> 1;
  const int fxa = 2 * (ci0 - t.whi[0]);
  const int fya = 2 * (cj0 - t.whi[1]);
  const __amdgpu_buffer_rsrc_t rf =
      make_rsrc(t.fbar.p + comp * t.fcs, t.fbytes);
  const __amdgpu_buffer_rsrc_t rc =
      make_rsrc(t.cbar.p + comp * t.ccs, t.cbytes);

  // Per-thread window slots: byte offset inside a fine plane (or -1 when the
  // point lies outside the window or the fine region) and LDS slot.
  int soff[CXP_ADJ_R], lslot[CXP_ADJ_R];
#pragma unroll
  for (int r = 0; r < CXP_ADJ_R; ++r) {
    const int idx = tid + 256 * r;
    const int fy = idx / W;
    const int fx = idx - fy * W;
    const int gx = fxa + fx, gy = fya + fy;
    const bool in = idx < W * H && gx >= t.rlo[0] && gx <= t.rhi[0] &&
                    gy >= t.rlo[1] && gy <= t.rhi[1];
    soff[r] = in ? 8 * (int)((gx - t.fbar.lo[0]) +
                             (long long)(gy - t.fbar.lo[1]) * t.fbar.sy)
                 : -1;
    lslot[r] = idx < W * H
                   ? (fx & 1) * (CXP_ADJ_WMAX * CXP_ADJ_HALF) + fy * hw + (fx >> 1)
                   : -1;
  }
  const int cx = tid & 15, cy = tid >> 4;
  const int sidx = t.sid[0], sidy = t.sid[1], sidz = t.sid[2];

  const int mA = kca - t.whi[2], mB = kcb - t.wlo[2];
  // fine plane fz -> byte base, or -1 outside the region
  auto plane = [&](int fz) -> int {
    return (fz >= t.rlo[2] && fz <= t.rhi[2])
               ? 8 * (int)((long long)(fz - t.fbar.lo[2]) * t.fbar.sz)
               : -1;
  };
  // pre[0]: even fine planes 2m, pre[1]: odd ones 2m+1, each loaded one
  // coarse step (two fine planes) ahead
  double pre[2][CXP_ADJ_R];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int pb = plane(2 * mA + h);
#pragma unroll
    for (int r = 0; r < CXP_ADJ_R; ++r)
      pre[h][r] = bld(rf, (pb >= 0 && soff[r] >= 0) ? pb + soff[r] : CXP_OOB);
  }
  double acc[9];
#pragma unroll
  for (int q = 0; q < 9; ++q)
    acc[q] = 0.0;

  const int ci = ci0 + cx, cj = cj0 + cy;
  for (int m = mA; m <= mB; ++m) {
    // coarse plane m + wlo is completed by this step; its crse_bar values
    // are loaded first, so waiting for them later does not wait for the
    // prefetches issued below (the vector memory counter is in order)
    const int kc = m + t.wlo[2];
    const bool own = kc >= kca && ci <= t.ohi[0] && cj <= t.ohi[1];
    const int co = own ? 8 * (int)didx(t.cbar, ci, cj, kc) : CXP_OOB;
    const double cold = bld(rc, co);
    auto fine_plane = [&](auto offc) {
      constexpr int off = decltype(offc)::value;
      // this plane's window -> LDS
#pragma unroll
      for (int r = 0; r < CXP_ADJ_R; ++r)
        if (lslot[r] >= 0)
          Fs[lslot[r]] = pre[off][r];
      // the plane of this parity one coarse step ahead -> registers
      {
        const int nz = 2 * (m + 1) + off;
        const int pb = nz <= 2 * mB + 1 ? plane(nz) : -1;
#pragma unroll
        for (int r = 0; r < CXP_ADJ_R; ++r)
          pre[off][r] =
              bld(rf, (pb >= 0 && soff[r] >= 0) ? pb + soff[r] : CXP_OOB);
      }
      __syncthreads();
      // x^T over every window row
      auto xrows = [&](auto sc) {
        constexpr int S = decltype(sc)::value;
        for (int idx = tid; idx < H * CXP_ADJ_T; idx += 256) {
          const int row = idx >> 4, qx = idx & 15;
          const double *Fe = Fs + row * hw;
          Qs[idx] = adj_x<S>(Fe, Fe + CXP_ADJ_WMAX * CXP_ADJ_HALF, qx);
        }
      };
      if constexpr (SU >= 0) {
        xrows(std::integral_constant<int, SU>{});
      } else {
        switch (sidx) {
#define CXP_AX(S)                                                              \
  case S:                                                                      \
    xrows(std::integral_constant<int, S>{});                                   \
    break;
          CXP_ADJ_LINEAR_SIDS(CXP_AX)
#undef CXP_AX
        }
      }
      __syncthreads();
      double s = 0.0;
      if constexpr (SU >= 0) {
        s = adj_y<SU>(Qs, cx, cy);
        AdjZ<SU, off, 0>::add(acc, s);
      } else {
        switch (sidy) {
#define CXP_AY(S)                                                              \
  case S:                                                                      \
    s = adj_y<S>(Qs, cx, cy);                                                  \
    break;
          CXP_ADJ_LINEAR_SIDS(CXP_AY)
#undef CXP_AY
        }
        switch (sidz) {
#define CXP_AZ(S)                                                              \
  case S:                                                                      \
    AdjZ<S, off, 0>::add(acc, s);                                              \
    break;
          CXP_ADJ_LINEAR_SIDS(CXP_AZ)
#undef CXP_AZ
        }
      }
    };
    fine_plane(std::integral_constant<int, 0>{});
    fine_plane(std::integral_constant<int, 1>{});
    // coarse plane m + wlo is complete
    bst(rc, co, cold + acc[0]);
#pragma unroll
    for (int q = 0; q < 8; ++q)
      acc[q] = acc[q + 1];
    acc[8] = 0.0;
  }
}

// Uniform-stencil linear adjoint, z-transpose first (routines and their
// description in adj_zfirst.hpp).  Fine loads are raw buffer loads one
// parent plane ahead (points outside the fine region read as 0); the old
// crse_bar value of a complete plane is loaded before that plane's LDS
// passes.
// Tile b of task t (ntx x nty column tiles, z-chunks of kchunk planes);
// Zs, Ys: the workgroup's LDS buffers (AzGeom<S>::NZ and ::NY doubles).
template <int S>
__device__ __forceinline__ void adj_zf_tile(const AdjTask &t, int b, int ntx,
                                            int nty, int kchunk, double *Zs,
                                            double *Ys) {
  using G = AzGeom<S>;
  const int tid = threadIdx.x;
  const long long comp = blockIdx.y;
  int i0, j0, kca;
  az_tile(t.olo, b, ntx, nty, kchunk, i0, j0, kca);
  const int kcb = min(t.ohi[2], kca + kchunk - 1);
  const __amdgpu_buffer_rsrc_t rf =
      make_rsrc(t.fbar.p + comp * t.fcs, t.fbytes);
  const __amdgpu_buffer_rsrc_t rc =
      make_rsrc(t.cbar.p + comp * t.ccs, t.cbytes);
  const int fxa = 2 * (i0 - G::whi), fya = 2 * (j0 - G::whi);
  // owned columns: byte offset inside a fine plane, -1 outside the window
  // or the fine region
  int soff[G::R];
#pragma unroll
  for (int r = 0; r < G::R; ++r) {
    const int c = tid + AZ_NT * r;
    const int fy = c / G::W, fx = c - fy * G::W;
    const int gx = fxa + fx, gy = fya + fy;
    const bool in = c < G::NZ && gx >= t.rlo[0] && gx <= t.rhi[0] &&
                    gy >= t.rlo[1] && gy <= t.rhi[1];
    soff[r] = in ? 8 * (int)((gx - t.fbar.lo[0]) +
                             (long long)(gy - t.fbar.lo[1]) * t.fbar.sy)
                 : -1;
  }
  auto plane = [&](int fz) -> int {
    return (fz >= t.rlo[2] && fz <= t.rhi[2])
               ? 8 * (int)((long long)(fz - t.fbar.lo[2]) * t.fbar.sz)
               : -1;
  };
  const int pA = kca - G::whi, pB = kcb - G::wlo;
  double f[2][G::R];
  auto load = [&](int pz) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int pb = plane(2 * pz + h);
#pragma unroll
      for (int r = 0; r < G::R; ++r)
        f[h][r] = bld(rf, (pb >= 0 && soff[r] >= 0) ? pb + soff[r] : CXP_OOB);
    }
  };
  load(pA);
  double z[G::R][G::span + 1];
#pragma unroll
  for (int r = 0; r < G::R; ++r)
#pragma unroll
    for (int q = 0; q <= G::span; ++q)
      z[r][q] = 0.0;
  const int i = tid % AZ_T, j = tid / AZ_T;
  const bool own = i0 + i <= t.ohi[0] && j0 + j <= t.ohi[1];
  for (int pz = pA; pz <= pB; ++pz) {
    double g[2][G::R];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int r = 0; r < G::R; ++r)
        g[h][r] = f[h][r];
    if (pz < pB)
      load(pz + 1);
#pragma unroll
    for (int r = 0; r < G::R; ++r)
      az_zstep<S>(z[r], g[0][r], g[1][r]);
    const int k = pz + G::wlo;  // complete now
    if (k >= kca) {             // uniform over the workgroup
      const int co = own ? 8 * (int)didx(t.cbar, i0 + i, j0 + j, k) : CXP_OOB;
      const double cold = bld(rc, co);
#pragma unroll
      for (int r = 0; r < G::R; ++r) {
        const int c = tid + AZ_NT * r;
        if (c < G::NZ)
          Zs[c] = z[r][0];
      }
      __syncthreads();
      for (int e = tid; e < G::NY; e += AZ_NT)
        Ys[e] = az_y<S>(e, [&](int o) { return Zs[o]; });
      __syncthreads();
      // Zs is rewritten only after the next plane's z^T, Ys after the next
      // barrier: every read of both precedes the barrier above or this one
      const double x = az_x<S>(i, j, [&](int o) { return Ys[o]; });
      bst(rc, co, cold + x);
    }
#pragma unroll
    for (int r = 0; r < G::R; ++r) {
#pragma unroll
      for (int q = 0; q < G::span; ++q)
        z[r][q] = z[r][q + 1];
      z[r][G::span] = 0.0;
    }
  }
}

template <int S>
__global__ void __launch_bounds__(AZ_NT)
    k_adj_zf(AdjTask t, int ntx, int nty, int kchunk) {
  __shared__ double Zs[AzGeom<S>::NZ];
  __shared__ double Ys[AzGeom<S>::NY];
  adj_zf_tile<S>(t, xcd_block(blockIdx.x, gridDim.x), ntx, nty, kchunk, Zs,
                 Ys);
}

// Several regions' adjoints in one launch, region i owning blocks
// [start[i], start[i + 1]).  Their coarse output windows must be disjoint:
// every output point then takes the additions of one region only.
#define CXP_ADJ_MAX_REGIONS 8
struct AdjRegions {
  AdjTask t[CXP_ADJ_MAX_REGIONS];
  int start[CXP_ADJ_MAX_REGIONS + 1];
  int ntx[CXP_ADJ_MAX_REGIONS], nty[CXP_ADJ_MAX_REGIONS];
  int kchunk[CXP_ADJ_MAX_REGIONS];
  int n;
};

template <int S>
__global__ void __launch_bounds__(AZ_NT) k_adj_zf_regions(AdjRegions r) {
  __shared__ double Zs[AzGeom<S>::NZ];
  __shared__ double Ys[AzGeom<S>::NY];
  const int b = xcd_block(blockIdx.x, gridDim.x);
  int i = 0;
  while (i + 1 < r.n && b >= r.start[i + 1])
    ++i;
  adj_zf_tile<S>(r.t[i], b - r.start[i], r.ntx[i], r.nty[i], r.kchunk[i], Zs,
                 Ys);
}

// --------------------------------------------------------------------------
// ENO / linear-fallback adjoint

// Records the branch of every fine point of the region (per component in
// per-var mode, shared in per-group mode).
__global__ void __launch_bounds__(256)
    k_adj_decide(PointTask t, unsigned char *__restrict__ codes,
                 int per_group) {
  const long long npts = pt_npts(t);
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= npts)
    return;
  int f[3], ic[3], off[3], shift[3];
  pt_point(t, gid, f, ic, off);
  bool need;
  if (per_group) {
    double vals[16];
    pt_decide_group(t, off, ic, shift, vals, need);
    codes[gid] = adj_code(shift, need);
  } else {
    const long long comp = blockIdx.y;
    double val;
    pt_decide_var(t, comp, off, ic, shift, val, need);
    codes[comp * npts + gid] = adj_code(shift, need);
  }
}

struct AdjPointTask {
  PointTask pt;        // region and stencils (crse = crse_in)
  DFab cbar;           // coarse adjoint (accumulated), first component
  DFab fbar;           // fine adjoint (input), first component
  long long ccs, fcs;  // component strides of cbar / fbar
  int sl[3];           // stencils of the fallback branch
  int olo[3], ohi[3];  // coarse output range
  int rlo_reach[3], rhi_reach[3];  // coarse offsets any branch reaches
  int per_group;
};

// w[d][branch][shift+1][off][o+4]
#define CXP_ADJ_PT_W (3 * 2 * 3 * 2 * 9)

// One thread per coarse output point and component: sum over the fine
// points whose recorded branch reads that coarse point.
__global__ void __launch_bounds__(256)
    k_adj_gather(AdjPointTask t, const unsigned char *__restrict__ codes) {
  __shared__ double w[CXP_ADJ_PT_W];
  for (int e = threadIdx.x; e < CXP_ADJ_PT_W; e += blockDim.x) {
    const int o = e % 9 - 4;
    const int off = (e / 9) % 2;
    const int sh = (e / 18) % 3 - 1;
    const int br = (e / 54) % 2;
    const int d = e / 108;
    const int sid = br ? t.sl[d] : t.pt.sid[d];
    const int shift = (sid == SID_ENO2 && !br) ? sh : 0;
    int lo, hi;
    stencil_radius(sid, shift, off, lo, hi);
    w[e] = (o >= lo && o <= hi && !(br && sh != 0))
               ? apply1d(sid, DeltaAcc{o}, shift, off)
               : 0.0;
  }
  __syncthreads();
  const int nx = t.ohi[0] - t.olo[0] + 1, ny = t.ohi[1] - t.olo[1] + 1;
  const long long nout = (long long)nx * ny * (t.ohi[2] - t.olo[2] + 1);
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nout)
    return;
  const long long comp = blockIdx.y;
  const int c[3] = {t.olo[0] + (int)(gid % nx),
                    t.olo[1] + (int)((gid / nx) % ny),
                    t.olo[2] + (int)(gid / ((long long)nx * ny))};
  // fine points whose stencil may read c: coarse index ic with
  // c - ic in [lo_reach, hi_reach]
  int flo[3], fhi[3];
  for (int d = 0; d < 3; ++d) {
    flo[d] = max(t.pt.rlo[d], 2 * (c[d] - t.rhi_reach[d]));
    fhi[d] = min(t.pt.rhi[d], 2 * (c[d] - t.rlo_reach[d]) + 1);
  }
  const long long npts = pt_npts(t.pt);
  const int rnx = t.pt.rhi[0] - t.pt.rlo[0] + 1;
  const int rny = t.pt.rhi[1] - t.pt.rlo[1] + 1;
  const unsigned char *cd = codes + (t.per_group ? 0 : comp * npts);
  DFab fb = t.fbar;
  fb.p += comp * t.fcs;
  double sum = 0.0;
  for (int fz = flo[2]; fz <= fhi[2]; ++fz)
    for (int fy = flo[1]; fy <= fhi[1]; ++fy)
      for (int fx = flo[0]; fx <= fhi[0]; ++fx) {
        const int code = cd[(fx - t.pt.rlo[0]) +
                            (long long)(fy - t.pt.rlo[1]) * rnx +
                            (long long)(fz - t.pt.rlo[2]) * rnx * rny];
        const int br = code >= 27 ? 1 : 0;
        const int sh[3] = {code % 3, (code / 3) % 3, (code / 9) % 3};
        const int fv[3] = {fx, fy, fz};
        double wt = 1.0;
        for (int d = 0; d < 3; ++d) {
          const int o = c[d] - (fv[d] >> 1);
          wt *= w[(((d * 2 + br) * 3 + (br ? 1 : sh[d])) * 2 + (fv[d] & 1)) *
                      9 +
                  o + 4];
        }
        if (wt != 0.0)
          sum += wt * gld(fb.p, didx(fb, fx, fy, fz));
      }
  DFab cb = t.cbar;
  cb.p += comp * t.ccs;
  const long long ci = didx(cb, c[0], c[1], c[2]);
  gst(cb.p, ci, gld(cb.p, ci) + sum);
}

// The same sums, a coarse tile per workgroup.  The branch codes of the fine
// points any of the tile's outputs can read are staged in LDS, and the
// weights come from a table indexed by branch code: T[code][d][off][o + 4]
// is the 1-D weight in direction d of a fine point of that code and fine
// offset off for the coarse point o away from its parent (code
// CXP_ADJ_NOCODE, outside the region, weighs 0).  Candidates are visited in
// the same order and weights multiplied in the same order as in
// k_adj_gather, so the sums are the same to the bit.
#define CXP_ADJ_TX 8
#define CXP_ADJ_TY 8
#define CXP_ADJ_TZ 4
#define CXP_ADJ_NOCODE 54
#define CXP_ADJ_TROW 54  // 3 directions x 2 offsets x 9 offsets o

// Fine window of a coarse tile [c0, c1] in direction d.
__host__ __device__ inline int adj_window(int c0, int c1, int rlo_reach,
                                          int rhi_reach, int flo, int fhi,
                                          int &w0, int &w1) {
  w0 = max(flo, 2 * (c0 - rhi_reach));
  w1 = min(fhi, 2 * (c1 - rlo_reach) + 1);
  return w1 - w0 + 1;
}

__global__ void __launch_bounds__(256)
    k_adj_gather_tiled(AdjPointTask t, const unsigned char *__restrict__ codes,
                       int ntx, int nty) {
  extern __shared__ double gsm[];
  double *T = gsm;
  unsigned char *cs = (unsigned char *)(gsm + (CXP_ADJ_NOCODE + 1) * CXP_ADJ_TROW);
  for (int e = threadIdx.x; e < (CXP_ADJ_NOCODE + 1) * CXP_ADJ_TROW;
       e += blockDim.x) {
    const int code = e / CXP_ADJ_TROW, r = e % CXP_ADJ_TROW;
    const int d = r / 18, off = (r / 9) % 2, o = r % 9 - 4;
    double v = 0.0;
    if (code < CXP_ADJ_NOCODE) {
      const int br = code >= 27 ? 1 : 0;
      const int pw = d == 0 ? 1 : d == 1 ? 3 : 9;
      const int sh = ((code % 27) / pw) % 3 - 1;
      const int sid = br ? t.sl[d] : t.pt.sid[d];
      const int shift = (sid == SID_ENO2 && !br) ? sh : 0;
      int lo, hi;
      stencil_radius(sid, shift, off, lo, hi);
      if (o >= lo && o <= hi)
        v = apply1d(sid, DeltaAcc{o}, shift, off);
    }
    T[e] = v;
  }
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int tv[3] = {b % ntx, (b / ntx) % nty, b / (ntx * nty)};
  const int tsz[3] = {CXP_ADJ_TX, CXP_ADJ_TY, CXP_ADJ_TZ};
  int c0[3], c1[3], w0[3], w1[3], wn[3];
  for (int d = 0; d < 3; ++d) {
    c0[d] = t.olo[d] + tsz[d] * tv[d];
    c1[d] = min(c0[d] + tsz[d] - 1, t.ohi[d]);
    wn[d] = adj_window(c0[d], c1[d], t.rlo_reach[d], t.rhi_reach[d],
                       t.pt.rlo[d], t.pt.rhi[d], w0[d], w1[d]);
  }
  const long long comp = blockIdx.y;
  const long long npts = pt_npts(t.pt);
  const int rnx = t.pt.rhi[0] - t.pt.rlo[0] + 1;
  const int rny = t.pt.rhi[1] - t.pt.rlo[1] + 1;
  const unsigned char *cd = codes + (t.per_group ? 0 : comp * npts);
  const bool any = wn[0] > 0 && wn[1] > 0 && wn[2] > 0;
  const int wvol = any ? wn[0] * wn[1] * wn[2] : 0;
  for (int e = threadIdx.x; e < wvol; e += blockDim.x) {
    const int x = e % wn[0], y = (e / wn[0]) % wn[1], z = e / (wn[0] * wn[1]);
    cs[e] = cd[(w0[0] + x - t.pt.rlo[0]) +
               (long long)(w0[1] + y - t.pt.rlo[1]) * rnx +
               (long long)(w0[2] + z - t.pt.rlo[2]) * rnx * rny];
  }
  __syncthreads();
  const int c[3] = {c0[0] + (int)threadIdx.x % CXP_ADJ_TX,
                    c0[1] + ((int)threadIdx.x / CXP_ADJ_TX) % CXP_ADJ_TY,
                    c0[2] + (int)threadIdx.x / (CXP_ADJ_TX * CXP_ADJ_TY)};
  if (c[0] > c1[0] || c[1] > c1[1] || c[2] > c1[2])
    return;
  int flo[3], fhi[3];
  for (int d = 0; d < 3; ++d) {
    flo[d] = max(w0[d], 2 * (c[d] - t.rhi_reach[d]));
    fhi[d] = min(w1[d], 2 * (c[d] - t.rlo_reach[d]) + 1);
  }
  DFab fb = t.fbar;
  fb.p += comp * t.fcs;
  double sum = 0.0;
  for (int fz = flo[2]; fz <= fhi[2]; ++fz) {
    const int rz = 36 + (fz & 1) * 9 + c[2] - (fz >> 1) + 4;
    for (int fy = flo[1]; fy <= fhi[1]; ++fy) {
      const int ry = 18 + (fy & 1) * 9 + c[1] - (fy >> 1) + 4;
      const unsigned char *crow =
          cs + ((fz - w0[2]) * wn[1] + (fy - w0[1])) * wn[0];
      for (int fx = flo[0]; fx <= fhi[0]; ++fx) {
        const double *tc = T + crow[fx - w0[0]] * CXP_ADJ_TROW;
        const double wt =
            tc[(fx & 1) * 9 + c[0] - (fx >> 1) + 4] * tc[ry] * tc[rz];
        if (wt != 0.0)
          sum += wt * gld(fb.p, didx(fb, fx, fy, fz));
      }
    }
  }
  DFab cb = t.cbar;
  cb.p += comp * t.ccs;
  const long long ci = didx(cb, c[0], c[1], c[2]);
  gst(cb.p, ci, gld(cb.p, ci) + sum);
}

inline DFab adj_dfab(const cxp_fab *f, int comp) {
  DFab v;
  v.p = f->data + (int64_t)comp * f->comp_stride;
  for (int d = 0; d < 3; ++d)
    v.lo[d] = f->lo[d];
  v.sy = f->hi[0] - f->lo[0] + 1;
  v.sz = (long long)v.sy * (f->hi[1] - f->lo[1] + 1);
  return v;
}

inline long long adj_component_bytes(const cxp_fab *f) {
  return 8LL * (f->hi[0] - f->lo[0] + 1) * (f->hi[1] - f->lo[1] + 1) *
         (f->hi[2] - f->lo[2] + 1);
}

// The linear adjoint's task for region [rlo, rhi]: fine_bar in, crse_bar
// accumulated over the coarse output window.
inline AdjTask adj_task(const OpDesc &op, const cxp_fab *crse_bar,
                        int crse_comp, const cxp_fab *fine_bar, int fine_comp,
                        const int rlo[3], const int rhi[3]) {
  AdjTask t;
  t.fbar = adj_dfab(fine_bar, fine_comp);
  t.cbar = adj_dfab(crse_bar, crse_comp);
  for (int d = 0; d < 3; ++d) {
    t.rlo[d] = rlo[d];
    t.rhi[d] = rhi[d];
    t.sid[d] = op.sid[d];
    t.wlo[d] = sid_wlo(op.sid[d]);
    t.whi[d] = sid_whi(op.sid[d]);
    t.olo[d] = (rlo[d] >> 1) + sid_wlo(op.sid[d]);
    t.ohi[d] = (rhi[d] >> 1) + sid_whi(op.sid[d]);
  }
  t.fcs = fine_bar->comp_stride;
  t.ccs = crse_bar->comp_stride;
  // component volumes; the caller has checked both stay below 2 GiB
  t.fbytes = (int)adj_component_bytes(fine_bar);
  t.cbytes = (int)adj_component_bytes(crse_bar);
  return t;
}

// The z-first kernel's grid for task t: column tiles and z-chunks.  Chunks
// of at least 8 planes when that still gives 512 workgroups (a chunk
// re-reads span_z parent planes; at 96^3 -> 192^3 VVV order 5, chunks of 8
// took 30 us, of 13 32 us, of 26 43 us, of 4 37 us); thinner outputs (the
// slabs of a ghost shell, a few tiles across) take chunks down to 2
// planes, which took the two-level wave's backward from 3.64 to 3.3 ms per
// coarse step.
inline void adj_zf_grid(const AdjTask &t, int ncomp, int &ntx, int &nty,
                        int &kz, int &nkc) {
  ntx = (t.ohi[0] - t.olo[0] + AZ_T) / AZ_T;
  nty = (t.ohi[1] - t.olo[1] + AZ_T) / AZ_T;
  const int nk = t.ohi[2] - t.olo[2] + 1;
  const long long cols = (long long)ntx * nty * ncomp;
  kz = (int)((nk * cols + 1023) / 1024);
  const int kmin = cols * ((nk + 7) / 8) >= 512 ? 8 : 2;
  kz = kz < kmin ? kmin : (kz > nk ? nk : kz);
  nkc = (nk + kz - 1) / kz;
}

inline int adjoint_launch(const OpDesc &op, const cxp_fab *crse_in,
                          const cxp_fab *crse_bar, int crse_comp,
                          const cxp_fab *fine_bar, int fine_comp, int ncomp,
                          const int rlo[3], const int rhi[3], int per_group,
                          hipStream_t s) {
  (void)per_group;
  // ENO and fallback operators are launched by adjoint_point_launch
  if (op_has_eno(op) || op_fallback_per_var(op) || op_fallback_per_group(op))
    return CXP_ERR_UNSUPPORTED;
  bool minmod = false;
  for (int d = 0; d < 3; ++d)
    minmod = minmod || op.sid[d] == SID_MINMOD;
  int olo[3], ohi[3];
  for (int d = 0; d < 3; ++d) {
    olo[d] = (rlo[d] >> 1) + sid_wlo(op.sid[d]);
    ohi[d] = (rhi[d] >> 1) + sid_whi(op.sid[d]);
  }
  if (!minmod) {
    const AdjTask t =
        adj_task(op, crse_bar, crse_comp, fine_bar, fine_comp, rlo, rhi);
    const int ntx = (ohi[0] - olo[0] + CXP_ADJ_T) / CXP_ADJ_T;
    const int nty = (ohi[1] - olo[1] + CXP_ADJ_T) / CXP_ADJ_T;
    const int nk = ohi[2] - olo[2] + 1;
    // z-chunks: about 1024 workgroups, but at least 8 coarse planes each
    // (a chunk re-reads span_z extra pairs of fine planes)
    const long long cols = (long long)ntx * nty * ncomp;
    int kchunk = (int)((nk * cols + 1023) / 1024);
    kchunk = kchunk < 8 ? 8 : kchunk;
    kchunk = kchunk > nk ? nk : kchunk;
    const int nkc = (nk + kchunk - 1) / kchunk;
    const dim3 grid(ntx * nty * nkc, ncomp);
    bool done = false;
    if (op.sid[0] == op.sid[1] && op.sid[1] == op.sid[2]) {
      // one stencil in every direction: the z-first kernel
      int ztx, zty, kz, zkc;
      adj_zf_grid(t, ncomp, ztx, zty, kz, zkc);
      const dim3 gz(ztx * zty * zkc, ncomp);
      done = true;
      switch (op.sid[0]) {
#define CXP_AL(S)                                                              \
  case S:                                                                      \
    hipLaunchKernelGGL(k_adj_zf<S>, gz, dim3(AZ_NT), 0, s, t, ztx, zty, kz);   \
    break;
        CXP_ADJ_LINEAR_SIDS(CXP_AL)
#undef CXP_AL
      default:
        done = false;
      }
    }
    if (!done)
      hipLaunchKernelGGL(k_adj_linear<-1>, grid, dim3(256), 0, s, t, ntx, nty,
                         kchunk);
    return hipGetLastError() == hipSuccess ? CXP_OK : CXP_ERR_HIP;
  }
  // MINMOD operators take the record-and-gather path (mm_*_launch), which
  // needs the forward input; the C ABI routes them there.
  (void)crse_in;
  return CXP_ERR_UNSUPPORTED;
}

// Linear adjoints of several regions (lo[3 i ..], hi[3 i ..]) into one
// crse_bar.  Regions run in order; a run of consecutive regions whose
// coarse output windows are pairwise disjoint shares one launch when the
// operator has one stencil in every direction (every output point then
// takes the same additions in the same order as region by region).
inline int adjoint_regions_launch(const OpDesc &op, const cxp_fab *crse_bar,
                                  int crse_comp, const cxp_fab *fine_bar,
                                  int fine_comp, int ncomp, int nregions,
                                  const int *lo, const int *hi, hipStream_t s) {
  bool uniform = op.sid[0] == op.sid[1] && op.sid[1] == op.sid[2];
  switch (op.sid[0]) {
#define CXP_AU(S)                                                              \
  case S:                                                                      \
    break;
    CXP_ADJ_LINEAR_SIDS(CXP_AU)
#undef CXP_AU
  default:
    uniform = false;
  }
  if (!uniform) {
    for (int i = 0; i < nregions; ++i)
      if (const int e = adjoint_launch(op, nullptr, crse_bar, crse_comp,
                                       fine_bar, fine_comp, ncomp, lo + 3 * i,
                                       hi + 3 * i, 0, s))
        return e;
    return CXP_OK;
  }
  auto meet = [&](const AdjTask &a, const AdjTask &b) {
    for (int d = 0; d < 3; ++d)
      if (a.ohi[d] < b.olo[d] || b.ohi[d] < a.olo[d])
        return false;
    return true;
  };
  int i = 0;
  while (i < nregions) {
    AdjRegions r{};
    r.start[0] = 0;
    r.n = 0;
    while (i < nregions && r.n < CXP_ADJ_MAX_REGIONS) {
      const AdjTask t = adj_task(op, crse_bar, crse_comp, fine_bar, fine_comp,
                                 lo + 3 * i, hi + 3 * i);
      bool apart = true;
      for (int k = 0; k < r.n; ++k)
        apart = apart && !meet(r.t[k], t);
      if (!apart)
        break;
      int nkc;
      r.t[r.n] = t;
      adj_zf_grid(t, ncomp, r.ntx[r.n], r.nty[r.n], r.kchunk[r.n], nkc);
      r.start[r.n + 1] = r.start[r.n] + r.ntx[r.n] * r.nty[r.n] * nkc;
      ++r.n;
      ++i;
    }
    const dim3 g((unsigned)r.start[r.n], ncomp);
    switch (op.sid[0]) {
#define CXP_AR(S)                                                              \
  case S:                                                                      \
    hipLaunchKernelGGL(k_adj_zf_regions<S>, g, dim3(AZ_NT), 0, s, r);          \
    break;
      CXP_ADJ_LINEAR_SIDS(CXP_AR)
#undef CXP_AR
    }
    if (hipGetLastError() != hipSuccess)
      return CXP_ERR_HIP;
  }
  return CXP_OK;
}

} // namespace cxp
