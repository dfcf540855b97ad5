// Adjoint of a linear rf=2 prolongation whose three directions share one
// stencil S (every VVV or CCC POLY / CONS operator), z-transpose first:
// crse_bar += Px^T Py^T Pz^T fine_bar.
//
// A workgroup of AZ_NT threads owns an AZ_T x AZ_T coarse (x, y) output tile
// and a chunk of coarse planes [kca, kcb].  Its fine window is the
// W x W = 2 (AZ_T + span)^2 fine columns whose x and y stencils reach the
// tile; thread t owns columns t + AZ_NT r.  It marches the parent planes
// pz = kca - whi .. kcb - wlo:
//   z^T  each owned column adds w[0][q] f(2 pz) + w[1][q] f(2 pz + 1) to a
//        register window of span + 1 coarse planes (window slot q is plane
//        pz + wlo + q); plane pz + wlo is then complete;
// and, for a complete plane k of the chunk,
//   its column values -> LDS Zs, y^T (fine rows -> coarse rows) -> LDS Ys,
//   x^T (fine x -> coarse x), one output per thread, added to crse_bar.
// So the two LDS passes and their two barriers run once per coarse plane,
// on data the z-transpose has already halved.  The weights are the
// forward's own (the weight of coarse offset q + wlo for fine offset off in
// Op1<S>::apply), compile-time constants with the zero ones skipped.
//
// The routines are host-and-device functions, so a host build runs the
// workgroups with checked accesses (carpetx_amd/host/adj_zfirst_host.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include "stencil1d.hpp"

namespace cxp {

// ---- weight tables: w[sid][off][q] is the weight of coarse offset
// q + wlo(sid) in the forward formula for fine offset off

struct DeltaAcc {
  int o;
  __host__ __device__ constexpr double operator()(int k) const {
    return k == o ? 1.0 : 0.0;
  }
};

template <int S> constexpr double lin_weight(int off, int o) {
  return Op1<S>::apply(DeltaAcc{o}, off);
}

constexpr double lin_weight_rt(int sid, int off, int o) {
  switch (sid) {
#define CXP_LW(S)                                                              \
  case S:                                                                      \
    return lin_weight<S>(off, o);
    CXP_LW(SID_VPOLY1) CXP_LW(SID_VPOLY3) CXP_LW(SID_VPOLY5) CXP_LW(SID_VPOLY7)
    CXP_LW(SID_CPOLY0) CXP_LW(SID_CPOLY1) CXP_LW(SID_CPOLY2) CXP_LW(SID_CPOLY3)
    CXP_LW(SID_CPOLY4) CXP_LW(SID_CPOLY5) CXP_LW(SID_CPOLY6) CXP_LW(SID_CPOLY7)
    CXP_LW(SID_VCONS0) CXP_LW(SID_VCONS1) CXP_LW(SID_CCONS0) CXP_LW(SID_CCONS1)
    CXP_LW(SID_CCONS2) CXP_LW(SID_CCONS3) CXP_LW(SID_CCONS4) CXP_LW(SID_CCONS5)
    CXP_LW(SID_CCONS6) CXP_LW(SID_CCONS7)
#undef CXP_LW
  }
  return 0.0;
}

struct AdjW {
  double w[SID_COUNT][2][9];
};

constexpr AdjW make_adjw() {
  AdjW t{};
  for (int s = 0; s < SID_COUNT; ++s)
    for (int off = 0; off < 2; ++off)
      for (int q = 0; q < 9; ++q) {
        const SidInfo si = sid_info(s);
        const int o = q + sid_wlo(s);
        const bool in = o >= si.lo[off] && o <= si.hi[off];
        t.w[s][off][q] = in ? lin_weight_rt(s, off, o) : 0.0;
      }
  return t;
}

constexpr AdjW kAdjWc = make_adjw();

// ---- the z-first tile

#define AZ_T 16    // coarse output tile edge (x and y)
#define AZ_NT 256  // threads per workgroup: one output column each

template <int S> struct AzGeom {
  static constexpr int wlo = sid_wlo(S), whi = sid_whi(S);
  static constexpr int span = whi - wlo;
  static constexpr int W = 2 * (AZ_T + span);              // fine window edge
  static constexpr int R = (W * W + AZ_NT - 1) / AZ_NT;   // columns per thread
  static constexpr int NZ = W * W;                         // Zs doubles
  static constexpr int NY = AZ_T * W;                      // Ys doubles
};

// z^T of one fine column for parent plane pz: f0, f1 are its fine adjoints
// at 2 pz and 2 pz + 1; z[q] accumulates coarse plane pz + wlo + q.
template <int S, int Q = 0>
__host__ __device__ __forceinline__ void az_zstep(double (&z)[AzGeom<S>::span + 1],
                                                  double f0, double f1) {
  if constexpr (Q <= AzGeom<S>::span) {
    constexpr double w0 = kAdjWc.w[S][0][Q], w1 = kAdjWc.w[S][1][Q];
    if constexpr (w0 != 0.0)
      z[Q] += w0 * f0;
    if constexpr (w1 != 0.0)
      z[Q] += w1 * f1;
    az_zstep<S, Q + 1>(z, f0, f1);
  }
}

// sum over off, q of w[off][q] src((2 (b + span - q) + off) * stride), off
// outer, zero weights skipped
template <int S, int OFF, int Q, class SRC>
__host__ __device__ __forceinline__ void az_taps(double &s, const SRC &src,
                                                 int b, int stride) {
  if constexpr (OFF < 2) {
    if constexpr (Q <= AzGeom<S>::span) {
      constexpr double w = kAdjWc.w[S][OFF][Q];
      if constexpr (w != 0.0)
        s += w * src((2 * (b + AzGeom<S>::span - Q) + OFF) * stride);
      az_taps<S, OFF, Q + 1>(s, src, b, stride);
    } else {
      az_taps<S, OFF + 1, 0>(s, src, b, stride);
    }
  }
}

// y^T entry e < NY of Ys (coarse row j = e / W, fine x = e % W) from the
// complete plane's column values zs(fy * W + fx).
template <int S, class ZS>
__host__ __device__ __forceinline__ double az_y(int e, const ZS &zs) {
  constexpr int W = AzGeom<S>::W;
  const int j = e / W, fx = e - j * W;
  double s = 0.0;
  az_taps<S, 0, 0>(s, [&](int o) { return zs(o + fx); }, j, W);
  return s;
}

// x^T output (i, j) of the tile from ys(j * W + fx).
template <int S, class YS>
__host__ __device__ __forceinline__ double az_x(int i, int j, const YS &ys) {
  constexpr int W = AzGeom<S>::W;
  double s = 0.0;
  az_taps<S, 0, 0>(s, [&](int o) { return ys(j * W + o); }, i, 1);
  return s;
}

// Tile b of ntx x nty x (z-chunks of kchunk planes) over the output window
// starting at olo: origin (i0, j0) and chunk start kca.
__host__ __device__ inline void az_tile(const int olo[3], int b, int ntx,
                                        int nty, int kchunk, int &i0, int &j0,
                                        int &kca) {
  i0 = olo[0] + AZ_T * (b % ntx);
  j0 = olo[1] + AZ_T * ((b / ntx) % nty);
  kca = olo[2] + kchunk * (b / (ntx * nty));
}

} // namespace cxp
