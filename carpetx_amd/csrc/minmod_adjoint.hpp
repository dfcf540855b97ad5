// Adjoint of the MINMOD operators (CC MINMOD 1, impl.hxx:846-878, in the
// cell-centred directions of the `minmod` family; vertex-centred
// directions are POLY p, linear).
//
// The forward evaluates, x innermost (impl.hxx:1206-1216),
//   X(fx, j, k)   = Sx_ox( c(px + ., j, k) )
//   XY(fx, fy, k) = Sy_oy( X(fx, py + ., k) )
//   f(fx, fy, fz) = Sz_oz( XY(fx, fy, pz + .) )
// with p = f >> 1 and o = f & 1 per direction.  A MINMOD application
// c + s delta (s = -1 at o = 0, +1 at o = 1) takes one of three branches,
//   delta = 0,  (c - c-) / 4  or  (c+ - c) / 4,
// and its branch depends only on the three values it reads, not on o.  So
// each MINMOD direction has one branch per point of the grid its pass reads
// along it:
//   Bx(px, j, k)   from c(px - 1 .. px + 1, j, k),
//   By(fx, py, k)  from X(fx, py - 1 .. py + 1, k),
//   Bz(fx, fy, pz) from XY(fx, fy, pz - 1 .. pz + 1).
// The record pass writes these codes (one byte each); the gather sums, for
// each coarse point (i, j, k), over the fine points F whose stencils reach
// it,
//   wz(oz, k - pz; Bz(F)) * wy(oy, j - py; By(fx, py, k))
//     * wx(ox, i - px; Bx(px, j, k)) * fbar(F),
// the Jacobian entry of f(F) by the chain rule: f(F) depends on plane k
// through XY(fx, fy, k) alone, which depends on row j through X(fx, j, k)
// alone.  Linear directions take their fixed weights.  No atomics.
//
// The kernels (k_mm_record_col, k_mm_gather_tiled for CCC operators and
// k_mm_gather otherwise, in adjoint.hpp) record by columns and gather a
// tile at a time, the transposes applied as z^T, y^T, x^T (below); the
// per-point routines (mm_record_one, mm_gather_one) state the same sums one
// entry at a time.  All are host-and-device functions over accessor
// functors, so a host build runs them with bounds-checked accessors
// (carpetx_amd/host/minmod_adjoint_host.cpp) and the kernels with plain
// ones.
#pragma once

#include <hip/hip_runtime.h>

#include "stencil1d.hpp"

namespace cxp {

enum : unsigned char { MM_ZERO = 0, MM_MINUS = 1, MM_PLUS = 2 };

// One component of a FAB in the level's index space (x fastest).
struct DFabC {
  double *p;
  int lo[3];
  long long sy, sz;
};
__host__ __device__ inline long long dcidx(const DFabC &f, int i, int j,
                                           int k) {
  return (long long)(i - f.lo[0]) + (long long)(j - f.lo[1]) * f.sy +
         (long long)(k - f.lo[2]) * f.sz;
}

// The branch of minmod on (c-, c, c+), as the forward's arithmetic takes it
// (stencil1d.hpp minmod_apply): no slope when the one-sided differences'
// signs differ, else the smaller difference (the minus side on a tie).
// ... from the one-sided differences dm = c - c-, dp = c+ - c
__host__ __device__ inline unsigned char mm_code_d(double dm, double dp) {
  const double k = copysign(0.125, dm) + copysign(0.125, dp);
  if (k == 0.0)
    return MM_ZERO;
  return fabs(dm) <= fabs(dp) ? MM_MINUS : MM_PLUS;
}
__host__ __device__ inline unsigned char mm_code(double cm, double c0,
                                                 double cp) {
  return mm_code_d(c0 - cm, cp - c0);
}

// d (c + s delta) / d c(a), a in {-1, 0, +1}, for branch code at offset off.
__host__ __device__ inline double mm_weight(unsigned char code, int off,
                                            int a) {
  const double s = off == 0 ? -0.25 : 0.25;  // s / 4
  if (code == MM_MINUS)
    return a == 0 ? 1.0 + s : (a == -1 ? -s : 0.0);
  if (code == MM_PLUS)
    return a == 0 ? 1.0 - s : (a == 1 ? s : 0.0);
  return a == 0 ? 1.0 : 0.0;
}

// Geometry of one MINMOD adjoint call (one component; the kernels offset
// the pointers per component).
struct MmGeom {
  int rlo[3], rhi[3];  // fine region
  int sid[3];
  int wlo[3], whi[3];  // stencil windows, sid_wlo / sid_whi
  int olo[3], ohi[3];  // coarse output window (r >> 1) + w
  // code arrays Bx, By, Bz: lower corner and extent (extent 0: the
  // direction is linear and has no codes), offset in the component's codes
  int blo[3][3], bn[3][3];
  long long boff[3];
  long long ncodes;  // codes per component
  // linear weights by direction, fine offset, coarse offset - wlo
  double w[3][2][9];
};

__host__ __device__ inline long long mm_code_index(const MmGeom &g, int which,
                                                   int x, int y, int z) {
  return g.boff[which] +
         ((long long)(z - g.blo[which][2]) * g.bn[which][1] +
          (y - g.blo[which][1])) *
             g.bn[which][0] +
         (x - g.blo[which][0]);
}

// Sets the code boxes and offsets of a geometry whose region, sids and
// windows are set (host).
inline void mm_layout(MmGeom &g) {
  long long off = 0;
  for (int w = 0; w < 3; ++w) {
    for (int d = 0; d < 3; ++d) {
      g.blo[w][d] = 0;
      g.bn[w][d] = 0;
    }
    g.boff[w] = off;
    if (g.sid[w] != SID_MINMOD)
      continue;
    for (int d = 0; d < 3; ++d) {
      int lo, hi;
      if (d < w) {  // fine indices (passes before w made them fine)
        lo = g.rlo[d];
        hi = g.rhi[d];
      } else if (d == w) {  // parents
        lo = g.rlo[d] >> 1;
        hi = g.rhi[d] >> 1;
      } else {  // coarse indices the later passes read
        lo = g.olo[d];
        hi = g.ohi[d];
      }
      g.blo[w][d] = lo;
      g.bn[w][d] = hi - lo + 1;
    }
    off += (long long)g.bn[w][0] * g.bn[w][1] * g.bn[w][2];
  }
  g.ncodes = off;
}

// The 1-D stencils of the `minmod` family (impl_minmod.cxx): MINMOD in
// cell-centred directions, POLY 1/3/5 in vertex-centred ones.  Dispatch
// over these four only keeps the nested passes small.
__host__ __device__ inline bool mm_sid_ok(int sid) {
  return sid == SID_MINMOD || sid == SID_VPOLY1 || sid == SID_VPOLY3 ||
         sid == SID_VPOLY5;
}
template <class F>
__host__ __device__ inline double mm_apply(int sid, const F &c, int off) {
  switch (sid) {
  case SID_MINMOD:
    return Op1<SID_MINMOD>::apply(c, off);
  case SID_VPOLY1:
    return Op1<SID_VPOLY1>::apply(c, off);
  case SID_VPOLY3:
    return Op1<SID_VPOLY3>::apply(c, off);
  case SID_VPOLY5:
    return Op1<SID_VPOLY5>::apply(c, off);
  }
  return 0.0;
}

// The geometry of a call over fine region [rlo, rhi] with stencils sid
// (each mm_sid_ok).
inline void mm_setup(MmGeom &g, const int sid[3], const int rlo[3],
                     const int rhi[3]) {
  for (int d = 0; d < 3; ++d) {
    g.rlo[d] = rlo[d];
    g.rhi[d] = rhi[d];
    g.sid[d] = sid[d];
    g.wlo[d] = sid_wlo(sid[d]);
    g.whi[d] = sid_whi(sid[d]);
    g.olo[d] = (rlo[d] >> 1) + g.wlo[d];
    g.ohi[d] = (rhi[d] >> 1) + g.whi[d];
    for (int off = 0; off < 2; ++off)
      for (int q = 0; q < 9; ++q) {
        const int o = q + g.wlo[d];
        g.w[d][off][q] =
            sid[d] == SID_MINMOD || o > g.whi[d]
                ? 0.0
                : mm_apply(sid[d], [o](int m) { return m == o ? 1.0 : 0.0; },
                           off);
      }
  }
  mm_layout(g);
}

// X(fx, j, k): the x-pass on coarse row (j, k) at fine x.
template <class C>
__host__ __device__ inline double mm_xval(const MmGeom &g, const C &c, int fx,
                                          int j, int k) {
  const int px = fx >> 1;
  return mm_apply(g.sid[0], [&](int o) { return c(px + o, j, k); }, fx & 1);
}

// XY(fx, fy, k): the y-pass over X at fine (x, y), coarse plane k.
template <class C>
__host__ __device__ inline double mm_xyval(const MmGeom &g, const C &c, int fx,
                                           int fy, int k) {
  const int py = fy >> 1;
  return mm_apply(g.sid[1], [&](int o) { return mm_xval(g, c, fx, py + o, k); },
                  fy & 1);
}

// Record: code number n (0 <= n < ncodes) from the forward input c.
template <class C>
__host__ __device__ inline unsigned char mm_record_one(const MmGeom &g,
                                                       const C &c,
                                                       long long n) {
  int w = 2;
  while (w > 0 && n < g.boff[w])
    --w;
  const long long r = n - g.boff[w];
  const int x = g.blo[w][0] + (int)(r % g.bn[w][0]);
  const int y = g.blo[w][1] + (int)((r / g.bn[w][0]) % g.bn[w][1]);
  const int z = g.blo[w][2] + (int)(r / ((long long)g.bn[w][0] * g.bn[w][1]));
  if (w == 0)
    return mm_code(c(x - 1, y, z), c(x, y, z), c(x + 1, y, z));
  if (w == 1)
    return mm_code(mm_xval(g, c, x, y - 1, z), mm_xval(g, c, x, y, z),
                   mm_xval(g, c, x, y + 1, z));
  return mm_code(mm_xyval(g, c, x, y, z - 1), mm_xyval(g, c, x, y, z),
                 mm_xyval(g, c, x, y, z + 1));
}

// Gather: (P^T fbar)(i, j, k) for a coarse point of the output window.
// codes(w, x, y, z) returns the code of array w (0: Bx, 1: By, 2: Bz) at
// (x, y, z), fbar(fx, fy, fz) the fine adjoint.
template <class B, class F>
__host__ __device__ inline double mm_gather_one(const MmGeom &g, const B &codes,
                                                const F &fbar, int i, int j,
                                                int k) {
  const int c3[3] = {i, j, k};
  int flo[3], fhi[3];
  for (int d = 0; d < 3; ++d) {
    // fine points whose parent p has c - p in [wlo, whi]
    flo[d] = 2 * (c3[d] - g.whi[d]);
    fhi[d] = 2 * (c3[d] - g.wlo[d]) + 1;
    flo[d] = flo[d] < g.rlo[d] ? g.rlo[d] : flo[d];
    fhi[d] = fhi[d] > g.rhi[d] ? g.rhi[d] : fhi[d];
  }
  const bool mx = g.sid[0] == SID_MINMOD, my = g.sid[1] == SID_MINMOD,
             mz = g.sid[2] == SID_MINMOD;
  double acc = 0.0;
  for (int fz = flo[2]; fz <= fhi[2]; ++fz) {
    const int pz = fz >> 1, oz = fz & 1, e = k - pz;
    const double wzl = mz ? 0.0 : g.w[2][oz][e - g.wlo[2]];
    if (!mz && wzl == 0.0)
      continue;
    for (int fy = flo[1]; fy <= fhi[1]; ++fy) {
      const int py = fy >> 1, oy = fy & 1, b = j - py;
      const double wyl = my ? 0.0 : g.w[1][oy][b - g.wlo[1]];
      if (!my && wyl == 0.0)
        continue;
      for (int fx = flo[0]; fx <= fhi[0]; ++fx) {
        const int px = fx >> 1, ox = fx & 1, a = i - px;
        const double wx =
            mx ? mm_weight(codes(0, px, j, k), ox, a)
               : g.w[0][ox][a - g.wlo[0]];
        if (wx == 0.0)
          continue;
        const double wy =
            my ? mm_weight(codes(1, fx, py, k), oy, b) : wyl;
        const double wz =
            mz ? mm_weight(codes(2, fx, fy, pz), oz, e)
               : wzl;
        acc += (wz * wy) * wx * fbar(fx, fy, fz);
      }
    }
  }
  return acc;
}


// ---------------------------------------------------------------------------
// Faster record: one thread per column of codes, marching in z.
//
// Columns: Bx(px, j, .), By(fx, py, .), Bz(fx, fy, .) in that order; a Bz
// column keeps XY of three consecutive parents, so each code costs one new
// XY (9 coarse loads) instead of three.  The values and branches are those
// of mm_record_one.

// Column pieces of array w: (x, y) columns times z-chunks of zc codes.
__host__ __device__ inline long long mm_columns(const MmGeom &g, int w,
                                                int zc) {
  return (long long)g.bn[w][0] * g.bn[w][1] * ((g.bn[w][2] + zc - 1) / zc);
}

// Piece n of all arrays (Bx pieces first, x fastest within a chunk);
// store(w, x, y, z, code) stores one code.
template <class C, class S>
__host__ __device__ inline void mm_record_column(const MmGeom &g, const C &c,
                                                 const S &store, long long n,
                                                 int zc) {
  int w = 0;
  while (w < 2 && n >= mm_columns(g, w, zc)) {
    n -= mm_columns(g, w, zc);
    ++w;
  }
  if (n >= mm_columns(g, w, zc))
    return;
  const long long cols = (long long)g.bn[w][0] * g.bn[w][1];
  const long long col = n % cols;
  const int ch = (int)(n / cols);
  const int x = g.blo[w][0] + (int)(col % g.bn[w][0]);
  const int y = g.blo[w][1] + (int)(col / g.bn[w][0]);
  const int z0 = g.blo[w][2] + ch * zc;
  const int zend = g.blo[w][2] + g.bn[w][2] - 1;
  const int z1 = z0 + zc - 1 < zend ? z0 + zc - 1 : zend;
  if (w == 0) {
    for (int z = z0; z <= z1; ++z)
      store(0, x, y, z, mm_code(c(x - 1, y, z), c(x, y, z), c(x + 1, y, z)));
  } else if (w == 1) {
    for (int z = z0; z <= z1; ++z)
      store(1, x, y, z,
            mm_code(mm_xval(g, c, x, y - 1, z), mm_xval(g, c, x, y, z),
                    mm_xval(g, c, x, y + 1, z)));
  } else {
    double a = mm_xyval(g, c, x, y, z0 - 1), b = mm_xyval(g, c, x, y, z0);
    for (int z = z0; z <= z1; ++z) {
      const double n1 = mm_xyval(g, c, x, y, z + 1);
      store(2, x, y, z, mm_code(a, b, n1));
      a = b;
      b = n1;
    }
  }
}

// ---------------------------------------------------------------------------
// Tiled adjoint of the operators MINMOD in all three directions (CCC).
//
// One MINMOD pass, f(2p + o) = c(p) + (2o - 1) delta(p) with delta(p) one of
// 0, (c(p) - c(p - 1)) / 4, (c(p + 1) - c(p)) / 4 by the branch at p, has
// for the fine adjoints g0 = fbar(2p), g1 = fbar(2p + 1) the transpose
//   d = (g1 - g0) / 4,  a = d on MINUS else 0,  b = d on PLUS else 0,
//   cbar(p - 1) -= a,  cbar(p) += ((g0 + g1) + a) - b,  cbar(p + 1) += b,
// so out(i) = (b(i - 1) + c(i)) - a(i + 1) with c = ((g0 + g1) + a) - b
// (mt_pair).  The tiles apply it as z^T, y^T, x^T:
//   z^T: parent pz adds -a, c, b to the XYbar planes pz - 1, pz, pz + 1,
//        kept per footprint position in registers (mt_zt); after pz, plane
//        pz - 1 is complete and goes to LDS (the order of the adds is that
//        of the out(i) formula);
//   y^T: Xbar(fx, j, k) from the XYbar rows of parents j - 1 .. j + 1;
//   x^T: crse_bar(i, j, k) += from the Xbar columns of parents i - 1 .. i + 1.
// Fine points outside the region have a zero adjoint, so a parent outside
// the region gives a = b = c = 0 whatever its branch: only the fine loads
// and the output stores are masked, and branches of such parents may be
// anything.
//
// A workgroup of MT_T threads owns the outputs [i0, i0 + 14) x [j0, j0 + 14)
// of the planes [k0, k1].  Its fine footprint, the fine points whose parents
// are within one of the tile, is 32 x 32 (fx = 2 i0 - 2 + lx, fy likewise):
// two positions per thread, lx = tid & 31, ly = (tid >> 5) + 16 m.  LDS
// index conventions: xy[ly * 32 + lx]; xb[lj * 32 + lx] (j = j0 + lj); By of
// parent py = j0 - 1 + lp at [lp * 32 + lx]; Bx of parent px = i0 - 1 + lp
// on row j0 + lj at [lj * 16 + lp].
//
// Two kernels use these routines: the gather from recorded codes (parents
// marched pz = k0 - 1 .. k1 + 1) and the fused adjoint at crse_in (below),
// which recomputes the codes.  Both give the same bits.  The steps are
// host-and-device functions, so a host build runs the same indexing with
// checked accessors (the threads of a workgroup one after the other between
// barriers).

#define MT_T 512
#define MT_O 14                  // outputs per tile edge
#define MT_FP 32                 // footprint edge, 2 MT_O + 4
#define MT_NPOS (MT_FP * MT_FP / MT_T)  // footprint positions per thread
#define MT_PB 16                 // parents per tile edge, MT_O + 2
#define MT_CW 18                 // coarse window edge and X rows, MT_O + 4
#define MT_NY (MT_O * MT_FP)     // y^T outputs (448)
#define MT_NX (MT_O * MT_O)      // x^T outputs (196)
#define MT_RING 4

struct MtTile {
  int i0, j0, k0, k1;
};

// Tile b of the output window [olo, ohi]: x, y tiles of MT_O, z-chunks of kc.
__host__ __device__ inline MtTile mt_tile(const int olo[3], const int ohi[3],
                                          int b, int ntx, int nty, int kc) {
  MtTile t;
  t.i0 = olo[0] + MT_O * (b % ntx);
  t.j0 = olo[1] + MT_O * ((b / ntx) % nty);
  t.k0 = olo[2] + kc * (b / (ntx * nty));
  t.k1 = t.k0 + kc - 1 < ohi[2] ? t.k0 + kc - 1 : ohi[2];
  return t;
}

__host__ __device__ inline void mt_pair(double g0, double g1,
                                        unsigned char code, double &a,
                                        double &b, double &c) {
  const double d = 0.25 * (g1 - g0);
  a = code == MM_MINUS ? d : 0.0;
  b = code == MM_PLUS ? d : 0.0;
  c = ((g0 + g1) + a) - b;
}

// z^T state of one footprint position before parent pz: s[0] the sum so
// far of plane pz - 1 (b(pz - 2) + c(pz - 1)), s[1] that of plane pz
// (b(pz - 1)); plane pz + 1 has none yet.
struct MtWin {
  double s[MT_NPOS][2];
};

__host__ __device__ inline void mt_win_zero(MtWin &w) {
  for (int m = 0; m < MT_NPOS; ++m)
    w.s[m][0] = w.s[m][1] = 0.0;
}

// z^T of parent pz at position m (fine adjoints g0 of plane 2 pz, g1 of
// 2 pz + 1; zeros for a parent outside the region): returns plane pz - 1,
// now complete, and moves the window on by one plane.
__host__ __device__ inline double mt_zt(MtWin &w, int m, unsigned char code,
                                        double g0, double g1) {
  double a, b, c;
  mt_pair(g0, g1, code, a, b, c);
  const double done = w.s[m][0] - a;
  w.s[m][0] = w.s[m][1] + c;
  w.s[m][1] = b;
  return done;
}

// y^T output e < MT_NY (lx = e & 31, j = j0 + (e >> 5)): xy_ld(q) reads
// XYbar, byc(lx, lp) the By code of parent j0 - 1 + lp.
template <class XL, class BC>
__host__ __device__ inline double mt_yt(int e, const XL &xy_ld,
                                        const BC &byc) {
  const int lx = e & (MT_FP - 1), lj = e / MT_FP;
  double a[3], b[3], c[3];
  for (int t = 0; t < 3; ++t) {  // parent j - 1 + t: fine rows 2 (lj + t) + o
    const int r = 2 * (lj + t);
    mt_pair(xy_ld(r * MT_FP + lx), xy_ld((r + 1) * MT_FP + lx), byc(lx, lj + t),
            a[t], b[t], c[t]);
  }
  return (b[0] + c[1]) - a[2];
}

// All y^T outputs of thread tid: xb_st(e, mt_yt(e, ...)).
template <class XL, class BC, class BS>
__host__ __device__ inline void mt_yt_all(int tid, const XL &xy_ld,
                                          const BC &byc, const BS &xb_st) {
#pragma unroll
  for (int n = 0; n < (MT_NY + MT_T - 1) / MT_T; ++n) {
    const int e = tid + MT_T * n;
    if (e >= MT_NY)
      break;
    xb_st(e, mt_yt(e, xy_ld, byc));
  }
}

// x^T output (i0 + li, j0 + lj): xb_ld(q) reads Xbar, bxc(lp, lj) the Bx
// code of parent i0 - 1 + lp on row j0 + lj.
template <class BL, class BC>
__host__ __device__ inline double mt_xt(int li, int lj, const BL &xb_ld,
                                        const BC &bxc) {
  double a[3], b[3], c[3];
  for (int t = 0; t < 3; ++t) {
    const int x = lj * MT_FP + 2 * (li + t);
    mt_pair(xb_ld(x), xb_ld(x + 1), bxc(li + t, lj), a[t], b[t], c[t]);
  }
  return (b[0] + c[1]) - a[2];
}

// ---------------------------------------------------------------------------
// Fused CCC adjoint at crse_in: the tiled adjoint with its codes recomputed
// inside the workgroup from the coarse data (no code array).  Step s of the
// march (coarse plane s, from k0 - 2 to k1 + 3; "produce" steps s <= k1 + 2,
// "own" steps k0 <= s <= k1) runs
//   A  coarse plane s -> cpl (x in [i0 - 2, i0 + 16), y likewise; zero
//      outside the output window, which holds every coarse point the
//      region's values read);
//      barrier
//   B  X(., ., s) -> xpl (the footprint's fine x, rows j0 - 2 .. j0 + 15);
//      own: Bx(., ., s) -> bx ring slot s % 4;
//      y^T of plane s - 3 (xy, by ring) -> xb;
//      barrier
//   C  own: By(., ., s) -> by ring slot s % 4;
//      XY(., ., s) -> the thread's XY window (last value and difference),
//      giving Bz of parent s - 1;
//      x^T of plane s - 3 (xb, bx ring) -> crse_bar;
//      z^T of parent s - 1 (fine loads issued in A); XYbar plane s - 2,
//      now complete, -> xy.
// X and XY are the forward's minmod_apply on the same values and the codes
// mm_code's, so they are the recorded codes bit for bit wherever a fine
// point of the region depends on them.  Ring slots are written in one phase
// and read three phases of their kind later, with barriers between.

// XY of the planes before s at one footprint position, as the Bz code
// needs them: v the last plane's value, dm its difference to the one
// before.
struct MtXY {
  double v[MT_NPOS], dm[MT_NPOS];
};

__host__ __device__ inline void mt_xy_zero(MtXY &x) {
  for (int m = 0; m < MT_NPOS; ++m)
    x.v[m] = x.dm[m] = 0.0;
}

// A: the value cpl[q] (q < MT_CW^2) of coarse plane s: c(x, y, s) for
// points of the output window [olo, ohi], else 0.  Each thread loads
// q = tid + MT_T n, n < MT_NCL.
#define MT_NCL ((MT_CW * MT_CW + MT_T - 1) / MT_T)
template <class C>
__host__ __device__ inline double mt_coarse(const int olo[3], const int ohi[3],
                                            const MtTile &t, int q, int s,
                                            const C &c) {
  const int x = t.i0 - 2 + q % MT_CW, y = t.j0 - 2 + q / MT_CW;
  const bool in = s >= olo[2] && s <= ohi[2] && x >= olo[0] && x <= ohi[0] &&
                  y >= olo[1] && y <= ohi[1];
  return in ? c(x, y, s) : 0.0;
}

// B: X of the plane (xpl_st(q, v), q = r * 32 + lx, row j0 - 2 + r) and,
// when bx is set, its Bx codes (bx_st(q, code)).
template <class CL, class XS, class BS>
__host__ __device__ inline void mt_x(int tid, const CL &cpl_ld,
                                     const XS &xpl_st, bool bx,
                                     const BS &bx_st) {
#pragma unroll
  for (int n = 0; n < (MT_FP * MT_CW + MT_T - 1) / MT_T; ++n) {
    const int e = tid + MT_T * n;
    if (e >= MT_FP * MT_CW)
      break;
    const int lx = e & (MT_FP - 1), r = e / MT_FP;
    const int q = r * MT_CW + (lx >> 1);  // coarse x of parent - 1
    xpl_st(e, minmod_apply(cpl_ld(q), cpl_ld(q + 1), cpl_ld(q + 2), lx & 1));
  }
  if (!bx)
    return;
#pragma unroll
  for (int n = 0; n < (MT_PB * MT_O + MT_T - 1) / MT_T; ++n) {
    const int e = tid + MT_T * n;
    if (e >= MT_PB * MT_O)
      break;
    const int lp = e & (MT_PB - 1), lj = e / MT_PB;
    const int q = (lj + 2) * MT_CW + lp;
    bx_st(e, mm_code(cpl_ld(q), cpl_ld(q + 1), cpl_ld(q + 2)));
  }
}

// C: the By codes of the plane (by_st(q, code)).
template <class XL, class BS>
__host__ __device__ inline void mt_by(int tid, const XL &xpl_ld,
                                      const BS &by_st) {
#pragma unroll
  for (int n = 0; n < (MT_FP * MT_PB + MT_T - 1) / MT_T; ++n) {
    const int e = tid + MT_T * n;
    if (e >= MT_FP * MT_PB)
      break;
    const int lx = e & (MT_FP - 1), lp = e / MT_FP;
    by_st(e, mm_code(xpl_ld(lp * MT_FP + lx), xpl_ld((lp + 1) * MT_FP + lx),
                     xpl_ld((lp + 2) * MT_FP + lx)));
  }
}

// C: XY of plane s at the thread's positions; code[m] gets Bz of parent
// s - 1 (mm_code of XY at s - 2, s - 1, s) and the window moves on.
template <class XL>
__host__ __device__ inline void mt_xy(int tid, const XL &xpl_ld, MtXY &xy,
                                      unsigned char (&code)[MT_NPOS]) {
  const int lx = tid & (MT_FP - 1);
  for (int m = 0; m < MT_NPOS; ++m) {
    const int ly = tid / MT_FP + (MT_T / MT_FP) * m;
    const int r = ly >> 1;
    const double v = minmod_apply(xpl_ld(r * MT_FP + lx),
                                  xpl_ld((r + 1) * MT_FP + lx),
                                  xpl_ld((r + 2) * MT_FP + lx), ly & 1);
    const double dp = v - xy.v[m];
    code[m] = mm_code_d(xy.dm[m], dp);
    xy.dm[m] = dp;
    xy.v[m] = v;
  }
}

// ---------------------------------------------------------------------------
// Tiled record of a CCC operator's codes: the fused march's phases A, B and
// C for steps s = k0 - 1 .. k1 + 1, storing the codes a tile owns instead
// of keeping them in LDS: Bx of parents i0 .. i0 + 13 on its rows, By of
// fine x 2 i0 .. 2 i0 + 27 and parents j0 .. j0 + 13, Bz of fine x, y in
// those ranges and parents k0 .. k1 (each code of the arrays belongs to
// one tile).  The predicates below say which entry of mt_x's / mt_by's
// index space, or which footprint position, is stored.

__host__ __device__ inline bool mt_owns_bx(const MmGeom &g, const MtTile &t,
                                           int q, int &px, int &j) {
  const int lp = q % MT_PB, lj = q / MT_PB;
  px = t.i0 - 1 + lp;
  j = t.j0 + lj;
  return lp >= 1 && lp <= MT_O && px >= (g.rlo[0] >> 1) &&
         px <= (g.rhi[0] >> 1) && j <= g.ohi[1];
}

__host__ __device__ inline bool mt_owns_by(const MmGeom &g, const MtTile &t,
                                           int q, int &fx, int &py) {
  const int lx = q % MT_FP, lp = q / MT_FP;
  fx = 2 * t.i0 - 2 + lx;
  py = t.j0 - 1 + lp;
  return lx >= 2 && lx < 2 + 2 * MT_O && lp >= 1 && lp <= MT_O &&
         fx >= g.rlo[0] && fx <= g.rhi[0] && py >= (g.rlo[1] >> 1) &&
         py <= (g.rhi[1] >> 1);
}

__host__ __device__ inline bool mt_owns_bz(const MmGeom &g, const MtTile &t,
                                           int tid, int m, int &fx, int &fy) {
  const int lx = tid % MT_FP, ly = tid / MT_FP + (MT_T / MT_FP) * m;
  fx = 2 * t.i0 - 2 + lx;
  fy = 2 * t.j0 - 2 + ly;
  return lx >= 2 && lx < 2 + 2 * MT_O && ly >= 2 && ly < 2 + 2 * MT_O &&
         fx >= g.rlo[0] && fx <= g.rhi[0] && fy >= g.rlo[1] && fy <= g.rhi[1];
}

} // namespace cxp
