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
// The record pass writes these codes (one byte each); the gather pass runs
// one thread per coarse point (i, j, k) and sums, over the fine points F
// whose stencils reach it,
//   wz(oz, k - pz; Bz(F)) * wy(oy, j - py; By(fx, py, k))
//     * wx(ox, i - px; Bx(px, j, k)) * fbar(F),
// the Jacobian entry of f(F) by the chain rule: f(F) depends on plane k
// through XY(fx, fy, k) alone, which depends on row j through X(fx, j, k)
// alone.  Linear directions take their fixed weights.  No atomics.
//
// The per-point routines are host-and-device functions over accessor
// functors, so a host build runs them with bounds-checked accessors
// (carpetx_amd/host/minmod_adjoint_host.cpp) and the kernels (k_mm_record,
// k_mm_gather in adjoint.hpp) with plain ones.
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
__host__ __device__ inline unsigned char mm_code(double cm, double c0,
                                                 double cp) {
  const double dm = c0 - cm;
  const double dp = cp - c0;
  const double k = copysign(0.125, dm) + copysign(0.125, dp);
  if (k == 0.0)
    return MM_ZERO;
  return fabs(dm) <= fabs(dp) ? MM_MINUS : MM_PLUS;
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
// Tiled gather for operators MINMOD in all three directions (CCC).
//
// A workgroup of MM_T threads owns the coarse outputs i in [i0, i0 + 16),
// j in [j0, j0 + 16) of planes k in [k0, k1], and marches the parents pz
// from k0 - 1 to k1 + 1.  Its fine footprint is fx in [2 i0 - 2, 2 i0 + 34)
// and fy likewise (36 x 36, the fine points whose parents are within one
// of the tile); thread t holds footprint positions t + MM_T m.
//   z^T: each parent pz adds, for both of its fine planes, w(oz, e; Bz) *
//        fbar to XYbar(fx, fy, pz + e), e = -1, 0, 1, kept in a 3-slot
//        register window; after pz, plane pz - 1 is complete;
//   y^T: Xbar(fx, j, k) = sum over the 6 fy whose parents are within one of
//        j of w(oy, j - py; By(fx, py, k)) * XYbar(fx, fy, k)   (LDS);
//   x^T: cbar(i, j, k) += sum over the 6 fx of w(ox, i - px; Bx(px, j, k))
//        * Xbar(fx, j, k)                                       (LDS).
// The steps are host-and-device functions of (thread, state), so the host
// build runs the same indexing with checked accessors (one thread after the
// other between the barriers).

#define MM_T 256
#define MM_TILE 16
#define MM_FP (2 * MM_TILE + 4)         // footprint extent (36)
#define MM_NPOS ((MM_FP * MM_FP + MM_T - 1) / MM_T)  // positions per thread

struct MmTile {
  int i0, j0, k0, k1;
};

struct MmWin {
  double s[MM_NPOS][3];  // XYbar of planes pz - 1, pz, pz + 1
};

__host__ __device__ inline void mm_win_zero(MmWin &w) {
  for (int m = 0; m < MM_NPOS; ++m)
    w.s[m][0] = w.s[m][1] = w.s[m][2] = 0.0;
}

// z^T contributions of parent pz.  codes(w, x, y, z), fbar(x, y, z).
template <class B, class F>
__host__ __device__ inline void mm_zt_add(const MmGeom &g, const MmTile &t,
                                          int tid, int pz, MmWin &win,
                                          const B &codes, const F &fbar) {
  const int plo = g.rlo[2] >> 1, phi = g.rhi[2] >> 1;
  if (pz < plo || pz > phi)
    return;
  for (int m = 0; m < MM_NPOS; ++m) {
    const int q = tid + MM_T * m;
    if (q >= MM_FP * MM_FP)
      break;
    const int fx = 2 * t.i0 - 2 + q % MM_FP, fy = 2 * t.j0 - 2 + q / MM_FP;
    if (fx < g.rlo[0] || fx > g.rhi[0] || fy < g.rlo[1] || fy > g.rhi[1])
      continue;
    const unsigned char code = codes(2, fx, fy, pz);
    for (int oz = 0; oz < 2; ++oz) {
      const int fz = 2 * pz + oz;
      if (fz < g.rlo[2] || fz > g.rhi[2])
        continue;
      const double f = fbar(fx, fy, fz);
      for (int e = -1; e <= 1; ++e)
        win.s[m][e + 1] += mm_weight(code, oz, e) * f;
    }
  }
}

// Stores the completed plane (slot 0) with xy_st(q, v), q < MM_FP^2, when
// store is set, and rotates the window.
template <class XS>
__host__ __device__ inline void mm_zt_flush(int tid, MmWin &win, bool store,
                                            const XS &xy_st) {
  for (int m = 0; m < MM_NPOS; ++m) {
    const int q = tid + MM_T * m;
    if (store && q < MM_FP * MM_FP)
      xy_st(q, win.s[m][0]);
    win.s[m][0] = win.s[m][1];
    win.s[m][1] = win.s[m][2];
    win.s[m][2] = 0.0;
  }
}

// y^T of plane k into xb (MM_FP * MM_TILE values, fx fastest, then j - j0):
// xy_ld(idx) reads the plane, xb_st(q, v) writes.
template <class B, class XL, class BS>
__host__ __device__ inline void mm_yt(const MmGeom &g, const MmTile &t,
                                      int tid, int k, const XL &xy_ld,
                                      const BS &xb_st, const B &codes) {
  for (int q = tid; q < MM_FP * MM_TILE; q += MM_T) {
    const int lx = q % MM_FP, lj = q / MM_FP;
    const int fx = 2 * t.i0 - 2 + lx, j = t.j0 + lj;
    double s = 0.0;
    if (fx >= g.rlo[0] && fx <= g.rhi[0] && j <= g.ohi[1]) {
      const int flo = 2 * j - 2 > g.rlo[1] ? 2 * j - 2 : g.rlo[1];
      const int fhi = 2 * j + 3 < g.rhi[1] ? 2 * j + 3 : g.rhi[1];
      for (int fy = flo; fy <= fhi; ++fy) {
        const int py = fy >> 1;
        s += mm_weight(codes(1, fx, py, k), fy & 1, j - py) *
             xy_ld((fy - (2 * t.j0 - 2)) * MM_FP + lx);
      }
    }
    xb_st(q, s);
  }
}

// x^T of plane k: cbar_add(i, j, k, value) for the thread's output;
// xb_ld(idx) reads y^T's values.
template <class B, class BL, class A>
__host__ __device__ inline void mm_xt(const MmGeom &g, const MmTile &t,
                                      int tid, int k, const BL &xb_ld,
                                      const B &codes, const A &cbar_add) {
  const int i = t.i0 + tid % MM_TILE, j = t.j0 + tid / MM_TILE;
  if (i > g.ohi[0] || j > g.ohi[1])
    return;
  const int flo = 2 * i - 2 > g.rlo[0] ? 2 * i - 2 : g.rlo[0];
  const int fhi = 2 * i + 3 < g.rhi[0] ? 2 * i + 3 : g.rhi[0];
  double s = 0.0;
  for (int fx = flo; fx <= fhi; ++fx) {
    const int px = fx >> 1;
    s += mm_weight(codes(0, px, j, k), fx & 1, i - px) *
         xb_ld((j - t.j0) * MM_FP + (fx - (2 * t.i0 - 2)));
  }
  cbar_add(i, j, k, s);
}

// Tiles of the output window: x, y tiles of MM_TILE and z-chunks of kc.
__host__ __device__ inline MmTile mm_tile(const MmGeom &g, int b, int ntx,
                                          int nty, int kc) {
  MmTile t;
  t.i0 = g.olo[0] + MM_TILE * (b % ntx);
  t.j0 = g.olo[1] + MM_TILE * ((b / ntx) % nty);
  t.k0 = g.olo[2] + kc * (b / (ntx * nty));
  t.k1 = t.k0 + kc - 1 < g.ohi[2] ? t.k0 + kc - 1 : g.ohi[2];
  return t;
}

} // namespace cxp
