// CCC ENO (CC ENO order 3 in all three directions, per-var, no fallback):
// prolongate_3d_rf2_impl.hxx:1417-1517 (shift search) and :1519-1530 (the
// shifted tensor stencil).
//
// The search for a fine point with coarse parent p scores each shift
// s in {-1, 0, 1}^3 by
//   (1 + 2^-26 (|sx| + |sy| + |sz|)) * max(Mx, My, Mz)(p + s),
//   Mx(q) = max over dj, dk in [-1, 1] of |D2x(q + (0, dj, dk))|
// (D2x the second undivided difference along x; My, Mz alike): with ENO in
// every direction the stencil lines are those of the shifted centre for
// either fine offset, so the score reads one coarse field, Mall, at p + s,
// and all eight children of p take the same shift.  A workgroup owns
// 8 x 8 x PZ parents, one per thread (PZ = 4 for the forward, 8 for the
// codes-only kernel; EcGeom below):
//   A  the coarse window, parents - 2 .. parents + 2 (12 x 12 x (PZ + 4))
//                                                                    -> LDS
//   B  |D2| along x, y, z where Mall needs them                      -> LDS
//   C  Mall on parents - 1 .. parents + 1 (10 x 10 x (PZ + 2))       -> LDS
//   D  per parent: the scan over the 27 shifts in the reference's
//      order (sz outermost, sx innermost; the first strictly smallest
//      score wins, none finite leaves shift 0);
//   E  per parent: the 27 coarse values around p + s and the 8 children,
//      each by the reference's nested x, y, z passes (eno2_apply, the
//      forward's own routine), so the values are the generic kernel's bit
//      for bit.
// |D2| is fmax(0, |undiv2|) and the maxima are fmax chains from 0, as the
// generic search takes them (fmax is exact and ignores NaN).
//
// The adjoint from per-fine-point codes (the shift of each point's parent)
// gathers per coarse point q over the parents p within 2 of it whose
// shifted stencil covers q (t = q - p - s(p) in [-1, 1]^3):
//   crse_bar(q) += sum_p sum_o wz(sz, oz, tz) wy(sy, oy, ty) wx(sx, ox, tx)
//                  fine_bar(2 p + o),
// parents in ascending z, y, x order and, per parent, oz outermost, ox
// innermost.
//
// The steps are host-and-device functions over accessors, so a host build
// runs them with checked accesses (carpetx_amd/host/eno_ccc_host.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include "stencil1d.hpp"

namespace cxp {

#define EC_PX 8
#define EC_PY 8
#define EC_CX (EC_PX + 4)  // coarse window
#define EC_CY (EC_PY + 4)
#define EC_MX (EC_PX + 2)  // Mall window
#define EC_MY (EC_PY + 2)

// A tile of 8 x 8 x PZ parents, one thread each: the forward takes PZ = 4
// (its eight children per thread hold many registers), the codes-only
// kernel PZ = 8 (less window per parent: 3.4 coarse values and 8.4 |D2|
// entries instead of 4.5 and 10.9).
template <int PZ> struct EcGeom {
  static constexpr int T = EC_PX * EC_PY * PZ;     // threads
  static constexpr int CZ = PZ + 4, MZ = PZ + 2;   // window, Mall planes
  static constexpr int NC = EC_CX * EC_CY * CZ;    // window values
  static constexpr int NDX = EC_MX * EC_CY * CZ;   // |D2| along x
  static constexpr int NDY = EC_CX * EC_MY * CZ;   // along y
  static constexpr int NDZ = EC_CX * EC_CY * MZ;   // along z
  static constexpr int ND = NDX + NDY + NDZ;
  static constexpr int NM = EC_MX * EC_MY * MZ;    // Mall values
};
#define EC_PZ 4        // the forward's tile depth
#define EC_PZ_CODES 8  // the codes-only kernel's
#define EC_T (EcGeom<EC_PZ>::T)
#define EC_NC (EcGeom<EC_PZ>::NC)                  // 1152
#define EC_NDX (EcGeom<EC_PZ>::NDX)                // 960
#define EC_ND (EcGeom<EC_PZ>::ND)
#define EC_NM (EcGeom<EC_PZ>::NM)                  // 600

// Parent tile b of the parents [plo, phi]: origin p0.
template <int PZ = EC_PZ>
__host__ __device__ inline void ec_tile(const int plo[3], int b, int ntx,
                                        int nty, int p0[3]) {
  p0[0] = plo[0] + EC_PX * (b % ntx);
  p0[1] = plo[1] + EC_PY * ((b / ntx) % nty);
  p0[2] = plo[2] + PZ * (b / (ntx * nty));
}

__host__ __device__ inline int ec_cidx(int x, int y, int z) {
  return (z * EC_CY + y) * EC_CX + x;
}

// A: window value q < EC_NC: c(x, y, z) inside the coarse window
// [olo, ohi] (which holds every value the region's points read), else 0.
template <class C>
__host__ __device__ inline double ec_coarse(const int olo[3], const int ohi[3],
                                            const int p0[3], int q,
                                            const C &c) {
  const int x = p0[0] - 2 + q % EC_CX, y = p0[1] - 2 + (q / EC_CX) % EC_CY,
            z = p0[2] - 2 + q / (EC_CX * EC_CY);
  const bool in = x >= olo[0] && x <= ohi[0] && y >= olo[1] && y <= ohi[1] &&
                  z >= olo[2] && z <= ohi[2];
  return in ? c(x, y, z) : 0.0;
}

__host__ __device__ inline double ec_undiv2(double m, double c, double p) {
  double dd = -2.0 * c;
  dd += 1.0 * (m + p);
  return dd;
}

// B: |D2| entry e < ND (x array, then y, then z), from the window.
template <int PZ = EC_PZ, class CL>
__host__ __device__ inline double ec_d2(int e, const CL &cl) {
  constexpr int NDX = EcGeom<PZ>::NDX, NDY = EcGeom<PZ>::NDY;
  int a, b, c, s;
  if (e < NDX) {  // (x', y', z') = window (x' + 1, y', z')
    a = e % EC_MX;
    b = (e / EC_MX) % EC_CY;
    c = e / (EC_MX * EC_CY);
    s = 1;
    const int q = ec_cidx(a, b, c);
    return fmax(0.0, fabs(ec_undiv2(cl(q), cl(q + s), cl(q + 2 * s))));
  }
  if (e < NDX + NDY) {  // window (x', y' + 1, z')
    e -= NDX;
    a = e % EC_CX;
    b = (e / EC_CX) % EC_MY;
    c = e / (EC_CX * EC_MY);
    s = EC_CX;
  } else {  // window (x', y', z' + 1)
    e -= NDX + NDY;
    a = e % EC_CX;
    b = (e / EC_CX) % EC_CY;
    c = e / (EC_CX * EC_CY);
    s = EC_CX * EC_CY;
  }
  const int q = ec_cidx(a, b, c);
  return fmax(0.0, fabs(ec_undiv2(cl(q), cl(q + s), cl(q + 2 * s))));
}

// C: one direction's part of Mall at window point m = (mx, my, mz) of the
// Mall box (parents - 1 + m): the fmax chain from 0 over the 3 x 3 |D2|
// entries across that direction (dl(e) reads that direction's array, from
// its own entry 0).
template <int DIR, class DL>
__host__ __device__ inline double ec_mall_dir(int q, const DL &dl) {
  const int mx = q % EC_MX, my = (q / EC_MX) % EC_MY, mz = q / (EC_MX * EC_MY);
  double dd = 0.0;
  for (int a = -1; a <= 1; ++a)
    for (int b = -1; b <= 1; ++b) {
      if constexpr (DIR == 0)  // a = dk, b = dj
        dd = fmax(dd, dl(((mz + 1 + a) * EC_CY + (my + 1 + b)) * EC_MX + mx));
      else if constexpr (DIR == 1)  // a = dk, b = di
        dd = fmax(dd, dl(((mz + 1 + a) * EC_MY + my) * EC_CX + (mx + 1 + b)));
      else  // a = dj, b = di
        dd = fmax(dd, dl((mz * EC_CY + (my + 1 + a)) * EC_CX + (mx + 1 + b)));
    }
  return dd;
}

// C: Mall at window point m from the three |D2| arrays (dl(e) reads entry
// e of the x array, then y, then z): max(Mx, My, Mz).
template <int PZ = EC_PZ, class DL>
__host__ __device__ inline double ec_mall(int q, const DL &dl) {
  constexpr int NDX = EcGeom<PZ>::NDX, NDY = EcGeom<PZ>::NDY;
  const double ddx = ec_mall_dir<0>(q, dl);
  const double ddy = ec_mall_dir<1>(q, [&](int e) { return dl(NDX + e); });
  const double ddz =
      ec_mall_dir<2>(q, [&](int e) { return dl(NDX + NDY + e); });
  return fmax(fmax(ddx, ddy), ddz);
}

// D: the shift of local parent l (ml(q) reads Mall), as the code
// (sx + 1) + 3 (sy + 1) + 9 (sz + 1).
template <class ML>
__host__ __device__ inline int ec_search(const int l[3], const ML &ml) {
  double best = __builtin_inf();
  int code = 13;  // shift 0
  for (int sk = -1; sk <= 1; ++sk)
    for (int sj = -1; sj <= 1; ++sj)
      for (int si = -1; si <= 1; ++si) {
        const double penalty =
            1.0 + 0x1p-26 * (double)(abs(si) + abs(sj) + abs(sk));
        const double dd =
            penalty * ml(((l[2] + 1 + sk) * EC_MY + (l[1] + 1 + sj)) * EC_MX +
                         (l[0] + 1 + si));
        if (dd < best) {
          best = dd;
          code = (si + 1) + 3 * (sj + 1) + 9 * (sk + 1);
        }
      }
  return code;
}

// E: the eight children of local parent l with shift code `code`, handed
// to out(ox, oy, oz, value) one x offset at a time (cl(q) reads the
// window), so only one offset's x-pass results are live.
template <class CL, class OUT>
__host__ __device__ inline void ec_children_each(const int l[3], int code,
                                                 const CL &cl, const OUT &out) {
  const int s[3] = {code % 3 - 1, (code / 3) % 3 - 1, code / 9 - 1};
  for (int ox = 0; ox < 2; ++ox) {
    double X[3][3];  // [dk - sz + 1][dj - sy + 1]
    for (int kk = 0; kk < 3; ++kk)
      for (int jj = 0; jj < 3; ++jj) {
        const int row = ec_cidx(l[0] + 2, l[1] + 2 + s[1] - 1 + jj,
                                l[2] + 2 + s[2] - 1 + kk);
        X[kk][jj] = eno2_apply([&](int o) { return cl(row + o); }, s[0], ox);
      }
    for (int oy = 0; oy < 2; ++oy) {
      double Y[3];
      for (int kk = 0; kk < 3; ++kk)
        Y[kk] = eno2_apply([&](int o) { return X[kk][o - s[1] + 1]; }, s[1],
                           oy);
      for (int oz = 0; oz < 2; ++oz)
        out(ox, oy, oz,
            eno2_apply([&](int o) { return Y[o - s[2] + 1]; }, s[2], oz));
    }
  }
}

// E: the eight children as v[oz][oy][ox].
template <class CL>
__host__ __device__ inline void ec_children(const int l[3], int code,
                                            const CL &cl, double (&v)[2][2][2]) {
  ec_children_each(l, code, cl, [&](int ox, int oy, int oz, double val) {
    v[oz][oy][ox] = val;
  });
}

// Weight of coarse offset s + t (t in [-1, 1]) in eno2_apply(., s, o).
__host__ __device__ inline double ec_w(int s, int o, int t) {
  return o == 0 ? eno2_w(s + 1, t + 1) : eno2_w(1 - s, 1 - t);
}

// The adjoint's sum for coarse point q: code(p) returns the shift code of
// parent p (p within the region's parents), fb(x, y, z) the fine adjoint
// (0 outside the region).
template <class CD, class FB>
__host__ __device__ inline double ec_adj_one(const int q[3], const int plo[3],
                                             const int phi[3], const CD &code,
                                             const FB &fb) {
  double acc = 0.0;
  for (int pz = q[2] - 2; pz <= q[2] + 2; ++pz) {
    if (pz < plo[2] || pz > phi[2])
      continue;
    for (int py = q[1] - 2; py <= q[1] + 2; ++py) {
      if (py < plo[1] || py > phi[1])
        continue;
      for (int px = q[0] - 2; px <= q[0] + 2; ++px) {
        if (px < plo[0] || px > phi[0])
          continue;
        const int c = code(px, py, pz);
        const int t[3] = {q[0] - px - (c % 3 - 1), q[1] - py - ((c / 3) % 3 - 1),
                          q[2] - pz - (c / 9 - 1)};
        if (t[0] < -1 || t[0] > 1 || t[1] < -1 || t[1] > 1 || t[2] < -1 ||
            t[2] > 1)
          continue;
        const int s[3] = {c % 3 - 1, (c / 3) % 3 - 1, c / 9 - 1};
        double g = 0.0;
        for (int oz = 0; oz < 2; ++oz) {
          double gy = 0.0;
          for (int oy = 0; oy < 2; ++oy) {
            double gx = 0.0;
            for (int ox = 0; ox < 2; ++ox)
              gx += ec_w(s[0], ox, t[0]) *
                    fb(2 * px + ox, 2 * py + oy, 2 * pz + oz);
            gy += ec_w(s[1], oy, t[1]) * gx;
          }
          g += ec_w(s[2], oz, t[2]) * gy;
        }
        acc += g;
      }
    }
  }
  return acc;
}

// ---- the tiled adjoint
//
// A workgroup of EA_NT threads owns EA_TX x EA_TY coarse output columns
// (one per thread) of a z-chunk of the output window [plo - 2, phi + 2]
// and marches the parent planes pz = qz0 - 2 .. qz1 + 2.  Per plane:
//   L  each parent slot of the (EA_TX + 4) x (EA_TY + 4) parents around the
//      tile: from its eight fine adjoints, the 18 partial sums
//        H[oz][ty][tx] = sum_oy wy(sy, oy, ty) sum_ox wx(sx, ox, tx) fine_bar
//      and an info word: which of the 25 columns within 2 of the parent its
//      shifted stencil covers, the H index base and sz          -> LDS
//   G  per output column: the 25 parent slots within 2 in x and y; a slot
//      covering the column adds its H[0], H[1] at (tx, ty) to the sums B of
//      its z-shift sz; then, for each sz and tz, output plane pz + sz + tz
//      of a five-plane register window gains wz(sz, 0, tz) B[sz][0] +
//      wz(sz, 1, tz) B[sz][1] (the z weights are compile-time constants);
//      plane pz - 2 is then complete.
// The sums are the Jacobian transpose of ec_children's values for the
// recorded shifts (ec_adj_one), grouped by parent plane and z-shift.
#define EA_TX 16
#ifndef EA_TY
#define EA_TY 16
#endif
#define EA_NT (EA_TX * EA_TY)
#define EA_PX (EA_TX + 4)
#define EA_PY (EA_TY + 4)
#define EA_NP (EA_PX * EA_PY)  // parent slots per plane
#define EA_NH 18               // H values per parent

// Tile b of the output window [olo, ...] with ntx x nty column tiles and
// z-chunks of zc planes: origin q0.
__host__ __device__ inline void ea_tile(const int olo[3], int b, int ntx,
                                        int nty, int zc, int q0[3]) {
  q0[0] = olo[0] + EA_TX * (b % ntx);
  q0[1] = olo[1] + EA_TY * ((b / ntx) % nty);
  q0[2] = olo[2] + zc * (b / (ntx * nty));
}

// The info word of a parent with shift code c: bit (dx + 2) + 5 (dy + 2)
// is set when the parent's shifted stencil covers the column at offset
// (dx, dy) from it (|dx - sx| <= 1 and |dy - sy| <= 1); bits 25-28 hold
// kb = 4 - 3 sy - sx, so H index k(tx, ty) = kb + 3 dy + dx; bits 29-30
// hold sz + 1.  0: no parent.
__host__ __device__ inline unsigned ea_info(int c) {
  const int sx = c % 3 - 1, sy = (c / 3) % 3 - 1, sz = c / 9 - 1;
  const unsigned mask = 0x1CE7u << ((sx + 1) + 5 * (sy + 1));  // 3 x 3 block
  return mask | (unsigned)(4 - 3 * sy - sx) << 25 | (unsigned)(sz + 1) << 29;
}

// L, loads: parent slot n (< EA_NP) of plane pz around tile origin q0.
// Returns its shift code and its eight fine adjoints in f[oz][oy][ox], or
// -1 when the parent is outside [plo, phi] (f untouched).
// code(px, py, pz) reads a parent's shift code, kids(px, py, pz, f) its
// children's fine adjoints (0 outside the region).
template <class CD, class KD>
__host__ __device__ inline int ea_load(const int q0[3], int n, int pz,
                                       const int plo[3], const int phi[3],
                                       const CD &code, const KD &kids,
                                       double (&f)[2][2][2]) {
  const int px = q0[0] - 2 + n % EA_PX, py = q0[1] - 2 + n / EA_PX;
  if (px < plo[0] || px > phi[0] || py < plo[1] || py > phi[1] ||
      pz < plo[2] || pz > phi[2])
    return -1;
  kids(px, py, pz, f);
  return code(px, py, pz);
}

// L, sums: the info word of a parent with code c and fine adjoints f, and
// its 18 partial sums into h (w(s, o, t) = ec_w(s, o, t)).  c < 0: info 0,
// h untouched.
template <class W>
__host__ __device__ inline unsigned ea_sums(int c, const double (&f)[2][2][2],
                                            const W &w, double (&h)[EA_NH]) {
  if (c < 0)
    return 0u;
  const int sx = c % 3 - 1, sy = (c / 3) % 3 - 1;
#pragma unroll
  for (int tx = -1; tx <= 1; ++tx) {
    const double wx0 = w(sx, 0, tx), wx1 = w(sx, 1, tx);
#pragma unroll
    for (int oz = 0; oz < 2; ++oz) {
      const double gx0 = wx0 * f[oz][0][0] + wx1 * f[oz][0][1];
      const double gx1 = wx0 * f[oz][1][0] + wx1 * f[oz][1][1];
#pragma unroll
      for (int ty = -1; ty <= 1; ++ty)
        h[oz * 9 + (ty + 1) * 3 + (tx + 1)] =
            w(sy, 0, ty) * gx0 + w(sy, 1, ty) * gx1;
    }
  }
  return ea_info(c);
}

// G: output column (lx, ly) of the tile.  cl(n) reads slot n's info word;
// hp(b, h0, h1) reads the H pair at byte offset b = 16 i of the pair array
// (pair i = k (EA_NP + 1) + n holds H[0][k], H[1][k] of slot n; pair EA_NP
// holds zeros); acc[u] is output plane pz - 2 + u.  A row of five slots at
// a time: their info words, then for each slot and each z-shift class the
// pair it adds to that class's sums: its own where the slot covers the
// column and has that class, else the zero pair.  So every sum takes one
// read and one add per slot, with no branches and no selects of values;
// offsets are formed in bytes, so masking them is the only index work.
template <class CL, class HP>
__host__ __device__ inline void ea_gather(int lx, int ly, const CL &cl,
                                          const HP &hp, double (&acc)[5]) {
  constexpr int PB = 16 * (EA_NP + 1);  // bytes per H index k
  constexpr int ZB = 16 * EA_NP;        // the zero pair
  // slot of offset (dx, dy): nb + (2 - dy) EA_PX + (2 - dx), offsets >= 0
  const int nb = ly * EA_PX + lx;
  double b[3][2] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};  // [sz + 1][oz]
  // rows one after another (not unrolled: an unrolled body keeps every
  // row's reads live at once)
#pragma unroll 1
  for (int dy = 2; dy >= -2; --dy) {
    unsigned inf[5];
#pragma unroll
    for (int dx = 2; dx >= -2; --dx)
      inf[2 - dx] = cl(nb + (2 - dy) * EA_PX + (2 - dx));
    int off[5][3];
#pragma unroll
    for (int dx = 2; dx >= -2; --dx) {
      const int n = nb + (2 - dy) * EA_PX + (2 - dx);
      const unsigned info = inf[2 - dx];
      // all ones when the slot covers the column, else 0
      const unsigned vm =
          0u - ((info >> ((dx + 2) + 5 * (dy + 2))) & 1u);
      // own pair minus the zero pair, in bytes
      const int own = (int)((info >> 25) & 15u) * PB +
                      16 * (n + (3 * dy + dx) * (EA_NP + 1)) - ZB;
      // byte c of sel is 0xff for the slot's class c (none when it does
      // not cover the column)
      const unsigned sel = (0xffu << (8u * (info >> 29))) & vm;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const unsigned m = (unsigned)((int)(sel << (24 - 8 * c)) >> 24);
        off[2 - dx][c] = ZB + (int)((unsigned)own & m);
      }
    }
#pragma unroll
    for (int i = 0; i < 5; ++i)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        double h0, h1;
        hp(off[i][c], h0, h1);
        b[c][0] += h0;
        b[c][1] += h1;
      }
  }
#pragma unroll
  for (int sz = -1; sz <= 1; ++sz)
#pragma unroll
    for (int tz = -1; tz <= 1; ++tz)
      acc[sz + tz + 2] += ec_w(sz, 0, tz) * b[sz + 1][0] +
                          ec_w(sz, 1, tz) * b[sz + 1][1];
}

} // namespace cxp
