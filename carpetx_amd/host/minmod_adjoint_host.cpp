// Host build of the MINMOD adjoint's per-point routines
// (carpetx_amd/csrc/minmod_adjoint.hpp): the same record and gather code the
// device kernels run, here over host arrays with every access checked
// against the box it addresses.  It is the CPU check of the kernels'
// indexing and arithmetic (tests/test_minmod_adjoint_host.py compares its
// result with the oracle's Jacobian), not a product path.
#include <cstdio>
#include <vector>

#include "../csrc/minmod_adjoint.hpp"
#include "../csrc/ops.hpp"

using namespace cxp;

namespace {

bool inside(const int lo[3], const int hi[3], int i, int j, int k) {
  return i >= lo[0] && i <= hi[0] && j >= lo[1] && j <= hi[1] && k >= lo[2] &&
         k <= hi[2];
}

long long at(const int lo[3], const int hi[3], int i, int j, int k) {
  const long long nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1;
  return (i - lo[0]) + nx * ((j - lo[1]) + ny * (long long)(k - lo[2]));
}

// An LDS array of a host-emulated workgroup: every element keeps the plane
// it was written for and the phase (barrier interval) of its last write and
// read.  A read must see a value of the expected plane written in an
// earlier phase; no element may be written in a phase that reads it, or
// twice in one phase.
struct Lds {
  const char *name;
  std::vector<double> v;
  std::vector<long long> tag;
  std::vector<int> wph, rph;
  Lds(const char *n, size_t size)
      : name(n), v(size), tag(size, -(1LL << 40)), wph(size, -1),
        rph(size, -1) {}
};

struct LdsCheck {
  int phase = 0;
  long long bad = 0;
  void report(const char *what, const char *name, long long q, long long tag) {
    if (bad++ < 8)
      std::fprintf(stderr, "minmod tiled check: %s %s [%lld] plane %lld phase %d\n",
                   what, name, q, tag, phase);
  }
  void st(Lds &b, long long q, double v, long long tag) {
    if (q < 0 || q >= (long long)b.v.size()) {
      report("store out of range", b.name, q, tag);
      return;
    }
    if (b.rph[q] == phase || b.wph[q] == phase)
      report("store races a read or store of its phase", b.name, q, tag);
    b.v[q] = v;
    b.tag[q] = tag;
    b.wph[q] = phase;
  }
  double ld(Lds &b, long long q, long long tag) {
    if (q < 0 || q >= (long long)b.v.size()) {
      report("load out of range", b.name, q, tag);
      return 0.0;
    }
    if (b.wph[q] >= phase || b.tag[q] != tag)
      report("load of a stale or unsynchronised value", b.name, q, tag);
    b.rph[q] = phase;
    return b.v[q];
  }
};

} // namespace

extern "C" {

// cbar += P^T fbar for the `minmod` operator of the given order and index
// type, linearised at crse (coarse box [clo, chi], which must hold
// CoarseBox(region)); cbar has crse's box, fbar the region's.  Returns the
// number of out-of-range or unwritten accesses (0 when the indexing is
// sound), or -1 for an unknown operator; *ncodes gets the codes per
// component.
long long cxp_host_minmod_adjoint(int order, const int it[3],
                                  const double *crse, const int clo[3],
                                  const int chi[3], const double *fbar,
                                  const int rlo[3], const int rhi[3],
                                  double *cbar, long long *ncodes) {
  OpDesc od;
  if (op_lookup("minmod", order, it, &od))
    return -1;
  for (int d = 0; d < 3; ++d)
    if (!mm_sid_ok(od.sid[d]))
      return -1;
  MmGeom g;
  mm_setup(g, od.sid, rlo, rhi);
  *ncodes = g.ncodes;
  long long bad = 0;
  auto report = [&](const char *what, int w, int i, int j, int k) {
    if (bad++ < 8)
      std::fprintf(stderr, "minmod adjoint host check: %s %d (%d,%d,%d)\n",
                   what, w, i, j, k);
  };
  auto cv = [&](int i, int j, int k) {
    if (!inside(clo, chi, i, j, k)) {
      report("coarse read outside crse", -1, i, j, k);
      return 0.0;
    }
    return crse[at(clo, chi, i, j, k)];
  };
  std::vector<unsigned char> codes((size_t)g.ncodes, 255);
  for (long long n = 0; n < g.ncodes; ++n)
    codes[(size_t)n] = mm_record_one(g, cv, n);
  auto cd = [&](int w, int x, int y, int z) -> unsigned char {
    int blo[3], bhi[3];
    for (int d = 0; d < 3; ++d) {
      blo[d] = g.blo[w][d];
      bhi[d] = g.blo[w][d] + g.bn[w][d] - 1;
    }
    if (!inside(blo, bhi, x, y, z)) {
      report("code outside its array", w, x, y, z);
      return 0;
    }
    const long long m = mm_code_index(g, w, x, y, z);
    if (m < 0 || m >= g.ncodes || codes[(size_t)m] == 255) {
      report("code not recorded", w, x, y, z);
      return 0;
    }
    return codes[(size_t)m];
  };
  auto fv = [&](int x, int y, int z) {
    if (!inside(rlo, rhi, x, y, z)) {
      report("fine read outside the region", -1, x, y, z);
      return 0.0;
    }
    return fbar[at(rlo, rhi, x, y, z)];
  };
  for (int k = g.olo[2]; k <= g.ohi[2]; ++k)
    for (int j = g.olo[1]; j <= g.ohi[1]; ++j)
      for (int i = g.olo[0]; i <= g.ohi[0]; ++i) {
        if (!inside(clo, chi, i, j, k)) {
          report("coarse write outside crse", -1, i, j, k);
          continue;
        }
        cbar[at(clo, chi, i, j, k)] += mm_gather_one(g, cd, fv, i, j, k);
      }
  return bad;
}

// The same adjoint through the kernels' other routines: codes recorded by
// columns in z-chunks of zc (each code must be written exactly once), and,
// for operators MINMOD in all three directions, the tiled gather with
// z-chunks of kc planes, its workgroup's threads run one after the other
// between its barriers, every register-window, LDS, code, fine and coarse
// access checked.  Other operators use the per-point gather.  Returns as
// cxp_host_minmod_adjoint.
long long cxp_host_minmod_adjoint_tiled(int order, const int it[3],
                                        const double *crse, const int clo[3],
                                        const int chi[3], const double *fbar,
                                        const int rlo[3], const int rhi[3],
                                        double *cbar, int zc, int kc,
                                        long long *ncodes) {
  OpDesc od;
  if (op_lookup("minmod", order, it, &od) || zc < 1 || kc < 1)
    return -1;
  for (int d = 0; d < 3; ++d)
    if (!mm_sid_ok(od.sid[d]))
      return -1;
  MmGeom g;
  mm_setup(g, od.sid, rlo, rhi);
  *ncodes = g.ncodes;
  long long bad = 0;
  auto report = [&](const char *what, int w, int i, int j, int k) {
    if (bad++ < 8)
      std::fprintf(stderr, "minmod adjoint tiled check: %s %d (%d,%d,%d)\n",
                   what, w, i, j, k);
  };
  auto cv = [&](int i, int j, int k) {
    if (!inside(clo, chi, i, j, k)) {
      report("coarse read outside crse", -1, i, j, k);
      return 0.0;
    }
    return crse[at(clo, chi, i, j, k)];
  };
  std::vector<unsigned char> codes((size_t)g.ncodes, 255);
  auto code_at = [&](int w, int x, int y, int z) -> long long {
    int blo[3], bhi[3];
    for (int d = 0; d < 3; ++d) {
      blo[d] = g.blo[w][d];
      bhi[d] = g.blo[w][d] + g.bn[w][d] - 1;
    }
    if (!inside(blo, bhi, x, y, z))
      return -1;
    return mm_code_index(g, w, x, y, z);
  };
  auto store = [&](int w, int x, int y, int z, unsigned char v) {
    const long long m = code_at(w, x, y, z);
    if (m < 0 || m >= g.ncodes) {
      report("code store outside its array", w, x, y, z);
      return;
    }
    if (codes[(size_t)m] != 255)
      report("code stored twice", w, x, y, z);
    codes[(size_t)m] = v;
  };
  long long pieces = 0;
  for (int w = 0; w < 3; ++w)
    pieces += mm_columns(g, w, zc);
  for (long long n = 0; n < pieces; ++n)
    mm_record_column(g, cv, store, n, zc);
  for (long long m = 0; m < g.ncodes; ++m)
    if (codes[(size_t)m] == 255)
      report("code never stored", -1, (int)m, 0, 0);
  auto cd = [&](int w, int x, int y, int z) -> unsigned char {
    const long long m = code_at(w, x, y, z);
    if (m < 0 || m >= g.ncodes || codes[(size_t)m] == 255) {
      report("code read outside its array or unrecorded", w, x, y, z);
      return 0;
    }
    return codes[(size_t)m];
  };
  auto fv = [&](int x, int y, int z) {
    if (!inside(rlo, rhi, x, y, z)) {
      report("fine read outside the region", -1, x, y, z);
      return 0.0;
    }
    return fbar[at(rlo, rhi, x, y, z)];
  };
  auto add = [&](int i, int j, int k, double v) {
    if (!inside(clo, chi, i, j, k) || i < g.olo[0] || j < g.olo[1] ||
        k < g.olo[2] || i > g.ohi[0] || j > g.ohi[1] || k > g.ohi[2]) {
      report("coarse write outside the output window", -1, i, j, k);
      return;
    }
    cbar[at(clo, chi, i, j, k)] += v;
  };
  const bool all_minmod = od.sid[0] == SID_MINMOD && od.sid[1] == SID_MINMOD &&
                          od.sid[2] == SID_MINMOD;
  if (!all_minmod) {
    for (int k = g.olo[2]; k <= g.ohi[2]; ++k)
      for (int j = g.olo[1]; j <= g.ohi[1]; ++j)
        for (int i = g.olo[0]; i <= g.ohi[0]; ++i)
          add(i, j, k, mm_gather_one(g, cd, fv, i, j, k));
    return bad;
  }
  // k_mm_gather_tiled
  LdsCheck ck;
  const int ntx = (g.ohi[0] - g.olo[0] + MT_O) / MT_O;
  const int nty = (g.ohi[1] - g.olo[1] + MT_O) / MT_O;
  const int nkc = (g.ohi[2] - g.olo[2] + kc) / kc;
  const int plo[3] = {rlo[0] >> 1, rlo[1] >> 1, rlo[2] >> 1};
  const int phi[3] = {rhi[0] >> 1, rhi[1] >> 1, rhi[2] >> 1};
  auto in_region = [&](int fx, int fy) {
    return fx >= rlo[0] && fx <= rhi[0] && fy >= rlo[1] && fy <= rhi[1];
  };
  for (int b = 0; b < ntx * nty * nkc; ++b) {
    const MtTile t = mt_tile(g.olo, g.ohi, b, ntx, nty, kc);
    Lds xy("xy", MT_FP * MT_FP), xb("xb", MT_NY);
    std::vector<MtWin> win(MT_T);
    for (auto &w : win)
      mt_win_zero(w);
    for (int pz = t.k0 - 1; pz <= t.k1 + 1; ++pz) {
      const bool zin = pz >= plo[2] && pz <= phi[2];
      const int fz0 = 2 * pz;
      const int k = pz - 1;
      const bool out = k >= t.k0;
      for (int tid = 0; tid < MT_T; ++tid)
        for (int m = 0; m < MT_NPOS; ++m) {
          const int fx = 2 * t.i0 - 2 + (tid & (MT_FP - 1));
          const int fy = 2 * t.j0 - 2 + tid / MT_FP + (MT_T / MT_FP) * m;
          double g0 = 0.0, g1 = 0.0;
          unsigned char code = MM_ZERO;
          if (zin && in_region(fx, fy)) {
            code = cd(2, fx, fy, pz);
            if (fz0 >= rlo[2])
              g0 = fv(fx, fy, fz0);
            if (fz0 + 1 <= rhi[2])
              g1 = fv(fx, fy, fz0 + 1);
          }
          const double done = mt_zt(win[tid], m, code, g0, g1);
          if (out)
            ck.st(xy, tid + MT_T * m, done, k);
        }
      if (!out)
        continue;
      ++ck.phase;
      auto byc = [&](int lx, int lp) -> unsigned char {
        const int fx = 2 * t.i0 - 2 + lx, py = t.j0 - 1 + lp;
        if (fx < rlo[0] || fx > rhi[0] || py < plo[1] || py > phi[1])
          return MM_ZERO;
        return cd(1, fx, py, k);
      };
      for (int tid = 0; tid < MT_T; ++tid)
        for (int e = tid; e < MT_NY; e += MT_T)
          ck.st(xb, e,
                mt_yt(e, [&](int q) { return ck.ld(xy, q, k); }, byc), k);
      ++ck.phase;
      auto bxc = [&](int lp, int lj) -> unsigned char {
        const int px = t.i0 - 1 + lp, j = t.j0 + lj;
        if (px < plo[0] || px > phi[0] || j > g.ohi[1])
          return MM_ZERO;
        return cd(0, px, j, k);
      };
      for (int tid = 0; tid < MT_NX; ++tid) {
        const int li = tid % MT_O, lj = tid / MT_O;
        const double v =
            mt_xt(li, lj, [&](int q) { return ck.ld(xb, q, k); }, bxc);
        if (t.i0 + li <= g.ohi[0] && t.j0 + lj <= g.ohi[1])
          add(t.i0 + li, t.j0 + lj, k, v);
      }
    }
  }
  return bad + ck.bad;
}

// The fused CCC adjoint (k_mm_fused) on the host: the workgroups of z-chunks
// of kc planes one after the other, the threads of a phase one after the
// other, barriers as phase boundaries; LDS accesses checked as in the tiled
// gather, fine and coarse accesses as above.  Returns the number of failed
// checks, or -1 for an operator that is not MINMOD in all three directions.
long long cxp_host_minmod_adjoint_fused(int order, const int it[3],
                                        const double *crse, const int clo[3],
                                        const int chi[3], const double *fbar,
                                        const int rlo[3], const int rhi[3],
                                        double *cbar, int kc) {
  OpDesc od;
  if (op_lookup("minmod", order, it, &od) || kc < 1)
    return -1;
  if (!(od.sid[0] == SID_MINMOD && od.sid[1] == SID_MINMOD &&
        od.sid[2] == SID_MINMOD))
    return -1;
  MmGeom g;
  mm_setup(g, od.sid, rlo, rhi);
  LdsCheck ck;
  auto cv = [&](int i, int j, int k) {
    if (!inside(clo, chi, i, j, k)) {
      ck.report("coarse read outside crse", "crse", i, k);
      return 0.0;
    }
    return crse[at(clo, chi, i, j, k)];
  };
  auto fv = [&](int x, int y, int z) {
    if (!inside(rlo, rhi, x, y, z)) {
      ck.report("fine read outside the region", "fbar", x, z);
      return 0.0;
    }
    return fbar[at(rlo, rhi, x, y, z)];
  };
  auto add = [&](int i, int j, int k, double v) {
    if (!inside(clo, chi, i, j, k) || i < g.olo[0] || j < g.olo[1] ||
        k < g.olo[2] || i > g.ohi[0] || j > g.ohi[1] || k > g.ohi[2]) {
      ck.report("coarse write outside the output window", "cbar", i, k);
      return;
    }
    cbar[at(clo, chi, i, j, k)] += v;
  };
  const int ntx = (g.ohi[0] - g.olo[0] + MT_O) / MT_O;
  const int nty = (g.ohi[1] - g.olo[1] + MT_O) / MT_O;
  const int nkc = (g.ohi[2] - g.olo[2] + kc) / kc;
  const int plo2 = rlo[2] >> 1, phi2 = rhi[2] >> 1;
  const int nbx = MT_PB * MT_O, nby = MT_FP * MT_PB;
  for (int b = 0; b < ntx * nty * nkc; ++b) {
    const MtTile t = mt_tile(g.olo, g.ohi, b, ntx, nty, kc);
    Lds cpl("cpl", MT_CW * MT_CW), xpl("xpl", MT_FP * MT_CW),
        xy("xy", MT_FP * MT_FP), xb("xb", MT_NY), bxr("bx", MT_RING * nbx),
        byr("by", MT_RING * nby);
    std::vector<MtWin> win(MT_T);
    std::vector<MtXY> xyv(MT_T);
    for (int tid = 0; tid < MT_T; ++tid) {
      mt_win_zero(win[tid]);
      mt_xy_zero(xyv[tid]);
    }
    for (int s = t.k0 - 2; s <= t.k1 + 3; ++s) {
      const bool produce = s <= t.k1 + 2;
      const bool own = s >= t.k0 && s <= t.k1;
      const int k = s - 3;
      const long long bxs = (long long)(s & (MT_RING - 1)) * nbx;
      const long long bys = (long long)(s & (MT_RING - 1)) * nby;
      const long long bxk = (long long)(k & (MT_RING - 1)) * nbx;
      const long long byk = (long long)(k & (MT_RING - 1)) * nby;
      // A
      if (produce)
        for (int tid = 0; tid < MT_T; ++tid)
          for (int n = 0; n < MT_NCL; ++n) {
            const int q = tid + MT_T * n;
            if (q < MT_CW * MT_CW)
              ck.st(cpl, q, mt_coarse(g.olo, g.ohi, t, q, s, cv), s);
          }
      ++ck.phase;
      // B
      for (int tid = 0; tid < MT_T; ++tid) {
        if (produce)
          mt_x(tid, [&](int q) { return ck.ld(cpl, q, s); },
               [&](int q, double v) { ck.st(xpl, q, v, s); }, own,
               [&](int q, unsigned char c) {
                 if (q < 0 || q >= nbx)
                   ck.report("bx store outside its slot", "bx", q, s);
                 ck.st(bxr, bxs + q, c, s);
               });
        if (k >= t.k0)
          for (int e = tid; e < MT_NY; e += MT_T)
            ck.st(xb, e,
                  mt_yt(e, [&](int q) { return ck.ld(xy, q, k); },
                        [&](int lx, int lp) {
                          const long long q = (long long)lp * MT_FP + lx;
                          if (q < 0 || q >= nby)
                            ck.report("by load outside its slot", "by", q, k);
                          return (unsigned char)ck.ld(byr, byk + q, k);
                        }),
                  k);
      }
      ++ck.phase;
      // C
      for (int tid = 0; tid < MT_T; ++tid) {
        unsigned char bz[MT_NPOS] = {};
        if (produce) {
          if (own)
            mt_by(tid, [&](int q) { return ck.ld(xpl, q, s); },
                  [&](int q, unsigned char c) {
                    if (q < 0 || q >= nby)
                      ck.report("by store outside its slot", "by", q, s);
                    ck.st(byr, bys + q, c, s);
                  });
          mt_xy(tid, [&](int q) { return ck.ld(xpl, q, s); }, xyv[tid], bz);
        }
        if (k >= t.k0 && tid < MT_NX) {
          const int li = tid % MT_O, lj = tid / MT_O;
          const double v = mt_xt(
              li, lj, [&](int q) { return ck.ld(xb, q, k); },
              [&](int lp, int ljj) {
                const long long q = (long long)ljj * MT_PB + lp;
                if (q < 0 || q >= nbx)
                  ck.report("bx load outside its slot", "bx", q, k);
                return (unsigned char)ck.ld(bxr, bxk + q, k);
              });
          if (t.i0 + li <= g.ohi[0] && t.j0 + lj <= g.ohi[1])
            add(t.i0 + li, t.j0 + lj, k, v);
        }
        const int pz = s - 1;
        const bool zt =
            pz >= t.k0 - 1 && pz <= t.k1 + 1 && pz >= plo2 && pz <= phi2;
        const int fz0 = 2 * pz;
        const bool store = s - 2 >= t.k0 && s - 2 <= t.k1;
        for (int m = 0; m < MT_NPOS; ++m) {
          const int fx = 2 * t.i0 - 2 + (tid & (MT_FP - 1));
          const int fy = 2 * t.j0 - 2 + tid / MT_FP + (MT_T / MT_FP) * m;
          const bool in = zt && fx >= rlo[0] && fx <= rhi[0] && fy >= rlo[1] &&
                          fy <= rhi[1];
          const double g0 = in && fz0 >= rlo[2] ? fv(fx, fy, fz0) : 0.0;
          const double g1 = in && fz0 + 1 <= rhi[2] ? fv(fx, fy, fz0 + 1) : 0.0;
          const double done = mt_zt(win[tid], m, bz[m], g0, g1);
          if (store)
            ck.st(xy, tid + MT_T * m, done, s - 2);
        }
      }
    }
  }
  return ck.bad;
}

// The branch codes of one component recorded three ways: mode 0 per point
// (mm_record_one), 1 by columns in z-chunks of kc (k_mm_record_col), 2 by
// tiles with z-chunks of kc (k_mm_record_tiled, CCC only), into out
// (ncodes bytes, *ncodes set).  Modes 1 and 2 check that every code is
// written exactly once and mode 2 its LDS accesses.  Returns the number of
// failed checks, -1 for an unknown operator or mode.
long long cxp_host_minmod_codes(int order, const int it[3], const double *crse,
                                const int clo[3], const int chi[3],
                                const int rlo[3], const int rhi[3], int mode,
                                int kc, unsigned char *out, long long cap,
                                long long *ncodes) {
  OpDesc od;
  if (op_lookup("minmod", order, it, &od) || kc < 1 || mode < 0 || mode > 2)
    return -1;
  for (int d = 0; d < 3; ++d)
    if (!mm_sid_ok(od.sid[d]))
      return -1;
  const bool all_minmod = od.sid[0] == SID_MINMOD &&
                          od.sid[1] == SID_MINMOD && od.sid[2] == SID_MINMOD;
  if (mode == 2 && !all_minmod)
    return -1;
  MmGeom g;
  mm_setup(g, od.sid, rlo, rhi);
  *ncodes = g.ncodes;
  if (g.ncodes > cap)
    return -1;
  LdsCheck ck;
  auto cv = [&](int i, int j, int k) {
    if (!inside(clo, chi, i, j, k)) {
      ck.report("coarse read outside crse", "crse", i, k);
      return 0.0;
    }
    return crse[at(clo, chi, i, j, k)];
  };
  std::vector<int> writes((size_t)g.ncodes, 0);
  auto store = [&](int w, int x, int y, int z, unsigned char v) {
    int blo[3], bhi[3];
    for (int d = 0; d < 3; ++d) {
      blo[d] = g.blo[w][d];
      bhi[d] = g.blo[w][d] + g.bn[w][d] - 1;
    }
    if (!inside(blo, bhi, x, y, z)) {
      ck.report("code store outside its array", "codes", x, w);
      return;
    }
    const long long m = mm_code_index(g, w, x, y, z);
    ++writes[(size_t)m];
    out[m] = v;
  };
  if (mode == 0) {
    for (long long n = 0; n < g.ncodes; ++n)
      out[n] = mm_record_one(g, cv, n);
    return ck.bad;
  }
  if (mode == 1) {
    long long pieces = 0;
    for (int w = 0; w < 3; ++w)
      pieces += mm_columns(g, w, kc);
    for (long long n = 0; n < pieces; ++n)
      mm_record_column(g, cv, store, n, kc);
  } else {
    const int ntx = (g.ohi[0] - g.olo[0] + MT_O) / MT_O;
    const int nty = (g.ohi[1] - g.olo[1] + MT_O) / MT_O;
    const int nkc = (g.ohi[2] - g.olo[2] + kc) / kc;
    const int plo2 = rlo[2] >> 1, phi2 = rhi[2] >> 1;
    for (int b = 0; b < ntx * nty * nkc; ++b) {
      const MtTile t = mt_tile(g.olo, g.ohi, b, ntx, nty, kc);
      Lds cpl("cpl", MT_CW * MT_CW), xpl("xpl", MT_FP * MT_CW);
      std::vector<MtXY> xyv(MT_T);
      for (auto &x : xyv)
        mt_xy_zero(x);
      for (int s = t.k0 - 1; s <= t.k1 + 1; ++s) {
        const bool own = s >= t.k0 && s <= t.k1;
        for (int tid = 0; tid < MT_T; ++tid)
          for (int n = 0; n < MT_NCL; ++n) {
            const int q = tid + MT_T * n;
            if (q < MT_CW * MT_CW)
              ck.st(cpl, q, mt_coarse(g.olo, g.ohi, t, q, s, cv), s);
          }
        ++ck.phase;
        for (int tid = 0; tid < MT_T; ++tid)
          mt_x(tid, [&](int q) { return ck.ld(cpl, q, s); },
               [&](int q, double v) { ck.st(xpl, q, v, s); }, own,
               [&](int q, unsigned char c) {
                 int px, j;
                 if (mt_owns_bx(g, t, q, px, j))
                   store(0, px, j, s, c);
               });
        ++ck.phase;
        for (int tid = 0; tid < MT_T; ++tid) {
          if (own)
            mt_by(tid, [&](int q) { return ck.ld(xpl, q, s); },
                  [&](int q, unsigned char c) {
                    int fx, py;
                    if (mt_owns_by(g, t, q, fx, py))
                      store(1, fx, py, s, c);
                  });
          unsigned char bz[MT_NPOS];
          mt_xy(tid, [&](int q) { return ck.ld(xpl, q, s); }, xyv[tid], bz);
          const int pz = s - 1;
          if (pz >= t.k0 && pz <= t.k1 && pz >= plo2 && pz <= phi2)
            for (int m = 0; m < MT_NPOS; ++m) {
              int fx, fy;
              if (mt_owns_bz(g, t, tid, m, fx, fy))
                store(2, fx, fy, pz, bz[m]);
            }
        }
      }
    }
  }
  for (long long m = 0; m < g.ncodes; ++m)
    if (writes[(size_t)m] != 1)
      ck.report("code not written exactly once", "codes", m, writes[(size_t)m]);
  return ck.bad;
}

} // extern "C"
