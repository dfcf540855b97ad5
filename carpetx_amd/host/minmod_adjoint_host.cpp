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
  const int ntx = (g.ohi[0] - g.olo[0] + MM_TILE) / MM_TILE;
  const int nty = (g.ohi[1] - g.olo[1] + MM_TILE) / MM_TILE;
  const int nkc = (g.ohi[2] - g.olo[2] + kc) / kc;
  std::vector<double> xy(MM_FP * MM_FP), xb(MM_FP * MM_TILE);
  std::vector<char> xy_set(xy.size()), xb_set(xb.size());
  auto xy_st = [&](int q, double v) {
    if (q < 0 || q >= (int)xy.size()) {
      report("xy store out of range", q, 0, 0, 0);
      return;
    }
    xy[q] = v;
    xy_set[q] = 1;
  };
  auto xy_ld = [&](int q) {
    if (q < 0 || q >= (int)xy.size() || !xy_set[q]) {
      report("xy load out of range or unset", q, 0, 0, 0);
      return 0.0;
    }
    return xy[q];
  };
  auto xb_st = [&](int q, double v) {
    if (q < 0 || q >= (int)xb.size()) {
      report("xb store out of range", q, 0, 0, 0);
      return;
    }
    xb[q] = v;
    xb_set[q] = 1;
  };
  auto xb_ld = [&](int q) {
    if (q < 0 || q >= (int)xb.size() || !xb_set[q]) {
      report("xb load out of range or unset", q, 0, 0, 0);
      return 0.0;
    }
    return xb[q];
  };
  std::vector<MmWin> win(MM_T);
  for (int b = 0; b < ntx * nty * nkc; ++b) {
    const MmTile t = mm_tile(g, b, ntx, nty, kc);
    for (auto &w : win)
      mm_win_zero(w);
    for (int pz = t.k0 - 1; pz <= t.k1 + 1; ++pz) {
      for (int tid = 0; tid < MM_T; ++tid)
        mm_zt_add(g, t, tid, pz, win[tid], cd, fv);
      const int k = pz - 1;
      const bool out = k >= t.k0;
      std::fill(xy_set.begin(), xy_set.end(), 0);
      for (int tid = 0; tid < MM_T; ++tid)
        mm_zt_flush(tid, win[tid], out, xy_st);
      if (!out)
        continue;
      std::fill(xb_set.begin(), xb_set.end(), 0);
      for (int tid = 0; tid < MM_T; ++tid)
        mm_yt(g, t, tid, k, xy_ld, xb_st, cd);
      for (int tid = 0; tid < MM_T; ++tid)
        mm_xt(g, t, tid, k, xb_ld, cd, add);
    }
  }
  return bad;
}

} // extern "C"
