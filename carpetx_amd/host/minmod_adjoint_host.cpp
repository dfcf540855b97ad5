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

} // extern "C"
