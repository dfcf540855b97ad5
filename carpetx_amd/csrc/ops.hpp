// Operator registry: (prolongation_type, order, indextype) -> per-direction
// 1-D stencils.  Mirrors get_interpolator (CarpetX/src/driver.cxx:593-642)
// and the operator instantiations of
// CarpetX/src/prolongate_3d_rf2_impl_{poly,cons,ddf,eno,minmod,hermite,
// natural,poly_cons3lfb,poly_eno3lfb}.cxx.
#pragma once

#include <cstring>

#include "stencil1d.hpp"

namespace cxp {

struct OpDesc {
  int sid[3];   // per-direction 1-D operator
  int fb;       // 0 none, 1 linear fallback
  int family;   // index into kFamilies
  int order;    // family order (prolongation_order)
  int indextype[3];
};

static const char *const kFamilies[9] = {
    "interpolate", "conservative", "ddf",           "eno",          "minmod",
    "hermite",     "natural",      "poly-cons3lfb", "poly-eno3lfb"};

inline int vpoly_sid(int p) {
  switch (p) {
  case 1: return SID_VPOLY1;
  case 3: return SID_VPOLY3;
  case 5: return SID_VPOLY5;
  case 7: return SID_VPOLY7;
  }
  return -1;
}
inline int cpoly_sid(int p) { return p >= 0 && p <= 7 ? SID_CPOLY0 + p : -1; }
inline int ccons_sid(int p) { return p >= 0 && p <= 7 ? SID_CCONS0 + p : -1; }

// Returns 0 on success, -1 for an unknown type, -2 for an unsupported order
// (the two CCTK_VERROR cases of driver.cxx:620-625), -3 for a bad indextype.
inline int op_lookup(const char *type, int order, const int indextype[3],
                     OpDesc *out) {
  int fam = -1;
  for (int f = 0; f < 9; ++f)
    if (type && std::strcmp(type, kFamilies[f]) == 0)
      fam = f;
  if (fam < 0)
    return -1;
  bool ok = false;
  switch (fam) {
  case 0: case 2: case 6: // interpolate, ddf, natural
    ok = order == 1 || order == 3 || order == 5 || order == 7;
    break;
  case 1: // conservative
    ok = order == 0 || order == 1;
    break;
  case 3: // eno
    ok = order == 3 || order == 5;
    break;
  default: // minmod, hermite, poly-cons3lfb, poly-eno3lfb
    ok = order == 1 || order == 3 || order == 5;
  }
  if (!ok)
    return -2;
  OpDesc op{};
  op.family = fam;
  op.order = order;
  op.fb = (fam == 7 || fam == 8) ? 1 : 0;
  for (int d = 0; d < 3; ++d) {
    if (indextype[d] != 0 && indextype[d] != 1)
      return -3;
    op.indextype[d] = indextype[d];
    const bool cc = indextype[d] == 1;
    int s = -1;
    switch (fam) {
    case 0: s = cc ? cpoly_sid(order) : vpoly_sid(order); break;
    case 1: s = order == 0 ? (cc ? SID_CCONS0 : SID_VCONS0)
                           : (cc ? SID_CCONS2 : SID_VCONS1); break;
    case 2: s = cc ? ccons_sid(order - 1) : vpoly_sid(order); break;
    case 3: s = cc ? SID_ENO2 : vpoly_sid(order); break;
    case 4: s = cc ? SID_MINMOD : vpoly_sid(order); break;
    case 5: s = cc ? ccons_sid(order) : vpoly_sid(order); break; // HERMITE == POLY
    case 6: s = cc ? ccons_sid(order) : vpoly_sid(order); break;
    case 7: s = cc ? ccons_sid(order < 3 ? order : 3) : vpoly_sid(order); break;
    case 8: s = cc ? (order == 1 ? SID_CCONS0 : SID_ENO2) : vpoly_sid(order); break;
    }
    if (s < 0)
      return -2;
    op.sid[d] = s;
  }
  *out = op;
  return 0;
}

inline bool op_has_eno(const OpDesc &op) {
  return op.sid[0] == SID_ENO2 || op.sid[1] == SID_ENO2 || op.sid[2] == SID_ENO2;
}

// Fallback active (per-var :1535-1538): some CONS/ENO direction above order 1.
inline bool op_fallback_per_var(const OpDesc &op) {
  if (!op.fb)
    return false;
  for (int d = 0; d < 3; ++d) {
    const SidInfo s = sid_info(op.sid[d]);
    if ((s.intp == INTP_CONS || s.intp == INTP_ENO) && s.order > 1)
      return true;
  }
  return false;
}
// Fallback active (per-group :1990-1992): some CONS direction above order 1.
inline bool op_fallback_per_group(const OpDesc &op) {
  if (!op.fb)
    return false;
  for (int d = 0; d < 3; ++d) {
    const SidInfo s = sid_info(op.sid[d]);
    if (s.intp == INTP_CONS && s.order > 1)
      return true;
  }
  return false;
}

// The separable (tensor-pass) kernel applies when no ENO shift search and no
// fallback decision couples the directions.
inline bool op_is_separable(const OpDesc &op, bool per_group) {
  if (op_has_eno(op))
    return false;
  return per_group ? !op_fallback_per_group(op) : !op_fallback_per_var(op);
}

inline int floor2(int i) { return (i - (i & 1)) / 2; }

// CoarseBox (prolongate_3d_rf2_impl.hxx:1226-1254): amrex::coarsen by 2
// (floor; nodal hi rounds up when odd), grown by required_ghosts.
inline void op_coarse_box(const OpDesc &op, const int flo[3], const int fhi[3],
                          int clo[3], int chi[3]) {
  for (int d = 0; d < 3; ++d) {
    const SidInfo s = sid_info(op.sid[d]);
    int lo = floor2(flo[d]);
    int hi = floor2(fhi[d]);
    if (op.indextype[d] == 0 && (fhi[d] & 1))
      hi += 1;
    clo[d] = lo - s.ghosts;
    chi[d] = hi + s.ghosts;
  }
}

} // namespace cxp
