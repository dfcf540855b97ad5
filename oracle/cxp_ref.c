/*
 * cxp_ref.c -- the CPU oracle behind the C interface's own signatures.
 *
 * TEST INFRASTRUCTURE ONLY (see cxp_oracle.h and cxp_ref.h).  Each entry
 * point checks its arguments in the order the device library does and
 * returns the same CXP_* status, then calls the oracle:
 *   cxp_ref_op_lookup   -> or_lookup        (driver.cxx:593-642)
 *   cxp_ref_coarse_box  -> or_coarse_box    (impl.hxx:1226-1254)
 *   cxp_ref_prolongate  -> or_prolongate    (impl.hxx:2167-2188)
 *   cxp_ref_restrict    -> or_restrict      (schedule.cxx:2791-2816)
 */
#include "cxp_ref.h"

#include <pthread.h>
#include <string.h>

#include "cxp_oracle.h"

/* One immutable operator per (family, order, centring), like the device
 * library's handles. */
struct cxp_op_s {
  or_op op;
};

static const char *const kFamilies[9] = {
    "interpolate", "conservative", "ddf",           "eno",          "minmod",
    "hermite",     "natural",      "poly-cons3lfb", "poly-eno3lfb"};
static struct cxp_op_s g_ops[9][8][8];
static int g_valid[9][8][8];
static pthread_once_t g_once = PTHREAD_ONCE_INIT;

static void init_ops(void) {
  for (int f = 0; f < 9; ++f)
    for (int o = 0; o < 8; ++o)
      for (int c = 0; c < 8; ++c) {
        const int it[3] = {(c >> 2) & 1, (c >> 1) & 1, c & 1};
        g_valid[f][o][c] = or_lookup(kFamilies[f], o, it, &g_ops[f][o][c].op) == 0;
      }
}

int cxp_ref_op_lookup(const char *type, int order, const int indextype[3],
                      cxp_op *out) {
  if (!type || !indextype || !out)
    return CXP_ERR_ARG;
  for (int d = 0; d < 3; ++d)
    if (indextype[d] != 0 && indextype[d] != 1)
      return CXP_ERR_ARG;
  pthread_once(&g_once, init_ops);
  int fam = -1;
  for (int f = 0; f < 9; ++f)
    if (strcmp(type, kFamilies[f]) == 0)
      fam = f;
  if (fam < 0)
    return CXP_ERR_TYPE;
  const int c = (indextype[0] << 2) | (indextype[1] << 1) | indextype[2];
  if (order < 0 || order > 7 || !g_valid[fam][order][c])
    return CXP_ERR_ORDER;
  *out = &g_ops[fam][order][c];
  return CXP_OK;
}

int cxp_ref_coarse_box(cxp_op op, const int fine_lo[3], const int fine_hi[3],
                       int crse_lo[3], int crse_hi[3]) {
  if (!op || !fine_lo || !fine_hi || !crse_lo || !crse_hi)
    return CXP_ERR_ARG;
  or_coarse_box(&op->op, fine_lo, fine_hi, crse_lo, crse_hi);
  return CXP_OK;
}

static int fab_ok(const cxp_fab *f) {
  if (!f || !f->data || f->ncomp < 1 || f->comp_stride < 0)
    return 0;
  for (int d = 0; d < 3; ++d)
    if (f->hi[d] < f->lo[d])
      return 0;
  const int64_t vol = (int64_t)(f->hi[0] - f->lo[0] + 1) *
                      (f->hi[1] - f->lo[1] + 1) * (f->hi[2] - f->lo[2] + 1);
  return f->ncomp == 1 || f->comp_stride >= vol;
}

static int box_in(const int lo[3], const int hi[3], const int blo[3],
                  const int bhi[3]) {
  for (int d = 0; d < 3; ++d)
    if (lo[d] < blo[d] || hi[d] > bhi[d])
      return 0;
  return 1;
}

static or_fab to_or(const cxp_fab *f) {
  or_fab o;
  o.data = f->data;
  for (int d = 0; d < 3; ++d) {
    o.lo[d] = f->lo[d];
    o.hi[d] = f->hi[d];
  }
  o.comp_stride = f->comp_stride;
  return o;
}

int cxp_ref_prolongate(cxp_op op, const cxp_fab *crse, int crse_comp,
                       const cxp_fab *fine, int fine_comp, int ncomp,
                       const int region_lo[3], const int region_hi[3],
                       int per_group, void *stream) {
  (void)stream;
  if (!op || !region_lo || !region_hi || !fab_ok(crse) || !fab_ok(fine) ||
      ncomp < 1)
    return CXP_ERR_ARG;
  if (crse_comp < 0 || fine_comp < 0 || crse_comp + ncomp > crse->ncomp ||
      fine_comp + ncomp > fine->ncomp)
    return CXP_ERR_COMP;
  for (int d = 0; d < 3; ++d)
    if (region_hi[d] < region_lo[d])
      return CXP_ERR_ARG;
  if (!box_in(region_lo, region_hi, fine->lo, fine->hi))
    return CXP_ERR_REGION;
  int clo[3], chi[3];
  or_coarse_box(&op->op, region_lo, region_hi, clo, chi);
  if (!box_in(clo, chi, crse->lo, crse->hi))
    return CXP_ERR_COARSE;
  if (per_group && ncomp > 16)
    return CXP_ERR_NCOMP;
  const or_fab c = to_or(crse), f = to_or(fine);
  return or_prolongate(&op->op, &c, crse_comp, &f, fine_comp, ncomp,
                       region_lo, region_hi, per_group)
             ? CXP_ERR_ARG
             : CXP_OK;
}

int cxp_ref_restrict(const int indextype[3], const cxp_fab *fine,
                     int fine_comp, const cxp_fab *crse, int crse_comp,
                     int ncomp, const int rlo[3], const int rhi[3],
                     void *stream) {
  (void)stream;
  if (!indextype || !rlo || !rhi || !fab_ok(fine) || !fab_ok(crse) ||
      ncomp < 1)
    return CXP_ERR_ARG;
  if (fine_comp < 0 || crse_comp < 0 || fine_comp + ncomp > fine->ncomp ||
      crse_comp + ncomp > crse->ncomp)
    return CXP_ERR_COMP;
  for (int d = 0; d < 3; ++d) {
    if (indextype[d] != 0 && indextype[d] != 1)
      return CXP_ERR_ARG;
    if (rhi[d] < rlo[d])
      return CXP_ERR_ARG;
  }
  if (!box_in(rlo, rhi, crse->lo, crse->hi))
    return CXP_ERR_REGION;
  for (int d = 0; d < 3; ++d)
    if (2 * rlo[d] < fine->lo[d] || 2 * rhi[d] + indextype[d] > fine->hi[d])
      return CXP_ERR_COARSE;
  const or_fab f = to_or(fine), c = to_or(crse);
  return or_restrict(indextype, &f, fine_comp, &c, crse_comp, ncomp, rlo, rhi)
             ? CXP_ERR_ARG
             : CXP_OK;
}
