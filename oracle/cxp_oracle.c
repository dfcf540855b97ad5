/*
 * cxp_oracle.c -- CPU restatement of CarpetX's rf=2 prolongation and
 * restriction.  TEST INFRASTRUCTURE ONLY (see cxp_oracle.h).
 *
 * Built twice by oracle/Makefile:
 *   liboracle_strict.so  -O2 -ffp-contract=off        (the parity checker)
 *   liboracle_fast.so    -O3 -ffp-contract=fast -march=x86-64-v3 -fopenmp
 *                        (the CPU baseline, flags of the reference CI,
 *                         scripts/actions-cpu-real64.cfg:21,42-49)
 *
 * What pins each function (tests/test_oracle_*.py):
 *   weight tables      transcription of prolongate_3d_rf2_impl.hxx:56-384
 *                      (tests/golden/coeffs_reference.json) AND the exact
 *                      rational derivation in carpetx_amd/csrc/gen_coeffs.py
 *   or_interp1d        known-answer tests restated from test_interp1d
 *                      (prolongate_3d_rf2_impl.hxx:882-1174): exact
 *                      reproduction of x^q, exact conservation/integral
 *   tensor product     separable closed form: P(p(x)q(y)r(z)) equals the
 *                      product of 1-D results; polynomial exactness
 *   ENO shift search   hand-built fixtures (step, kink, linear ties), and a
 *                      brute-force 1-D smoothness argmin on tiny inputs
 *   fallback (LFB)     hand-built fixtures: monotone data never falls back,
 *                      a step always does, next to closed forms
 *   or_coarse_box      exhaustive closed-form check on boxes with negative lo
 *   or_restrict        closed form (injection / mean of children) and
 *                      restrict(prolongate(u)) == u
 *   summation order inside AMReX average_down: parity unpinned (AMReX 25.11
 *                      is not vendored); immaterial at rel 1e-13.
 */
#include "cxp_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Weight tables (prolongate_3d_rf2_impl.hxx:56-384)                        */

static const double VC_POLY_1[2] = {1. / 2, 1. / 2};
static const double VC_POLY_3[4] = {-1. / 16, 9. / 16, 9. / 16, -1. / 16};
static const double VC_POLY_5[6] = {3. / 256,  -25. / 256, 75. / 128,
                                    75. / 128, -25. / 256, 3. / 256};
static const double VC_POLY_7[8] = {-5. / 2048,   49. / 2048,  -245. / 2048,
                                    1225. / 2048, 1225. / 2048, -245. / 2048,
                                    49. / 2048,   -5. / 2048};

static const double CC_POLY_0[1] = {1.};
static const double CC_POLY_1[2] = {1. / 4, 3. / 4};
static const double CC_POLY_2[3] = {5. / 32, 15. / 16, -3. / 32};
static const double CC_POLY_3[4] = {-5. / 128, 35. / 128, 105. / 128,
                                    -7. / 128};
static const double CC_POLY_4[5] = {-45. / 2048, 105. / 512, 945. / 1024,
                                    -63. / 512, 35. / 2048};
static const double CC_POLY_5[6] = {63. / 8192,   -495. / 8192, 1155. / 4096,
                                    3465. / 4096, -693. / 8192, 77. / 8192};
static const double CC_POLY_6[7] = {273. / 65536,  -1287. / 32768,
                                    15015. / 65536, 15015. / 16384,
                                    -9009. / 65536, 1001. / 32768,
                                    -231. / 65536};
static const double CC_POLY_7[8] = {
    -429. / 262144,   4095. / 262144,   -19305. / 262144, 75075. / 262144,
    225225. / 262144, -27027. / 262144, 5005. / 262144,   -495. / 262144};

static const double CC_CONS_0[1] = {1.};
static const double CC_CONS_1[2] = {1. / 4, 3. / 4};
static const double CC_CONS_2[3] = {1. / 8, 1., -1. / 8};
static const double CC_CONS_3[4] = {-3. / 64, 17. / 64, 55. / 64, -5. / 64};
static const double CC_CONS_4[5] = {-3. / 128, 11. / 64, 1., -11. / 64,
                                    3. / 128};
static const double CC_CONS_5[6] = {5. / 512,   -37. / 512, 69. / 256,
                                    231. / 256, -63. / 512, 7. / 512};
static const double CC_CONS_6[7] = {5. / 1024,    -11. / 256, 201. / 1024, 1.,
                                    -201. / 1024, 11. / 256,  -5. / 1024};
static const double CC_CONS_7[8] = {
    -35. / 16384,   325. / 16384,   -1439. / 16384, 4441. / 16384,
    15159. / 16384, -2481. / 16384, 459. / 16384,   -45. / 16384};

/* VC CONS (deprecated in the reference): coeffs0 / coeffs1 */
static const double VC_CONS0_0[1] = {1.};
static const double VC_CONS1_1[2] = {1. / 2, 1. / 2};

/* CC ENO 2: row m is the 3-cell conservative stencil over [m-2, m] */
static const double CC_ENO2[3][3] = {
    {-1. / 8, 1. / 2, 5. / 8}, {1. / 8, 1., -1. / 8}, {11. / 8, -1. / 2, 1. / 8}};

static const double *vc_poly_tab(int p) {
  switch (p) {
  case 1: return VC_POLY_1;
  case 3: return VC_POLY_3;
  case 5: return VC_POLY_5;
  case 7: return VC_POLY_7;
  }
  return NULL;
}
static const double *cc_poly_tab(int p) {
  switch (p) {
  case 0: return CC_POLY_0;
  case 1: return CC_POLY_1;
  case 2: return CC_POLY_2;
  case 3: return CC_POLY_3;
  case 4: return CC_POLY_4;
  case 5: return CC_POLY_5;
  case 6: return CC_POLY_6;
  case 7: return CC_POLY_7;
  }
  return NULL;
}
static const double *cc_cons_tab(int p) {
  switch (p) {
  case 0: return CC_CONS_0;
  case 1: return CC_CONS_1;
  case 2: return CC_CONS_2;
  case 3: return CC_CONS_3;
  case 4: return CC_CONS_4;
  case 5: return CC_CONS_5;
  case 6: return CC_CONS_6;
  case 7: return CC_CONS_7;
  }
  return NULL;
}

double or_coeff(int cent, int intp, int order, int row, int i) {
  const double *t = NULL;
  int n = order + 1;
  if (cent == OR_VC && (intp == OR_POLY || intp == OR_HERMITE))
    t = vc_poly_tab(order);
  else if (cent == OR_CC && intp == OR_POLY)
    t = cc_poly_tab(order);
  else if (cent == OR_CC && intp == OR_CONS)
    t = cc_cons_tab(order);
  else if (cent == OR_CC && intp == OR_ENO && order == 2 && row >= 0 &&
           row < 3 && i >= 0 && i < 3)
    return CC_ENO2[row][i];
  if (!t || i < 0 || i >= n)
    return NAN;
  return t[i];
}

/* ------------------------------------------------------------------------ */
/* 1-D operators (prolongate_3d_rf2_impl.hxx:388-878)                        */

int or_required_ghosts(int cent, int intp, int order) {
  if (intp == OR_ENO)
    return order; /* :783 */
  if (intp == OR_MINMOD)
    return 1; /* :847 */
  (void)cent;
  return (order + 1) / 2; /* :398, :462, :518, :585, :636 */
}

void or_stencil_radius(int cent, int intp, int P, int shift, int off,
                       int r[2]) {
  const int N = P + 1;
  if (cent == OR_VC && (intp == OR_POLY || intp == OR_HERMITE)) {
    /* :401-408, :520-528 */
    if (off == 0) {
      r[0] = 0;
      r[1] = 0;
    } else {
      const int i0 = N / 2 - off;
      r[0] = -i0;
      r[1] = P - i0;
    }
  } else if (cent == OR_CC && intp == OR_POLY) { /* :464-476 */
    const int i0 = N / 2;
    if (off == 0) {
      r[0] = -i0;
      r[1] = N - 1 - i0;
    } else {
      r[0] = -i0 + P % 2;
      r[1] = N - 1 - i0 + P % 2;
    }
  } else if (cent == OR_VC && intp == OR_CONS) { /* :587-601 */
    if (off == 0) {
      const int i0 = P / 2;
      r[0] = -i0;
      r[1] = P / 2 * 2 - i0;
    } else {
      const int i0 = (P + 1) / 2 - 1;
      r[0] = -i0;
      r[1] = (P + 1) / 2 * 2 - 1 - i0;
    }
  } else if (cent == OR_CC && intp == OR_CONS) { /* :638-650 */
    if (off == 0) {
      r[0] = -((P + 1) / 2);
      r[1] = P / 2;
    } else {
      r[0] = -(P / 2);
      r[1] = (P + 1) / 2;
    }
  } else if (cent == OR_CC && intp == OR_ENO) { /* :784-803 */
    const int i0 = N / 2 - shift;
    r[0] = -i0;
    r[1] = P - i0;
  } else if (cent == OR_CC && intp == OR_MINMOD) { /* :849-856 */
    r[0] = -1;
    r[1] = 1;
  } else {
    r[0] = 1;
    r[1] = 0;
  }
}

#define CAT(o) c[(ptrdiff_t)(o) * st]

double or_interp1d(int cent, int intp, int P, int shift, int off,
                   const double *c, ptrdiff_t st) {
  const int N = P + 1;
  double y;
  if (cent == OR_VC && (intp == OR_POLY || intp == OR_HERMITE)) {
    /* :410-456 and :538-578: off=0 copies, off=1 uses symmetric pairs */
    if (off == 0)
      return CAT(0);
    const double *cs = vc_poly_tab(P);
    const int i0 = N / 2 - off;
    y = 0;
    for (int i = 0; i < N / 2; ++i)
      y += cs[i] * (CAT(i - i0) + CAT(P - i - i0));
    return y;
  }
  if (cent == OR_CC && intp == OR_POLY) { /* :477-511 */
    const double *cs = cc_poly_tab(P);
    const int i0 = N / 2;
    y = 0;
    if (off == 0)
      for (int i = 0; i < N; ++i)
        y += cs[i] * CAT(i - i0);
    else
      for (int i = 0; i < N; ++i)
        y += cs[P - i] * CAT(i - i0 + P % 2);
    return y;
  }
  if (cent == OR_VC && intp == OR_CONS) { /* :602-630 */
    y = 0;
    if (off == 0) {
      const int i0 = P / 2;
      /* coeffs0 is {1} for both instantiated orders 0 and 1 */
      for (int i = 0; i < P / 2 * 2 + 1; ++i)
        y += VC_CONS0_0[i] * CAT(i - i0);
    } else {
      const int i0 = (P + 1) / 2 - 1;
      /* coeffs1 is empty for order 0 (:225), {1/2,1/2} for order 1 */
      for (int i = 0; i < (P + 1) / 2 * 2; ++i)
        y += VC_CONS1_1[i] * CAT(i - i0);
    }
    return y;
  }
  if (cent == OR_CC && intp == OR_CONS) { /* :651-715 */
    const double *cs = cc_cons_tab(P);
    const int i0 = N / 2;
    if (P % 2 == 0) {
      y = cs[P / 2] * CAT(P / 2 - i0);
      if (off == 0)
        for (int i = 0; i < P / 2; ++i)
          y += cs[i] * (CAT(i - i0) - CAT(P - i - i0));
      else
        for (int i = 0; i < P / 2; ++i)
          y += cs[i] * (CAT(P - i - i0) - CAT(i - i0));
    } else {
      y = 0;
      if (off == 0)
        for (int i = 0; i < N; ++i)
          y += cs[i] * CAT(i - i0);
      else
        for (int i = 0; i < N; ++i)
          y += cs[i] * CAT((P + 1) / 2 - i);
    }
    return y;
  }
  if (cent == OR_CC && intp == OR_ENO) { /* :804-841 (P == 2 only) */
    const int m = off == 0 ? shift + P / 2 : P / 2 - shift;
    const double *cs = CC_ENO2[m];
    const int i0 = N / 2 - shift;
    y = 0;
    if (off == 0)
      for (int i = 0; i <= P; ++i)
        y += cs[i] * CAT(i - i0);
    else
      for (int i = 0; i <= P; ++i)
        y += cs[P - i] * CAT(i - i0);
    return y;
  }
  if (cent == OR_CC && intp == OR_MINMOD) { /* :857-877 */
    const double cval = CAT(0);
    const double cval_minus = CAT(-1);
    const double cval_plus = CAT(1);
    const double delta_minus = cval - cval_minus;
    const double delta_plus = cval_plus - cval;
    const double delta =
        (copysign(0.125, delta_minus) + copysign(0.125, delta_plus)) *
        fmin(fabs(delta_minus), fabs(delta_plus));
    const double sign = off == 0 ? -1 : +1;
    return cval + sign * delta;
  }
  return NAN;
}

/* Undivided difference for CC ENO (:750-773): binomial weights. */
static double undivided_difference(int P, const double *c, ptrdiff_t st) {
  /* weights (-1)^i C(P,i), symmetric for even P */
  int ws[8];
  ws[0] = 1;
  for (int i = 1; i <= P; ++i)
    ws[i] = -ws[i - 1] * (P - i + 1) / i;
  const int i0 = P / 2;
  double dd = ws[P / 2] * CAT(0);
  for (int i = 0; i < P / 2; ++i)
    dd += ws[i] * (CAT(i - i0) + CAT(P - i - i0));
  return dd;
}

#undef CAT

/* ------------------------------------------------------------------------ */
/* Operator selection (driver.cxx:593-642; impl_*.cxx instantiations)        */

static void set_dir(or_op *op, int d, int cent, int intp, int order) {
  op->cent[d] = cent;
  op->intp[d] = intp;
  op->order[d] = order;
}

int or_lookup(const char *type, int order, const int indextype[3],
              or_op *out) {
  or_op op;
  memset(&op, 0, sizeof op);
  op.fb = OR_FB_NONE;
  enum { T_POLY, T_CONS, T_DDF, T_ENO, T_MINMOD, T_HERM, T_NAT, T_PC3, T_PE3 };
  int t;
  if (!strcmp(type, "interpolate"))
    t = T_POLY;
  else if (!strcmp(type, "conservative"))
    t = T_CONS;
  else if (!strcmp(type, "ddf"))
    t = T_DDF;
  else if (!strcmp(type, "eno"))
    t = T_ENO;
  else if (!strcmp(type, "minmod"))
    t = T_MINMOD;
  else if (!strcmp(type, "hermite"))
    t = T_HERM;
  else if (!strcmp(type, "natural"))
    t = T_NAT;
  else if (!strcmp(type, "poly-cons3lfb"))
    t = T_PC3;
  else if (!strcmp(type, "poly-eno3lfb"))
    t = T_PE3;
  else
    return -1;
  int ok;
  switch (t) {
  case T_POLY: case T_DDF: case T_NAT:
    ok = order == 1 || order == 3 || order == 5 || order == 7;
    break;
  case T_CONS:
    ok = order == 0 || order == 1;
    break;
  case T_ENO:
    ok = order == 3 || order == 5;
    break;
  default: /* minmod, hermite, poly-cons3lfb, poly-eno3lfb */
    ok = order == 1 || order == 3 || order == 5;
  }
  if (!ok)
    return -2;
  for (int d = 0; d < 3; ++d) {
    const int cc = indextype[d] != 0;
    const int cent = cc ? OR_CC : OR_VC;
    switch (t) {
    case T_POLY: /* impl_poly.cxx */
      set_dir(&op, d, cent, OR_POLY, order);
      break;
    case T_CONS: /* impl_cons.cxx: o0 CONS 0; o1 VC CONS 1, CC CONS 2 */
      set_dir(&op, d, cent, OR_CONS, order == 0 ? 0 : (cc ? 2 : 1));
      break;
    case T_DDF: /* impl_ddf.cxx */
      if (cc)
        set_dir(&op, d, cent, OR_CONS, order - 1);
      else
        set_dir(&op, d, cent, OR_POLY, order);
      break;
    case T_ENO: /* impl_eno.cxx */
      if (cc)
        set_dir(&op, d, cent, OR_ENO, 2);
      else
        set_dir(&op, d, cent, OR_POLY, order);
      break;
    case T_MINMOD: /* impl_minmod.cxx */
      if (cc)
        set_dir(&op, d, cent, OR_MINMOD, 1);
      else
        set_dir(&op, d, cent, OR_POLY, order);
      break;
    case T_HERM: /* impl_hermite.cxx */
      if (cc)
        set_dir(&op, d, cent, OR_CONS, order);
      else
        set_dir(&op, d, cent, OR_HERMITE, order);
      break;
    case T_NAT: /* impl_natural.cxx */
      if (cc)
        set_dir(&op, d, cent, OR_CONS, order);
      else
        set_dir(&op, d, cent, OR_POLY, order);
      break;
    case T_PC3: /* impl_poly_cons3lfb.cxx */
      if (cc)
        set_dir(&op, d, cent, OR_CONS, order < 3 ? order : 3);
      else
        set_dir(&op, d, cent, OR_POLY, order);
      op.fb = OR_FB_LINEAR;
      break;
    case T_PE3: /* impl_poly_eno3lfb.cxx */
      if (cc) {
        if (order == 1)
          set_dir(&op, d, cent, OR_CONS, 0);
        else
          set_dir(&op, d, cent, OR_ENO, 2);
      } else {
        set_dir(&op, d, cent, OR_POLY, order);
      }
      op.fb = OR_FB_LINEAR;
      break;
    }
  }
  *out = op;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Boxes                                                                     */

static int fdiv2(int i) { return (i - (i & 1)) / 2; } /* floor(i/2) == i>>1 */

void or_coarse_box(const or_op *op, const int flo[3], const int fhi[3],
                   int clo[3], int chi[3]) {
  for (int d = 0; d < 3; ++d) {
    /* amrex::coarsen: floor; nodal hi rounds up when not divisible */
    int lo = fdiv2(flo[d]);
    int hi = fdiv2(fhi[d]);
    if (op->cent[d] == OR_VC && (fhi[d] & 1))
      hi += 1;
    const int g = or_required_ghosts(op->cent[d], op->intp[d], op->order[d]);
    clo[d] = lo - g;
    chi[d] = hi + g;
  }
}

typedef struct {
  const double *p;
  int lo[3];
  int64_t sy, sz;
} view;

static inline ptrdiff_t vidx(const view *v, int i, int j, int k) {
  return (ptrdiff_t)(i - v->lo[0]) + (ptrdiff_t)(j - v->lo[1]) * v->sy +
         (ptrdiff_t)(k - v->lo[2]) * v->sz;
}

static view make_view(const or_fab *f, int comp) {
  view v;
  v.p = f->data + (ptrdiff_t)comp * f->comp_stride;
  for (int d = 0; d < 3; ++d)
    v.lo[d] = f->lo[d];
  v.sy = (int64_t)(f->hi[0] - f->lo[0] + 1);
  v.sz = v.sy * (int64_t)(f->hi[1] - f->lo[1] + 1);
  return v;
}

/* ------------------------------------------------------------------------ */
/* Tensor product (call_stencil_3d, :1206-1216): z outer, y, x innermost     */

typedef struct {
  int cent[3], intp[3], order[3];
} dirs3;

static double tensor(const dirs3 *s, const int shift[3], const int off[3],
                     const view *c, const int ic[3]) {
  int rx[2], ry[2], rz[2];
  or_stencil_radius(s->cent[1], s->intp[1], s->order[1], shift[1], off[1], ry);
  or_stencil_radius(s->cent[2], s->intp[2], s->order[2], shift[2], off[2], rz);
  (void)rx;
  double ybuf[32], xbuf[32];
  for (int dk = rz[0]; dk <= rz[1]; ++dk) {
    for (int dj = ry[0]; dj <= ry[1]; ++dj) {
      const double *p = c->p + vidx(c, ic[0], ic[1] + dj, ic[2] + dk);
      xbuf[16 + dj] = or_interp1d(s->cent[0], s->intp[0], s->order[0],
                                  shift[0], off[0], p, 1);
    }
    ybuf[16 + dk] = or_interp1d(s->cent[1], s->intp[1], s->order[1], shift[1],
                                off[1], &xbuf[16], 1);
  }
  return or_interp1d(s->cent[2], s->intp[2], s->order[2], shift[2], off[2],
                     &ybuf[16], 1);
}

/* ENO shift search (:1417-1517 per-var; :1855-1956 per-group over comps) */
static void eno_search(const dirs3 *s, const int off[3], const view *cs,
                       int ncomp, const int ic[3], int shift[3]) {
  int use_shift[3], maxshift[3];
  for (int d = 0; d < 3; ++d) {
    use_shift[d] = s->intp[d] == OR_ENO && s->order[d] > 0;
    maxshift[d] =
        use_shift[d]
            ? or_required_ghosts(s->cent[d], s->intp[d], s->order[d]) / 2
            : 0;
  }
  shift[0] = shift[1] = shift[2] = 0;
  double min_dd = INFINITY;
  for (int sk = -maxshift[2]; sk <= maxshift[2]; ++sk)
    for (int sj = -maxshift[1]; sj <= maxshift[1]; ++sj)
      for (int si = -maxshift[0]; si <= maxshift[0]; ++si) {
        const int sv[3] = {si, sj, sk};
        int r[3][2];
        for (int d = 0; d < 3; ++d) {
          or_stencil_radius(s->cent[d], s->intp[d], s->order[d], sv[d], off[d],
                            r[d]);
          if (use_shift[d]) {
            r[d][0] -= sv[d];
            r[d][1] -= sv[d];
          }
        }
        double ddx = 0, ddy = 0, ddz = 0;
        for (int comp = 0; comp < ncomp; ++comp) {
          const view *c = &cs[comp];
          if (use_shift[0])
            for (int dk = r[2][0]; dk <= r[2][1]; ++dk)
              for (int dj = r[1][0]; dj <= r[1][1]; ++dj) {
                const double *p =
                    c->p + vidx(c, ic[0] + si, ic[1] + sj + dj, ic[2] + sk + dk);
                ddx = fmax(ddx, fabs(undivided_difference(s->order[0], p, 1)));
              }
          if (use_shift[1])
            for (int dk = r[2][0]; dk <= r[2][1]; ++dk)
              for (int di = r[0][0]; di <= r[0][1]; ++di) {
                const double *p =
                    c->p + vidx(c, ic[0] + si + di, ic[1] + sj, ic[2] + sk + dk);
                ddy = fmax(ddy,
                           fabs(undivided_difference(s->order[1], p, c->sy)));
              }
          if (use_shift[2])
            for (int dj = r[1][0]; dj <= r[1][1]; ++dj)
              for (int di = r[0][0]; di <= r[0][1]; ++di) {
                const double *p =
                    c->p + vidx(c, ic[0] + si + di, ic[1] + sj + dj, ic[2] + sk);
                ddz = fmax(ddz,
                           fabs(undivided_difference(s->order[2], p, c->sz)));
              }
        }
        const double penalty =
            1 + sqrt(DBL_EPSILON) * (abs(si) + abs(sj) + abs(sk));
        const double dd = penalty * fmax(fmax(ddx, ddy), ddz);
        if (dd < min_dd) {
          min_dd = dd;
          shift[0] = si;
          shift[1] = sj;
          shift[2] = sk;
        }
      }
}

/* Slope sign-change test along direction d (:1609-1664 / :2026-2116). */
static int slope_changes(int d, const int sr[3][2], const view *c,
                         const int ic[3]) {
  int a = (d + 1) % 3, b = (d + 2) % 3;
  /* the reference loops the two other directions in (z, y) / (z, x) /
   * (y, x) order; the OR-reduction is order independent */
  if (a > b) {
    int t = a;
    a = b;
    b = t;
  }
  int need = 0;
  for (int tb = sr[b][0]; tb <= sr[b][1]; ++tb)
    for (int ta = sr[a][0]; ta <= sr[a][1]; ++ta) {
      int p0[3], p1[3];
      p0[a] = p1[a] = ic[a] + ta;
      p0[b] = p1[b] = ic[b] + tb;
      p0[d] = ic[d] + 0;
      p1[d] = ic[d] + 1;
      const double s0 = c->p[vidx(c, p1[0], p1[1], p1[2])] -
                        c->p[vidx(c, p0[0], p0[1], p0[2])];
      for (int t = sr[d][0] + 2; t <= sr[d][1]; ++t) {
        p0[d] = ic[d] + t - 1;
        p1[d] = ic[d] + t;
        const double sv = c->p[vidx(c, p1[0], p1[1], p1[2])] -
                          c->p[vidx(c, p0[0], p0[1], p0[2])];
        need |= sv * s0 < 0;
      }
    }
  return need;
}

static void minmax_box(const int sr[3][2], const view *c, const int ic[3],
                       double *minval, double *maxval) {
  for (int dk = sr[2][0]; dk <= sr[2][1]; ++dk)
    for (int dj = sr[1][0]; dj <= sr[1][1]; ++dj)
      for (int di = sr[0][0]; di <= sr[0][1]; ++di) {
        const double v = c->p[vidx(c, ic[0] + di, ic[1] + dj, ic[2] + dk)];
        *minval = fmin(*minval, v);
        *maxval = fmax(*maxval, v);
      }
}

/* Checks that every coarse point the operator may touch lies in crse. */
static int check_boxes(const or_op *op, const or_fab *crse, const or_fab *fine,
                       const int rlo[3], const int rhi[3]) {
  int clo[3], chi[3];
  for (int d = 0; d < 3; ++d) {
    if (rlo[d] > rhi[d])
      return -3;
    if (rlo[d] < fine->lo[d] || rhi[d] > fine->hi[d])
      return -4;
  }
  or_coarse_box(op, rlo, rhi, clo, chi);
  for (int d = 0; d < 3; ++d)
    if (clo[d] < crse->lo[d] || chi[d] > crse->hi[d])
      return -5;
  return 0;
}

static int fallback_active_per_var(const or_op *op) {
  if (op->fb == OR_FB_NONE)
    return 0;
  for (int d = 0; d < 3; ++d)
    if ((op->intp[d] == OR_CONS || op->intp[d] == OR_ENO) && op->order[d] > 1)
      return 1;
  return 0;
}

static int fallback_active_per_group(const or_op *op) {
  if (op->fb == OR_FB_NONE)
    return 0;
  for (int d = 0; d < 3; ++d)
    if (op->intp[d] == OR_CONS && op->order[d] > 1)
      return 1;
  return 0;
}

static void per_var_point(const or_op *op, const view *c, const int f[3],
                          double *out) {
  dirs3 s;
  int ic[3], off[3], shift[3] = {0, 0, 0};
  int any_shift = 0;
  for (int d = 0; d < 3; ++d) {
    s.cent[d] = op->cent[d];
    s.intp[d] = op->intp[d];
    s.order[d] = op->order[d];
    ic[d] = fdiv2(f[d]);
    off[d] = f[d] & 1;
    any_shift |= op->intp[d] == OR_ENO && op->order[d] > 0;
  }
  if (any_shift)
    eno_search(&s, off, c, 1, ic, shift);
  const double val = tensor(&s, shift, off, c, ic);
  if (!fallback_active_per_var(op)) {
    *out = val;
    return;
  }
  /* :1547-1574: linear value, CONS 1 in CONS/ENO directions */
  dirs3 sl = s;
  for (int d = 0; d < 3; ++d)
    if (s.intp[d] == OR_CONS || s.intp[d] == OR_ENO) {
      sl.intp[d] = OR_CONS;
      sl.order[d] = 1;
    }
  const int zero[3] = {0, 0, 0};
  const double val_lin = tensor(&sl, zero, off, c, ic);
  int sr[3][2];
  for (int d = 0; d < 3; ++d)
    or_stencil_radius(s.cent[d], s.intp[d], s.order[d], shift[d], off[d],
                      sr[d]);
  double minval = INFINITY, maxval = -INFINITY;
  minmax_box(sr, c, ic, &minval, &maxval);
  int need = val < minval || val > maxval;
  for (int d = 0; d < 3; ++d)
    if (s.intp[d] == OR_CONS || s.intp[d] == OR_ENO)
      need |= slope_changes(d, sr, c, ic);
  *out = need ? val_lin : val;
}

static void per_group_point(const or_op *op, const view *cs, int ncomp,
                            const int f[3], double *out /* [ncomp] */) {
  dirs3 s;
  int ic[3], off[3], shift[3] = {0, 0, 0};
  int any_shift = 0;
  for (int d = 0; d < 3; ++d) {
    s.cent[d] = op->cent[d];
    s.intp[d] = op->intp[d];
    s.order[d] = op->order[d];
    ic[d] = fdiv2(f[d]);
    off[d] = f[d] & 1;
    any_shift |= op->intp[d] == OR_ENO && op->order[d] > 0;
  }
  if (any_shift)
    eno_search(&s, off, cs, ncomp, ic, shift);
  double vals[16];
  for (int comp = 0; comp < ncomp; ++comp)
    vals[comp] = tensor(&s, shift, off, &cs[comp], ic);
  if (!fallback_active_per_group(op)) {
    for (int comp = 0; comp < ncomp; ++comp)
      out[comp] = vals[comp];
    return;
  }
  int sr[3][2];
  for (int d = 0; d < 3; ++d)
    or_stencil_radius(s.cent[d], s.intp[d], s.order[d], shift[d], off[d],
                      sr[d]);
  int need = 0;
  /* :2010-2024: min/max accumulate across components */
  double minval = INFINITY, maxval = -INFINITY;
  for (int comp = 0; comp < ncomp; ++comp) {
    minmax_box(sr, &cs[comp], ic, &minval, &maxval);
    need |= vals[comp] < minval || vals[comp] > maxval;
  }
  for (int d = 0; d < 3; ++d)
    if (s.intp[d] == OR_CONS)
      for (int comp = 0; comp < ncomp; ++comp)
        need |= slope_changes(d, sr, &cs[comp], ic);
  /* :2119-2137: same interpolation type, order 1 in CONS directions */
  dirs3 sl = s;
  for (int d = 0; d < 3; ++d)
    if (s.intp[d] == OR_CONS)
      sl.order[d] = 1;
  const int zero[3] = {0, 0, 0};
  for (int comp = 0; comp < ncomp; ++comp)
    out[comp] = need ? tensor(&sl, zero, off, &cs[comp], ic) : vals[comp];
}

int or_prolongate(const or_op *op, const or_fab *crse, int crse_comp,
                  const or_fab *fine, int fine_comp, int ncomp,
                  const int rlo[3], const int rhi[3], int per_group) {
  int err = check_boxes(op, crse, fine, rlo, rhi);
  if (err)
    return err;
  if (ncomp < 1)
    return -6;
  if (per_group && ncomp > 16)
    return -7; /* :1707-1708 */
  if (!per_group) {
    for (int comp = 0; comp < ncomp; ++comp) {
      const view c = make_view(crse, crse_comp + comp);
      view fv = make_view(fine, fine_comp + comp);
      double *fp = (double *)fv.p;
      for (int k = rlo[2]; k <= rhi[2]; ++k)
        for (int j = rlo[1]; j <= rhi[1]; ++j)
          for (int i = rlo[0]; i <= rhi[0]; ++i) {
            const int f[3] = {i, j, k};
            per_var_point(op, &c, f, &fp[vidx(&fv, i, j, k)]);
          }
    }
  } else {
    view cs[16], fs[16];
    for (int comp = 0; comp < ncomp; ++comp) {
      cs[comp] = make_view(crse, crse_comp + comp);
      fs[comp] = make_view(fine, fine_comp + comp);
    }
    for (int k = rlo[2]; k <= rhi[2]; ++k)
      for (int j = rlo[1]; j <= rhi[1]; ++j)
        for (int i = rlo[0]; i <= rhi[0]; ++i) {
          const int f[3] = {i, j, k};
          double out[16];
          per_group_point(op, cs, ncomp, f, out);
          for (int comp = 0; comp < ncomp; ++comp)
            ((double *)fs[comp].p)[vidx(&fs[comp], i, j, k)] = out[comp];
        }
  }
  return 0;
}

int or_prolongate_many(int nitems, const or_item *items, int per_group) {
  int status = 0;
#pragma omp parallel for schedule(dynamic, 1)
  for (int n = 0; n < nitems; ++n) {
    const or_item *it = &items[n];
    const int e = or_prolongate(it->op, &it->crse, it->crse_comp, &it->fine,
                                it->fine_comp, it->ncomp, it->rlo, it->rhi,
                                per_group);
    if (e) {
#pragma omp critical
      if (!status)
        status = e;
    }
  }
  return status;
}

/* ------------------------------------------------------------------------ */
/* Restriction (schedule.cxx:2791-2816 -> AMReX average_down family):       */
/* per direction VC = injection of fine 2i, CC = mean of fine 2i, 2i+1.     */

int or_restrict(const int indextype[3], const or_fab *fine, int fine_comp,
                const or_fab *crse, int crse_comp, int ncomp, const int rlo[3],
                const int rhi[3]) {
  int n[3], ncc = 0;
  for (int d = 0; d < 3; ++d) {
    n[d] = indextype[d] ? 2 : 1;
    ncc += indextype[d] != 0;
    if (rlo[d] > rhi[d])
      return -3;
    if (rlo[d] < crse->lo[d] || rhi[d] > crse->hi[d])
      return -4;
    if (2 * rlo[d] < fine->lo[d] || 2 * rhi[d] + n[d] - 1 > fine->hi[d])
      return -5;
  }
  const double scale = 1.0 / (double)(1 << ncc);
  for (int comp = 0; comp < ncomp; ++comp) {
    const view fv = make_view(fine, fine_comp + comp);
    view cv = make_view(crse, crse_comp + comp);
    double *cp = (double *)cv.p;
    for (int k = rlo[2]; k <= rhi[2]; ++k)
      for (int j = rlo[1]; j <= rhi[1]; ++j)
        for (int i = rlo[0]; i <= rhi[0]; ++i) {
          double v;
          if (ncc == 0) {
            v = fv.p[vidx(&fv, 2 * i, 2 * j, 2 * k)];
          } else {
            double s = 0;
            for (int kk = 0; kk < n[2]; ++kk)
              for (int jj = 0; jj < n[1]; ++jj)
                for (int ii = 0; ii < n[0]; ++ii)
                  s += fv.p[vidx(&fv, 2 * i + ii, 2 * j + jj, 2 * k + kk)];
            v = s * scale;
          }
          cp[vidx(&cv, i, j, k)] = v;
        }
  }
  return 0;
}

int or_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void or_set_num_threads(int n) {
#ifdef _OPENMP
  if (n > 0)
    omp_set_num_threads(n);
#else
  (void)n;
#endif
}
