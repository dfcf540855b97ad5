// 1-D rf=2 prolongation operators for gfx950 kernels and host code.
//
// Every operator is identified by a stencil id (Sid).  The arithmetic of each
// one reproduces the summation form of the corresponding reference functor
// (CarpetX/src/prolongate_3d_rf2_impl.hxx:388-878) term by term:
//   VC POLY   :410-456  off 0 copies, off 1 sums symmetric pairs
//   CC POLY   :477-511  plain ordered sum, mirrored for off 1
//   VC CONS   :602-630  plain ordered sum (deprecated operator)
//   CC CONS   :651-715  even order: centre + antisymmetric pairs; odd: sum
//   CC ENO 2  :804-841  shifted conservative row
//   CC MINMOD :857-877  minmod slope / 4
// so that a separable x -> y -> z evaluation, compiled without FMA
// contraction, is bit-identical to the reference's nested tensor product
// (call_stencil_3d, :1206-1216, evaluates x innermost for each (dj, dk)).
#pragma once

#include <hip/hip_runtime.h>

#include "cxp_coeffs.inc"

namespace cxp {

enum Sid : int {
  SID_VPOLY1 = 0,
  SID_VPOLY3,
  SID_VPOLY5,
  SID_VPOLY7,
  SID_CPOLY0,
  SID_CPOLY1,
  SID_CPOLY2,
  SID_CPOLY3,
  SID_CPOLY4,
  SID_CPOLY5,
  SID_CPOLY6,
  SID_CPOLY7,
  SID_VCONS0,
  SID_VCONS1,
  SID_CCONS0,
  SID_CCONS1,
  SID_CCONS2,
  SID_CCONS3,
  SID_CCONS4,
  SID_CCONS5,
  SID_CCONS6,
  SID_CCONS7,
  SID_MINMOD,
  SID_ENO2,
  SID_COUNT
};

// centring / interpolation codes (match include/cxp.h)
enum { CENT_VC = 0, CENT_CC = 1 };
enum { INTP_POLY = 0, INTP_HERMITE = 1, INTP_CONS = 2, INTP_ENO = 3, INTP_MINMOD = 4 };

struct SidInfo {
  int cent, intp, order;
  int ghosts;     // required_ghosts
  int lo[2], hi[2];  // stencil radius for off 0 / off 1 (shift 0)
};

// Radii per the reference's stencil_radius() (:401-408, :464-476, :587-601,
// :638-650, :784-803, :849-856) and required_ghosts (:398 ... :847).
__host__ __device__ constexpr SidInfo sid_info(int sid) {
  switch (sid) {
  case SID_VPOLY1: return {CENT_VC, INTP_POLY, 1, 1, {0, 0}, {0, 1}};
  case SID_VPOLY3: return {CENT_VC, INTP_POLY, 3, 2, {0, -1}, {0, 2}};
  case SID_VPOLY5: return {CENT_VC, INTP_POLY, 5, 3, {0, -2}, {0, 3}};
  case SID_VPOLY7: return {CENT_VC, INTP_POLY, 7, 4, {0, -3}, {0, 4}};
  case SID_CPOLY0: return {CENT_CC, INTP_POLY, 0, 0, {0, 0}, {0, 0}};
  case SID_CPOLY1: return {CENT_CC, INTP_POLY, 1, 1, {-1, 0}, {0, 1}};
  case SID_CPOLY2: return {CENT_CC, INTP_POLY, 2, 1, {-1, -1}, {1, 1}};
  case SID_CPOLY3: return {CENT_CC, INTP_POLY, 3, 2, {-2, -1}, {1, 2}};
  case SID_CPOLY4: return {CENT_CC, INTP_POLY, 4, 2, {-2, -2}, {2, 2}};
  case SID_CPOLY5: return {CENT_CC, INTP_POLY, 5, 3, {-3, -2}, {2, 3}};
  case SID_CPOLY6: return {CENT_CC, INTP_POLY, 6, 3, {-3, -3}, {3, 3}};
  case SID_CPOLY7: return {CENT_CC, INTP_POLY, 7, 4, {-4, -3}, {3, 4}};
  case SID_VCONS0: return {CENT_VC, INTP_CONS, 0, 0, {0, 1}, {0, 0}};
  case SID_VCONS1: return {CENT_VC, INTP_CONS, 1, 1, {0, 0}, {0, 1}};
  case SID_CCONS0: return {CENT_CC, INTP_CONS, 0, 0, {0, 0}, {0, 0}};
  case SID_CCONS1: return {CENT_CC, INTP_CONS, 1, 1, {-1, 0}, {0, 1}};
  case SID_CCONS2: return {CENT_CC, INTP_CONS, 2, 1, {-1, -1}, {1, 1}};
  case SID_CCONS3: return {CENT_CC, INTP_CONS, 3, 2, {-2, -1}, {1, 2}};
  case SID_CCONS4: return {CENT_CC, INTP_CONS, 4, 2, {-2, -2}, {2, 2}};
  case SID_CCONS5: return {CENT_CC, INTP_CONS, 5, 3, {-3, -2}, {2, 3}};
  case SID_CCONS6: return {CENT_CC, INTP_CONS, 6, 3, {-3, -3}, {3, 3}};
  case SID_CCONS7: return {CENT_CC, INTP_CONS, 7, 4, {-4, -3}, {3, 4}};
  case SID_MINMOD: return {CENT_CC, INTP_MINMOD, 1, 1, {-1, -1}, {1, 1}};
  case SID_ENO2: return {CENT_CC, INTP_ENO, 2, 2, {-1, -1}, {1, 1}};
  }
  return {0, 0, 0, 0, {1, 1}, {0, 0}};
}

// Union window over both offsets (shift 0).  Empty ranges (VC CONS 0 at
// off 1 reads nothing) are skipped.
__host__ __device__ constexpr int sid_wlo(int sid) {
  const SidInfo s = sid_info(sid);
  int w = 1 << 20;
  for (int o = 0; o < 2; ++o)
    if (s.lo[o] <= s.hi[o] && s.lo[o] < w)
      w = s.lo[o];
  return w;
}
__host__ __device__ constexpr int sid_whi(int sid) {
  const SidInfo s = sid_info(sid);
  int w = -(1 << 20);
  for (int o = 0; o < 2; ++o)
    if (s.lo[o] <= s.hi[o] && s.hi[o] > w)
      w = s.hi[o];
  return w;
}

// ---------------------------------------------------------------------------
// Weight tables as constexpr functions (fold into immediates when unrolled)

template <int SID> struct W;
#define CXP_W(SID, ARR)                                                        \
  template <> struct W<SID> {                                                  \
    __host__ __device__ static constexpr double c(int i) { return ARR[i]; }    \
  };
CXP_W(SID_VPOLY1, kVC_POLY_1)
CXP_W(SID_VPOLY3, kVC_POLY_3)
CXP_W(SID_VPOLY5, kVC_POLY_5)
CXP_W(SID_VPOLY7, kVC_POLY_7)
CXP_W(SID_CPOLY0, kCC_POLY_0)
CXP_W(SID_CPOLY1, kCC_POLY_1)
CXP_W(SID_CPOLY2, kCC_POLY_2)
CXP_W(SID_CPOLY3, kCC_POLY_3)
CXP_W(SID_CPOLY4, kCC_POLY_4)
CXP_W(SID_CPOLY5, kCC_POLY_5)
CXP_W(SID_CPOLY6, kCC_POLY_6)
CXP_W(SID_CPOLY7, kCC_POLY_7)
CXP_W(SID_CCONS0, kCC_CONS_0)
CXP_W(SID_CCONS1, kCC_CONS_1)
CXP_W(SID_CCONS2, kCC_CONS_2)
CXP_W(SID_CCONS3, kCC_CONS_3)
CXP_W(SID_CCONS4, kCC_CONS_4)
CXP_W(SID_CCONS5, kCC_CONS_5)
CXP_W(SID_CCONS6, kCC_CONS_6)
CXP_W(SID_CCONS7, kCC_CONS_7)
#undef CXP_W

__host__ __device__ constexpr double eno2_w(int m, int i) {
  return kCC_ENO2_0[3 * m + i];
}

// ---------------------------------------------------------------------------
// The operators.  `c(o)` returns the coarse value at offset o from the
// coarse point/cell that owns the fine point; `off` is the fine offset.

template <int SID> struct Op1;

template <int SID, int P> struct VPolyOp {
  template <class F>
  __host__ __device__ __attribute__((always_inline)) static constexpr double
  apply(const F &c, int off) {
    if (off == 0)
      return c(0);
    constexpr int N = P + 1;
    constexpr int i0 = N / 2 - 1;
    double y = 0;
#pragma unroll
    for (int i = 0; i < N / 2; ++i)
      y += W<SID>::c(i) * (c(i - i0) + c(P - i - i0));
    return y;
  }
};

template <int SID, int P> struct CPolyOp {
  template <class F>
  __host__ __device__ __attribute__((always_inline)) static constexpr double
  apply(const F &c, int off) {
    constexpr int N = P + 1;
    constexpr int i0 = N / 2;
    double y = 0;
    if (off == 0) {
#pragma unroll
      for (int i = 0; i < N; ++i)
        y += W<SID>::c(i) * c(i - i0);
    } else {
#pragma unroll
      for (int i = 0; i < N; ++i)
        y += W<SID>::c(P - i) * c(i - i0 + P % 2);
    }
    return y;
  }
};

template <int SID, int P> struct CConsOp {
  template <class F>
  __host__ __device__ __attribute__((always_inline)) static constexpr double
  apply(const F &c, int off) {
    constexpr int N = P + 1;
    constexpr int i0 = N / 2;
    double y = 0;
    if constexpr (P % 2 == 0) {
      y = W<SID>::c(P / 2) * c(P / 2 - i0);
      if (off == 0) {
#pragma unroll
        for (int i = 0; i < P / 2; ++i)
          y += W<SID>::c(i) * (c(i - i0) - c(P - i - i0));
      } else {
#pragma unroll
        for (int i = 0; i < P / 2; ++i)
          y += W<SID>::c(i) * (c(P - i - i0) - c(i - i0));
      }
    } else {
      y = 0;
      if (off == 0) {
#pragma unroll
        for (int i = 0; i < N; ++i)
          y += W<SID>::c(i) * c(i - i0);
      } else {
#pragma unroll
        for (int i = 0; i < N; ++i)
          y += W<SID>::c(i) * c((P + 1) / 2 - i);
      }
    }
    return y;
  }
};

template <> struct Op1<SID_VPOLY1> : VPolyOp<SID_VPOLY1, 1> {};
template <> struct Op1<SID_VPOLY3> : VPolyOp<SID_VPOLY3, 3> {};
template <> struct Op1<SID_VPOLY5> : VPolyOp<SID_VPOLY5, 5> {};
template <> struct Op1<SID_VPOLY7> : VPolyOp<SID_VPOLY7, 7> {};
template <> struct Op1<SID_CPOLY0> : CPolyOp<SID_CPOLY0, 0> {};
template <> struct Op1<SID_CPOLY1> : CPolyOp<SID_CPOLY1, 1> {};
template <> struct Op1<SID_CPOLY2> : CPolyOp<SID_CPOLY2, 2> {};
template <> struct Op1<SID_CPOLY3> : CPolyOp<SID_CPOLY3, 3> {};
template <> struct Op1<SID_CPOLY4> : CPolyOp<SID_CPOLY4, 4> {};
template <> struct Op1<SID_CPOLY5> : CPolyOp<SID_CPOLY5, 5> {};
template <> struct Op1<SID_CPOLY6> : CPolyOp<SID_CPOLY6, 6> {};
template <> struct Op1<SID_CPOLY7> : CPolyOp<SID_CPOLY7, 7> {};
template <> struct Op1<SID_CCONS0> : CConsOp<SID_CCONS0, 0> {};
template <> struct Op1<SID_CCONS1> : CConsOp<SID_CCONS1, 1> {};
template <> struct Op1<SID_CCONS2> : CConsOp<SID_CCONS2, 2> {};
template <> struct Op1<SID_CCONS3> : CConsOp<SID_CCONS3, 3> {};
template <> struct Op1<SID_CCONS4> : CConsOp<SID_CCONS4, 4> {};
template <> struct Op1<SID_CCONS5> : CConsOp<SID_CCONS5, 5> {};
template <> struct Op1<SID_CCONS6> : CConsOp<SID_CCONS6, 6> {};
template <> struct Op1<SID_CCONS7> : CConsOp<SID_CCONS7, 7> {};

// VC CONS order 0: copy at off 0 (as 0 + 1*c), nothing (0) at off 1.
template <> struct Op1<SID_VCONS0> {
  template <class F>
  __host__ __device__ __attribute__((always_inline)) static constexpr double
  apply(const F &c, int off) {
    double y = 0;
    if (off == 0)
      y += 1.0 * c(0);
    return y;
  }
};
// VC CONS order 1: copy at off 0, midpoint average at off 1.
template <> struct Op1<SID_VCONS1> {
  template <class F>
  __host__ __device__ __attribute__((always_inline)) static constexpr double
  apply(const F &c, int off) {
    double y = 0;
    if (off == 0) {
      y += 1.0 * c(0);
    } else {
      y += 0.5 * c(0);
      y += 0.5 * c(1);
    }
    return y;
  }
};

__host__ __device__ __attribute__((always_inline)) inline double
minmod_apply(double cval_minus, double cval, double cval_plus, int off) {
  const double delta_minus = cval - cval_minus;
  const double delta_plus = cval_plus - cval;
  const double delta =
      (copysign(0.125, delta_minus) + copysign(0.125, delta_plus)) *
      fmin(fabs(delta_minus), fabs(delta_plus));
  const double sign = off == 0 ? -1.0 : 1.0;
  return cval + sign * delta;
}

template <> struct Op1<SID_MINMOD> {
  template <class F>
  __host__ __device__ __attribute__((always_inline)) static double
  apply(const F &c, int off) {
    return minmod_apply(c(-1), c(0), c(1), off);
  }
};

// CC ENO order 2 with shift s in [-1, 1]: row m, support [s-1, s+1].
template <class F>
__host__ __device__ __attribute__((always_inline)) inline double
eno2_apply(const F &c, int shift, int off) {
  const int m = off == 0 ? shift + 1 : 1 - shift;
  const int i0 = 1 - shift;
  double y = 0;
  if (off == 0) {
#pragma unroll
    for (int i = 0; i <= 2; ++i)
      y += eno2_w(m, i) * c(i - i0);
  } else {
#pragma unroll
    for (int i = 0; i <= 2; ++i)
      y += eno2_w(m, 2 - i) * c(i - i0);
  }
  return y;
}

template <> struct Op1<SID_ENO2> {
  template <class F>
  __host__ __device__ __attribute__((always_inline)) static double
  apply(const F &c, int off) {
    return eno2_apply(c, 0, off);
  }
};

// Runtime dispatch over every Sid (shift applies to ENO only).
template <class F>
__host__ __device__ inline double apply1d(int sid, const F &c, int shift,
                                          int off) {
  switch (sid) {
#define CXP_CASE(S)                                                            \
  case S:                                                                      \
    return Op1<S>::apply(c, off);
    CXP_CASE(SID_VPOLY1) CXP_CASE(SID_VPOLY3) CXP_CASE(SID_VPOLY5)
    CXP_CASE(SID_VPOLY7) CXP_CASE(SID_CPOLY0) CXP_CASE(SID_CPOLY1)
    CXP_CASE(SID_CPOLY2) CXP_CASE(SID_CPOLY3) CXP_CASE(SID_CPOLY4)
    CXP_CASE(SID_CPOLY5) CXP_CASE(SID_CPOLY6) CXP_CASE(SID_CPOLY7)
    CXP_CASE(SID_VCONS0) CXP_CASE(SID_VCONS1) CXP_CASE(SID_CCONS0)
    CXP_CASE(SID_CCONS1) CXP_CASE(SID_CCONS2) CXP_CASE(SID_CCONS3)
    CXP_CASE(SID_CCONS4) CXP_CASE(SID_CCONS5) CXP_CASE(SID_CCONS6)
    CXP_CASE(SID_CCONS7) CXP_CASE(SID_MINMOD)
#undef CXP_CASE
  case SID_ENO2:
    return eno2_apply(c, shift, off);
  }
  return 0.0;
}

// Stencil radius with shift (ENO moves its support with the shift).
__host__ __device__ inline void stencil_radius(int sid, int shift, int off,
                                               int &lo, int &hi) {
  const SidInfo s = sid_info(sid);
  lo = s.lo[off];
  hi = s.hi[off];
  if (sid == SID_ENO2) {
    lo += shift;
    hi += shift;
  }
}

// The linearly-interpolated operator used by the FB_LINEAR fallback.
// per-var (:1547-1574): CONS or ENO directions become CONS 1.
__host__ __device__ constexpr int sid_linear_per_var(int sid) {
  const SidInfo s = sid_info(sid);
  if (s.intp == INTP_CONS || s.intp == INTP_ENO)
    return s.cent == CENT_CC ? SID_CCONS1 : SID_VCONS1;
  return sid;
}
// per-group (:2119-2137): CONS directions become CONS 1, others unchanged.
__host__ __device__ constexpr int sid_linear_per_group(int sid) {
  const SidInfo s = sid_info(sid);
  if (s.intp == INTP_CONS)
    return s.cent == CENT_CC ? SID_CCONS1 : SID_VCONS1;
  return sid;
}

} // namespace cxp
