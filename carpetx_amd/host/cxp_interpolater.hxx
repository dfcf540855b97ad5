// The amrex::Interpolater adapter over libcxp.so, for CarpetX's driver.
//
// It replaces the prolongate_3d_rf2<...> operator classes
// (CarpetX/src/prolongate_3d_rf2.hxx:33-121) with their static singletons
// (prolongate_3d_rf2_impl_*.cxx), and get_interpolator
// (CarpetX/src/driver.cxx:558-642).  It is written against AMReX's and the
// Cactus flesh's headers as CarpetX includes them; tests/amrex_shim/ holds
// the few declarations it uses so that it compiles and runs in this
// repository (tests/amrex_shim/interpolater_check.cpp).
#pragma once

#include <AMReX_Interpolater.H>
#include <cctk_Parameters.h>

#include <array>
#include <map>
#include <memory>
#include <mutex>
#include <string>

#include "cxp.h"

namespace CarpetX {

class CxpInterpolater final : public amrex::Interpolater {
  cxp_op op_;

  static void check(int status, const char *what) {
    if (status != CXP_OK)
      amrex::Abort(std::string(what) + ": " + cxp_status_string(status));
  }
  static cxp_fab fab(const amrex::FArrayBox &f) {
    cxp_fab r;
    r.data = const_cast<double *>(f.dataPtr());
    const amrex::Box &b = f.box();
    for (int d = 0; d < 3; ++d) {
      r.lo[d] = b.smallEnd(d);
      r.hi[d] = b.bigEnd(d);
    }
    r.comp_stride = b.numPts();  // FArrayBox: components are box-volume apart
    r.ncomp = f.nComp();
    return r;
  }

public:
  explicit CxpInterpolater(cxp_op op) : op_(op) {}
  CxpInterpolater(const char *type, int order, const int indextype[3]) {
    check(cxp_op_lookup(type, order, indextype, &op_), "cxp_op_lookup");
  }
  cxp_op op() const { return op_; }

  amrex::Box CoarseBox(const amrex::Box &fine, int ratio) override {
    if (ratio != 2)
      amrex::Abort("CxpInterpolater: ratio must be 2");  // impl.hxx:1229
    int lo[3], hi[3];
    check(cxp_coarse_box(op_, fine.smallEnd().getVect(),
                         fine.bigEnd().getVect(), lo, hi),
          "cxp_coarse_box");
    return amrex::Box(amrex::IntVect(lo), amrex::IntVect(hi), fine.ixType());
  }
  amrex::Box CoarseBox(const amrex::Box &fine,
                       const amrex::IntVect &ratio) override {
    if (ratio != amrex::IntVect(2))
      amrex::Abort("CxpInterpolater: ratio must be 2");
    return CoarseBox(fine, 2);
  }

  void interp(const amrex::FArrayBox &crse, int crse_comp,
              amrex::FArrayBox &fine, int fine_comp, int ncomp,
              const amrex::Box &fine_region, const amrex::IntVect &ratio,
              const amrex::Geometry &, const amrex::Geometry &,
              amrex::Vector<amrex::BCRec> const &, int, int,
              amrex::RunOn) override {
    DECLARE_CCTK_PARAMETERS;  // prolongate_per_group, impl.hxx:2179-2187
    if (ratio != amrex::IntVect(2))
      amrex::Abort("CxpInterpolater: ratio must be 2");
    const cxp_fab c = fab(crse), f = fab(fine);
    check(cxp_prolongate(op_, &c, crse_comp, &f, fine_comp, ncomp,
                         fine_region.smallEnd().getVect(),
                         fine_region.bigEnd().getVect(),
                         prolongate_per_group ? 1 : 0,
                         amrex::Gpu::gpuStream()),
          "cxp_prolongate");
    // asynchronous, like the reference's ParallelFor (impl.hxx:1685-1688)
  }

  void interp_face(const amrex::FArrayBox &crse, int crse_comp,
                   amrex::FArrayBox &fine, int fine_comp, int ncomp,
                   const amrex::Box &fine_region, const amrex::IntVect &ratio,
                   const amrex::IArrayBox &, const amrex::Geometry &cg,
                   const amrex::Geometry &fg,
                   amrex::Vector<amrex::BCRec> const &bcr, int,
                   amrex::RunOn r) override {
    // the reference ignores solve_mask and forwards (impl.hxx:2204-2207)
    interp(crse, crse_comp, fine, fine_comp, ncomp, fine_region, ratio, cg,
           fg, bcr, 0, 0, r);
  }
};

// One adapter per operator handle, never freed (the reference's
// singletons); handles are unique per (type, order, indextype).
inline amrex::Interpolater *cxp_interpolater_for(cxp_op op) {
  static std::mutex mu;
  static std::map<cxp_op, std::unique_ptr<CxpInterpolater>> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto &p = cache[op];
  if (!p)
    p = std::make_unique<CxpInterpolater>(op);
  return p.get();
}

// get_interpolator(prolongation_type, prolongation_order, indextype)
// (driver.cxx:593-626): aborts on an unknown type or order, as the
// reference's CCTK_VERROR does.
inline amrex::Interpolater *
get_interpolator(const std::string &prolongation_type,
                 const int prolongation_order,
                 const std::array<int, 3> &indextype) {
  cxp_op op;
  const int st = cxp_op_lookup(prolongation_type.c_str(), prolongation_order,
                               indextype.data(), &op);
  if (st != CXP_OK)
    amrex::Abort(std::string("get_interpolator: ") + cxp_status_string(st));
  return cxp_interpolater_for(op);
}

// get_interpolator(groupdata) (driver.cxx:558-591, 628-642): the group's
// tags may set prolongation_type and prolongation_order; the driver
// parameters fill in the rest.
inline amrex::Interpolater *
get_group_interpolator(const char *group_tags,
                       const std::array<int, 3> &indextype) {
  DECLARE_CCTK_PARAMETERS;
  cxp_op op;
  const int st = cxp_group_op_lookup(group_tags, prolongation_type,
                                     prolongation_order, indextype.data(),
                                     &op);
  if (st != CXP_OK)
    amrex::Abort(std::string("get_interpolator: ") + cxp_status_string(st));
  return cxp_interpolater_for(op);
}

} // namespace CarpetX
