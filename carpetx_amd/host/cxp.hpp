// cxp.hpp -- C++17 host API over libcxp.so's C ABI (include/cxp.h).
//
// Mirrors the reference's operator interface for the rf=2 transfer path so
// C++ callers read like CarpetX:
//   * get_interpolator(type, order, indextype)   CarpetX/src/driver.cxx:593-642
//   * Prolongate3DRF2::CoarseBox / interp / interp_face
//                                  CarpetX/src/prolongate_3d_rf2.hxx:76-120
//   * restrict_fab                 Restrict(), CarpetX/src/schedule.cxx:2791-2816
//   * WorkList / Graph             the FillPatchInterp loop over fine patches
//                                  (CarpetX/src/fillpatch.cxx:116-119)
//   * boxinbox_setup / error_est   BoxInBox_Setup (BoxInBox/src/boxinbox.cxx:
//                                  102-174) and ErrorEst's threshold test
//                                  (CarpetX/src/driver.cxx:1025-1080)
// FABs are views of device memory in AMReX FArrayBox layout.  Where the
// reference aborts (assert / CCTK_VERROR), these throw cxp::Error carrying
// the library status.
#pragma once

#include <array>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/cxp.h"
#include "../../include/cxp_tag.h"

namespace cxp {

class Error : public std::runtime_error {
public:
  Error(const std::string &what, int status)
      : std::runtime_error(what + ": " + cxp_status_string(status)),
        status_(status) {}
  int status() const { return status_; }

private:
  int status_;
};

inline void check(int status, const char *what) {
  if (status != CXP_OK)
    throw Error(what, status);
}

using IntVect = std::array<int, 3>;

// Inclusive index box [lo, hi] of a level.
struct Box {
  IntVect lo{0, 0, 0}, hi{-1, -1, -1};
  Box() = default;
  Box(IntVect l, IntVect h) : lo(l), hi(h) {}
  bool ok() const { return lo[0] <= hi[0] && lo[1] <= hi[1] && lo[2] <= hi[2]; }
  int64_t numPts() const {
    return ok() ? int64_t(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) *
                      (hi[2] - lo[2] + 1)
                : 0;
  }
  Box grow(int g) const {
    return Box({lo[0] - g, lo[1] - g, lo[2] - g}, {hi[0] + g, hi[1] + g, hi[2] + g});
  }
  bool contains(const Box &b) const {
    for (int d = 0; d < 3; ++d)
      if (b.lo[d] < lo[d] || b.hi[d] > hi[d])
        return false;
    return true;
  }
  bool operator==(const Box &o) const { return lo == o.lo && hi == o.hi; }
};

// Device-memory FAB view: component c of point (i,j,k) at
// data[c*comp_stride + (i-lo0) + (j-lo1)*nx + (k-lo2)*nx*ny].
struct Fab {
  double *data = nullptr;
  Box box;
  int ncomp = 1;
  int64_t comp_stride = 0;  // 0: the box volume
  cxp_fab c() const {
    cxp_fab f;
    f.data = data;
    for (int d = 0; d < 3; ++d) {
      f.lo[d] = box.lo[d];
      f.hi[d] = box.hi[d];
    }
    f.comp_stride = comp_stride ? comp_stride : box.numPts();
    f.ncomp = ncomp;
    return f;
  }
};

// One prolongation operator (an immutable, never-freed library handle, like
// the reference's static operator singletons).
class Prolongate3DRF2 {
public:
  explicit Prolongate3DRF2(cxp_op op) : op_(op) {}
  cxp_op handle() const { return op_; }

  Box CoarseBox(const Box &fine, int ratio = 2) const {
    if (ratio != 2)
      throw Error("CoarseBox", CXP_ERR_RATIO);
    Box c;
    check(cxp_coarse_box(op_, fine.lo.data(), fine.hi.data(), c.lo.data(),
                         c.hi.data()),
          "cxp_coarse_box");
    return c;
  }

  void interp(const Fab &crse, int crse_comp, const Fab &fine, int fine_comp,
              int ncomp, const Box &fine_region, int ratio = 2,
              bool per_group = false, void *stream = nullptr) const {
    if (ratio != 2)
      throw Error("interp", CXP_ERR_RATIO);
    const cxp_fab c = crse.c(), f = fine.c();
    check(cxp_prolongate(op_, &c, crse_comp, &f, fine_comp, ncomp,
                         fine_region.lo.data(), fine_region.hi.data(),
                         per_group ? 1 : 0, stream),
          "cxp_prolongate");
  }

  // The reference ignores solve_mask and forwards to interp
  // (prolongate_3d_rf2_impl.hxx:2204-2207).
  void interp_face(const Fab &crse, int crse_comp, const Fab &fine,
                   int fine_comp, int ncomp, const Box &fine_region,
                   int ratio = 2, bool per_group = false,
                   void *stream = nullptr) const {
    interp(crse, crse_comp, fine, fine_comp, ncomp, fine_region, ratio,
           per_group, stream);
  }

  // crse_bar += P^T fine_bar over CoarseBox(fine_region).
  void interp_adjoint(const Fab &crse_in, const Fab &crse_bar, int crse_comp,
                      const Fab &fine_bar, int fine_comp, int ncomp,
                      const Box &fine_region, void *stream = nullptr) const {
    const cxp_fab ci = crse_in.c(), cb = crse_bar.c(), fb = fine_bar.c();
    check(cxp_prolongate_adjoint(op_, &ci, &cb, crse_comp, &fb, fine_comp,
                                 ncomp, fine_region.lo.data(),
                                 fine_region.hi.data(), 0, stream),
          "cxp_prolongate_adjoint");
  }

  // Bytes of the branch record of a forward over fine_region (0 for
  // operators without ENO shifts or fallback).
  int64_t branch_bytes(const Box &fine_region, int ncomp,
                       bool per_group = false) const {
    int64_t n = 0;
    check(cxp_branch_bytes(op_, fine_region.lo.data(), fine_region.hi.data(),
                           ncomp, per_group ? 1 : 0, &n),
          "cxp_branch_bytes");
    return n;
  }

  // interp that also records every fine point's branch (device memory of
  // branch_bytes bytes; may be null when that is 0).
  void interp_record(const Fab &crse, int crse_comp, const Fab &fine,
                     int fine_comp, int ncomp, const Box &fine_region,
                     unsigned char *branches, bool per_group = false,
                     void *stream = nullptr) const {
    const cxp_fab c = crse.c(), f = fine.c();
    check(cxp_prolongate_record(op_, &c, crse_comp, &f, fine_comp, ncomp,
                                fine_region.lo.data(), fine_region.hi.data(),
                                per_group ? 1 : 0, branches, stream),
          "cxp_prolongate_record");
  }

  // crse_bar += P^T fine_bar from the branches of interp_record.
  void interp_adjoint_recorded(const unsigned char *branches,
                               const Fab &crse_bar, int crse_comp,
                               const Fab &fine_bar, int fine_comp, int ncomp,
                               const Box &fine_region, bool per_group = false,
                               void *stream = nullptr) const {
    const cxp_fab cb = crse_bar.c(), fb = fine_bar.c();
    check(cxp_prolongate_adjoint_recorded(
              op_, branches, &cb, crse_comp, &fb, fine_comp, ncomp,
              fine_region.lo.data(), fine_region.hi.data(), per_group ? 1 : 0,
              stream),
          "cxp_prolongate_adjoint_recorded");
  }

  // Per-direction (centring, interpolation, order, ghosts) and fallback.
  void describe(IntVect &cent, IntVect &intp, IntVect &order, IntVect &ghosts,
                int &fb) const {
    check(cxp_op_describe(op_, cent.data(), intp.data(), order.data(),
                          ghosts.data(), &fb),
          "cxp_op_describe");
  }

private:
  cxp_op op_;
};

// get_interpolator(prolongation_type, prolongation_order, indextype).
inline const Prolongate3DRF2 &get_interpolator(const std::string &type,
                                               int order,
                                               const IntVect &indextype) {
  static std::mutex mu;
  static std::map<std::tuple<std::string, int, int, int, int>,
                  std::unique_ptr<Prolongate3DRF2>>
      cache;
  std::lock_guard<std::mutex> lk(mu);
  auto &slot = cache[std::make_tuple(type, order, indextype[0], indextype[1],
                                     indextype[2])];
  if (!slot) {
    cxp_op op = nullptr;
    check(cxp_op_lookup(type.c_str(), order, indextype.data(), &op),
          "cxp_op_lookup");
    slot = std::make_unique<Prolongate3DRF2>(op);
  }
  return *slot;
}

// Restriction of `fine` onto coarse region `crse_region` of `crse`:
// injection in vertex-centred directions, mean of the two children in
// cell-centred ones.
inline void restrict_fab(const IntVect &indextype, const Fab &fine,
                         int fine_comp, const Fab &crse, int crse_comp,
                         int ncomp, const Box &crse_region,
                         void *stream = nullptr) {
  const cxp_fab f = fine.c(), c = crse.c();
  check(cxp_restrict(indextype.data(), &f, fine_comp, &c, crse_comp, ncomp,
                     crse_region.lo.data(), crse_region.hi.data(), stream),
        "cxp_restrict");
}

// A device-resident list of interp() calls, run by one persistent kernel.
// A refinement level: FABs, the valid box of each (valid boxes do not
// overlap) and the domain, the index range of the valid points.
struct Level {
  std::vector<Fab> fabs;
  std::vector<Box> valid;
  Box domain;
};

// Symmetry and boundary conditions of a level, per face (0 lower, 1 upper)
// and direction: CXP_BC_* (CarpetX/src/boundaries.hxx).  dirichlet and
// robin hold one value per appended component, parities three per
// component; the Robin condition's coordinates are x = xmin + i dx.
struct BoundaryConditions {
  int type[2][3] = {{CXP_BC_NONE, CXP_BC_NONE, CXP_BC_NONE},
                    {CXP_BC_NONE, CXP_BC_NONE, CXP_BC_NONE}};
  std::vector<double> dirichlet, robin;
  std::vector<IntVect> parities;
  std::array<double, 3> xmin{0.0, 0.0, 0.0}, dx{1.0, 1.0, 1.0};
};

namespace detail_ {
// C views of a Level and of BoundaryConditions, valid while they live.
struct CLevel {
  std::vector<cxp_fab> fabs;
  std::vector<int> valid;
  cxp_level l;
  explicit CLevel(const Level &L) {
    if (L.fabs.size() != L.valid.size())
      throw Error("Level: one valid box per FAB", CXP_ERR_ARG);
    for (const Fab &f : L.fabs)
      fabs.push_back(f.c());
    for (const Box &b : L.valid)
      for (const IntVect *v : {&b.lo, &b.hi})
        valid.insert(valid.end(), v->begin(), v->end());
    l.nfabs = (int)L.fabs.size();
    l.fabs = fabs.data();
    l.valid = valid.data();
    for (int d = 0; d < 3; ++d) {
      l.domain_lo[d] = L.domain.lo[d];
      l.domain_hi[d] = L.domain.hi[d];
    }
  }
};
struct CBc {
  std::vector<int> par;
  cxp_bc b{};
  explicit CBc(const BoundaryConditions *bc) {
    if (bc)
      set(*bc);
  }
  explicit CBc(const BoundaryConditions &bc) { set(bc); }
  void set(const BoundaryConditions &bc) {
    for (const IntVect &p : bc.parities)
      par.insert(par.end(), p.begin(), p.end());
    for (int side = 0; side < 2; ++side)
      for (int d = 0; d < 3; ++d)
        b.type[side][d] = bc.type[side][d];
    b.dirichlet_values = bc.dirichlet.empty() ? nullptr : bc.dirichlet.data();
    b.robin_values = bc.robin.empty() ? nullptr : bc.robin.data();
    b.parities = par.empty() ? nullptr : par.data();
    for (int d = 0; d < 3; ++d) {
      b.xmin[d] = bc.xmin[d];
      b.dx[d] = bc.dx[d];
    }
  }
};
} // namespace detail_

class WorkList {
  using CLevel = detail_::CLevel;
  using CBc = detail_::CBc;

public:
  WorkList() { check(cxp_worklist_create(&wl_), "cxp_worklist_create"); }
  ~WorkList() {
    if (wl_)
      cxp_worklist_destroy(wl_);
  }
  WorkList(const WorkList &) = delete;
  WorkList &operator=(const WorkList &) = delete;

  void append(const Prolongate3DRF2 &op, const Fab &crse, int crse_comp,
              const Fab &fine, int fine_comp, int ncomp, const Box &region) {
    const cxp_fab c = crse.c(), f = fine.c();
    check(cxp_worklist_append(wl_, op.handle(), &c, crse_comp, &f, fine_comp,
                              ncomp, region.lo.data(), region.hi.data()),
          "cxp_worklist_append");
  }
  // The ghost shell of `fine`'s box around `interior`.
  void append_shell(const Prolongate3DRF2 &op, const Fab &crse, int crse_comp,
                    const Fab &fine, int fine_comp, int ncomp,
                    const Box &interior) {
    const cxp_fab c = crse.c(), f = fine.c();
    check(cxp_worklist_append_shell(wl_, op.handle(), &c, crse_comp, &f,
                                    fine_comp, ncomp, interior.lo.data(),
                                    interior.hi.data()),
          "cxp_worklist_append_shell");
  }
  // The ghost fill of the whole fine level from the coarse level
  // (FillPatch_ProlongateGhosts); with bc, the coarse points outside the
  // coarse domain take the boundary values.
  void append_level(const Prolongate3DRF2 &op, const Level &crse,
                    int crse_comp, const Level &fine, int fine_comp, int ncomp,
                    const BoundaryConditions *bc = nullptr) {
    CLevel c(crse), f(fine);
    if (!bc) {
      check(cxp_worklist_append_level(wl_, op.handle(), &c.l, crse_comp, &f.l,
                                      fine_comp, ncomp),
            "cxp_worklist_append_level");
      return;
    }
    CBc b(*bc);
    check(cxp_worklist_append_level_bc(wl_, op.handle(), &c.l, crse_comp, &f.l,
                                       fine_comp, ncomp, &b.b),
          "cxp_worklist_append_level_bc");
  }
  // A new level's fill (FillPatch_NewLevel): whole fine boxes.
  void append_new_level(const Prolongate3DRF2 &op, const Level &crse,
                        int crse_comp, const Level &fine, int fine_comp,
                        int ncomp, const BoundaryConditions *bc = nullptr) {
    CLevel c(crse), f(fine);
    CBc b(bc);
    check(cxp_worklist_append_new_level(wl_, op.handle(), &c.l, crse_comp, &f.l,
                                        fine_comp, ncomp, bc ? &b.b : nullptr),
          "cxp_worklist_append_new_level");
  }
  // A regridded level's fill (FillPatch_RemakeLevel): the old fine level's
  // valid data where it exists, prolongation elsewhere.
  void append_remake_level(const Prolongate3DRF2 &op, const Level &crse,
                           int crse_comp, const Level &old_fine,
                           const Level &fine, int fine_comp, int ncomp,
                           const BoundaryConditions *bc = nullptr) {
    CLevel c(crse), o(old_fine), f(fine);
    CBc b(bc);
    check(cxp_worklist_append_remake_level(wl_, op.handle(), &c.l, crse_comp,
                                           &o.l, &f.l, fine_comp, ncomp,
                                           bc ? &b.b : nullptr),
          "cxp_worklist_append_remake_level");
  }
  // The level's own boundary conditions, in place, after the fills.
  void append_boundaries(const Level &level, const IntVect &indextype,
                         int comp, int ncomp, const BoundaryConditions &bc) {
    CLevel l(level);
    CBc b(bc);
    check(cxp_worklist_append_boundaries(wl_, &l.l, indextype.data(), comp,
                                         ncomp, &b.b),
          "cxp_worklist_append_boundaries");
  }
  // the reference's prolongate_per_group parameter for this list
  void set_per_group(bool per_group) {
    check(cxp_worklist_set_per_group(wl_, per_group ? 1 : 0),
          "cxp_worklist_set_per_group");
  }
  void upload(void *stream = nullptr) {
    check(cxp_worklist_upload(wl_, stream), "cxp_worklist_upload");
  }
  void run(void *stream = nullptr) {
    check(cxp_prolongate_batched(wl_, stream), "cxp_prolongate_batched");
  }
  void run_per_patch(void *stream = nullptr) {
    check(cxp_prolongate_per_patch(wl_, stream), "cxp_prolongate_per_patch");
  }
  void poison(double value, void *stream = nullptr) {
    check(cxp_worklist_poison(wl_, value, stream), "cxp_worklist_poison");
  }
  cxp_worklist handle() const { return wl_; }

private:
  cxp_worklist wl_ = nullptr;
};

// One box-in-box region at the level being decided (the parameters
// active_N, shape_N, num_levels_N, position_*_N and, per direction,
// radius_*_N[level] or radius_N[level]).
struct BoxInBoxRegion {
  bool active = true;
  int shape = CXP_BIB_CUBE;
  int num_levels = 1;
  std::array<double, 3> position{0.0, 0.0, 0.0};
  std::array<double, 3> radius{-1.0, -1.0, -1.0};
};

// regrid_error over `region` of a one-component FAB: 1 in the cells a
// region refines, 0 elsewhere.  x0 is the coordinate of cell 0's centre.
inline void boxinbox_setup(const std::vector<BoxInBoxRegion> &regions,
                           int level, const std::array<double, 3> &x0,
                           const std::array<double, 3> &dx, const Fab &err,
                           const Box &region, void *stream = nullptr) {
  std::vector<cxp_bib_region> r(regions.size() ? regions.size() : 1);
  for (size_t i = 0; i < regions.size(); ++i) {
    r[i].active = regions[i].active;
    r[i].shape = regions[i].shape;
    r[i].num_levels = regions[i].num_levels;
    for (int d = 0; d < 3; ++d) {
      r[i].position[d] = regions[i].position[d];
      r[i].radius[d] = regions[i].radius[d];
    }
  }
  check(cxp_boxinbox_tag(r.data(), int(regions.size()), level, x0.data(),
                         dx.data(), err.data, err.box.lo.data(),
                         err.box.hi.data(), region.lo.data(),
                         region.hi.data(), stream),
        "cxp_boxinbox_tag");
}

// ErrorEst's tags over `region` (err >= threshold): tags (device, one byte
// per cell of err's box, CXP_TAG_SET = amrex::TagBox::SET or CXP_TAG_CLEAR,
// or null) and box_out (device, 7 ints, or null).
inline void error_est(const Fab &err, const Box &region, double threshold,
                      signed char *tags, int *box_out, void *stream = nullptr) {
  check(cxp_error_est(err.data, err.box.lo.data(), err.box.hi.data(),
                      region.lo.data(), region.hi.data(), threshold, tags,
                      box_out, stream),
        "cxp_error_est");
}

// The 7 result words of error_est, copied to the host: the number of set
// cells and their bounding box (empty when none).
struct TagSummary {
  int64_t count = 0;
  Box box;
  explicit TagSummary(const int w[7]) : count(w[0]) {
    if (count > 0)
      box = Box({w[1], w[2], w[3]}, {w[4], w[5], w[6]});
  }
};

// A captured fill sequence (hipGraph).
class Graph {
public:
  template <typename F> Graph(F &&record, void *stream) {
    check(cxp_graph_capture_begin(stream), "cxp_graph_capture_begin");
    record();
    check(cxp_graph_capture_end(stream, &g_), "cxp_graph_capture_end");
  }
  ~Graph() {
    if (g_)
      cxp_graph_destroy(g_);
  }
  Graph(const Graph &) = delete;
  Graph &operator=(const Graph &) = delete;
  void launch(void *stream = nullptr) {
    check(cxp_graph_launch(g_, stream), "cxp_graph_launch");
  }

private:
  cxp_graph g_ = nullptr;
};

} // namespace cxp
