// Test shim: the few AMReX declarations carpetx_amd/host/cxp_interpolater.hxx
// uses (IntVect, Box, FArrayBox, IArrayBox, Geometry, BCRec, Vector, RunOn,
// Abort, Gpu::gpuStream, Interpolater), with AMReX's names and signatures
// (AMReX 25.11, which the reference builds against and which this image
// lacks).  FArrayBox wraps device memory the caller owns.  Abort throws, so
// that a check program can see the adapter's aborts.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

namespace amrex {

class IntVect {
  int v_[3];

public:
  IntVect() : v_{0, 0, 0} {}
  explicit IntVect(int s) : v_{s, s, s} {}
  IntVect(int a, int b, int c) : v_{a, b, c} {}
  explicit IntVect(const int *p) : v_{p[0], p[1], p[2]} {}
  const int *getVect() const { return v_; }
  int operator[](int d) const { return v_[d]; }
  bool operator==(const IntVect &o) const {
    return v_[0] == o.v_[0] && v_[1] == o.v_[1] && v_[2] == o.v_[2];
  }
  bool operator!=(const IntVect &o) const { return !(*this == o); }
};

struct IndexType {
  int t[3] = {0, 0, 0};  // 0 cell, 1 node (AMReX's CellIndex)
};

class Box {
  IntVect lo_, hi_;
  IndexType ix_;

public:
  Box() = default;
  Box(const IntVect &lo, const IntVect &hi, IndexType ix = {})
      : lo_(lo), hi_(hi), ix_(ix) {}
  const IntVect &smallEnd() const { return lo_; }
  const IntVect &bigEnd() const { return hi_; }
  int smallEnd(int d) const { return lo_[d]; }
  int bigEnd(int d) const { return hi_[d]; }
  IndexType ixType() const { return ix_; }
  long numPts() const {
    return long(hi_[0] - lo_[0] + 1) * (hi_[1] - lo_[1] + 1) *
           (hi_[2] - lo_[2] + 1);
  }
};

class FArrayBox {
  Box box_;
  int ncomp_;
  double *p_;

public:
  FArrayBox(const Box &b, int ncomp, double *device_data)
      : box_(b), ncomp_(ncomp), p_(device_data) {}
  const Box &box() const { return box_; }
  int nComp() const { return ncomp_; }
  double *dataPtr() { return p_; }
  const double *dataPtr() const { return p_; }
};

class IArrayBox {};
class Geometry {};
struct BCRec {};
template <class T> using Vector = std::vector<T>;
enum class RunOn { Host, Gpu, Device = Gpu };

[[noreturn]] inline void Abort(const std::string &msg) {
  throw std::runtime_error(msg);
}

namespace Gpu {
inline hipStream_t &current_stream() {
  static hipStream_t s = nullptr;
  return s;
}
inline hipStream_t gpuStream() { return current_stream(); }
} // namespace Gpu

class Interpolater {
public:
  virtual ~Interpolater() = default;
  virtual Box CoarseBox(const Box &fine, int ratio) = 0;
  virtual Box CoarseBox(const Box &fine, const IntVect &ratio) = 0;
  virtual void interp(const FArrayBox &crse, int crse_comp, FArrayBox &fine,
                      int fine_comp, int ncomp, const Box &fine_region,
                      const IntVect &ratio, const Geometry &crse_geom,
                      const Geometry &fine_geom, Vector<BCRec> const &bcr,
                      int actual_comp, int actual_state, RunOn gpu_or_cpu) = 0;
  virtual void interp_face(const FArrayBox &crse, int crse_comp,
                           FArrayBox &fine, int fine_comp, int ncomp,
                           const Box &fine_region, const IntVect &ratio,
                           const IArrayBox &solve_mask,
                           const Geometry &crse_geom,
                           const Geometry &fine_geom, Vector<BCRec> const &bcr,
                           int bccomp, RunOn gpu_or_cpu) = 0;
};

} // namespace amrex
