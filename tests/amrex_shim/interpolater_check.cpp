// CarpetX's Interpolater adapter (carpetx_amd/host/cxp_interpolater.hxx),
// compiled against the AMReX / Cactus shim in this directory, run on the
// GPU: the adapter's interp / interp_face / CoarseBox give exactly what the
// C ABI gives for the same arguments (prolongate_per_group honoured), a cubic
// is reproduced through it, its aborts fire where the reference's do, and
// get_interpolator / get_group_interpolator hand out one adapter per
// operator.  Prints one line per check and "interpolater check ok".
#include <cmath>
#include <cstdio>
#include <random>
#include <stdexcept>
#include <vector>

#include "cxp_interpolater.hxx"

namespace {

int failures = 0;

void expect(bool ok, const char *what) {
  std::printf("  %-64s %s\n", what, ok ? "ok" : "FAILED");
  if (!ok)
    ++failures;
}

bool hip_ok(hipError_t e) { return e == hipSuccess; }

struct DevFab {
  amrex::Box box;
  int ncomp;
  double *d = nullptr;
  DevFab(const amrex::Box &b, int nc) : box(b), ncomp(nc) {
    (void)hipMalloc(&d, sizeof(double) * b.numPts() * nc);
  }
  ~DevFab() { (void)hipFree(d); }
  amrex::FArrayBox fab() { return amrex::FArrayBox(box, ncomp, d); }
  void set(const std::vector<double> &h) {
    (void)hipMemcpy(d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice);
  }
  std::vector<double> get() const {
    std::vector<double> h(box.numPts() * ncomp);
    (void)hipMemcpy(h.data(), d, sizeof(double) * h.size(), hipMemcpyDeviceToHost);
    return h;
  }
};

amrex::Box grow(const amrex::Box &b, int g) {
  return amrex::Box(amrex::IntVect(b.smallEnd(0) - g, b.smallEnd(1) - g, b.smallEnd(2) - g),
                    amrex::IntVect(b.bigEnd(0) + g, b.bigEnd(1) + g, b.bigEnd(2) + g));
}

cxp_fab cfab(DevFab &f) {
  cxp_fab r;
  r.data = f.d;
  for (int d = 0; d < 3; ++d) {
    r.lo[d] = f.box.smallEnd(d);
    r.hi[d] = f.box.bigEnd(d);
  }
  r.comp_stride = f.box.numPts();
  r.ncomp = f.ncomp;
  return r;
}

template <class F> bool throws(F f) {
  try {
    f();
  } catch (const std::runtime_error &) {
    return true;
  }
  return false;
}

} // namespace

int main() {
  hipStream_t s;
  if (!hip_ok(hipStreamCreate(&s)))
    return 2;
  amrex::Gpu::current_stream() = s;
  const amrex::Geometry geom;
  const amrex::Vector<amrex::BCRec> bcr;
  const amrex::IArrayBox mask;
  const amrex::Box region(amrex::IntVect(-3, 1, -2), amrex::IntVect(20, 17, 15));
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> uni(-1, 1);

  struct Case {
    const char *type;
    int order;
    std::array<int, 3> it;
    int per_group;
  };
  const Case cases[] = {{"interpolate", 3, {0, 0, 0}, 0}, {"natural", 5, {1, 1, 1}, 0},
                        {"eno", 3, {1, 1, 1}, 0},         {"eno", 3, {1, 1, 1}, 1},
                        {"minmod", 3, {1, 0, 1}, 0},      {"poly-cons3lfb", 5, {1, 0, 1}, 1}};
  std::printf("Interpolater adapter through libcxp.so\n");
  for (const Case &c : cases) {
    cctk_shim::prolongate_per_group() = c.per_group;
    amrex::Interpolater *P = CarpetX::get_interpolator(c.type, c.order, c.it);
    const amrex::Box cb = P->CoarseBox(region, 2);
    cxp_op op;
    int lo[3], hi[3];
    (void)cxp_op_lookup(c.type, c.order, c.it.data(), &op);
    (void)cxp_coarse_box(op, region.smallEnd().getVect(), region.bigEnd().getVect(), lo, hi);
    bool cb_ok = true;
    for (int d = 0; d < 3; ++d)
      cb_ok = cb_ok && cb.smallEnd(d) == lo[d] && cb.bigEnd(d) == hi[d];
    cb_ok = cb_ok && P->CoarseBox(region, amrex::IntVect(2)).smallEnd() == cb.smallEnd();
    DevFab crse(grow(cb, 1), 2), f1(grow(region, 2), 2), f2(grow(region, 2), 2),
        f3(grow(region, 2), 2);
    std::vector<double> h(crse.box.numPts() * 2);
    for (double &x : h)
      x = uni(rng);
    crse.set(h);
    const std::vector<double> sentinel(f1.box.numPts() * 2, 1234.5);
    f1.set(sentinel);
    f2.set(sentinel);
    f3.set(sentinel);
    amrex::FArrayBox cf = crse.fab(), ff1 = f1.fab(), ff3 = f3.fab();
    P->interp(cf, 0, ff1, 0, 2, region, amrex::IntVect(2), geom, geom, bcr, 0, 0,
              amrex::RunOn::Gpu);
    P->interp_face(cf, 0, ff3, 0, 2, region, amrex::IntVect(2), mask, geom, geom, bcr, 0,
                   amrex::RunOn::Gpu);
    const cxp_fab ccf = cfab(crse), cf2 = cfab(f2);
    const int st = cxp_prolongate(op, &ccf, 0, &cf2, 0, 2, region.smallEnd().getVect(),
                                  region.bigEnd().getVect(), c.per_group, s);
    (void)hipStreamSynchronize(s);
    const std::vector<double> a = f1.get(), b = f2.get(), f = f3.get();
    char what[160];
    std::snprintf(what, sizeof what, "%s o%d (%d,%d,%d) pg=%d: CoarseBox, interp == C ABI",
                  c.type, c.order, c.it[0], c.it[1], c.it[2], c.per_group);
    expect(st == CXP_OK && cb_ok && a == b && a != sentinel, what);
    std::snprintf(what, sizeof what, "%s o%d pg=%d: interp_face forwards", c.type, c.order,
                  c.per_group);
    expect(f == a, what);
  }
  cctk_shim::prolongate_per_group() = 0;

  // a cubic through the adapter (VVV interpolate o3 reproduces it)
  {
    amrex::Interpolater *P = CarpetX::get_interpolator("interpolate", 3, {0, 0, 0});
    const amrex::Box cb = P->CoarseBox(region, 2);
    DevFab crse(cb, 1), fine(region, 1);
    auto poly = [](double x, double y, double z) {
      return (0.5 + x - 0.25 * x * x + 0.125 * x * x * x) * (1 - y + 0.5 * y * y) *
             (0.75 + 0.5 * z * z * z);
    };
    std::vector<double> h(cb.numPts());
    long n = 0;
    for (int k = cb.smallEnd(2); k <= cb.bigEnd(2); ++k)
      for (int j = cb.smallEnd(1); j <= cb.bigEnd(1); ++j)
        for (int i = cb.smallEnd(0); i <= cb.bigEnd(0); ++i)
          h[n++] = poly(i / 8.0, j / 8.0, k / 8.0);
    crse.set(h);
    amrex::FArrayBox cf = crse.fab(), ff = fine.fab();
    P->interp(cf, 0, ff, 0, 1, region, amrex::IntVect(2), geom, geom, bcr, 0, 0,
              amrex::RunOn::Gpu);
    (void)hipStreamSynchronize(s);
    const std::vector<double> g = fine.get();
    double err = 0, mx = 0;
    n = 0;
    for (int k = region.smallEnd(2); k <= region.bigEnd(2); ++k)
      for (int j = region.smallEnd(1); j <= region.bigEnd(1); ++j)
        for (int i = region.smallEnd(0); i <= region.bigEnd(0); ++i) {
          const double e = poly(i / 16.0, j / 16.0, k / 16.0);
          err = std::fmax(err, std::fabs(g[n++] - e));
          mx = std::fmax(mx, std::fabs(e));
        }
    std::printf("      max error %.3e of max %.3e\n", err, mx);
    expect(err <= 1e-13 * mx, "VVV interpolate o3 reproduces a tensor cubic");
  }

  // aborts where the reference aborts
  {
    amrex::Interpolater *P = CarpetX::get_interpolator("natural", 3, {1, 1, 1});
    DevFab crse(region, 1), fine(region, 1);
    amrex::FArrayBox cf = crse.fab(), ff = fine.fab();
    expect(throws([&] { P->CoarseBox(region, 4); }), "CoarseBox with ratio 4 aborts");
    expect(throws([&] {
             P->interp(cf, 0, ff, 0, 1, region, amrex::IntVect(3), geom, geom, bcr, 0, 0,
                       amrex::RunOn::Gpu);
           }),
           "interp with ratio 3 aborts");
    expect(throws([&] {
             P->interp(cf, 0, ff, 0, 1, region, amrex::IntVect(2), geom, geom, bcr, 0, 0,
                       amrex::RunOn::Gpu);
           }),
           "interp with a coarse FAB short of CoarseBox aborts");
    expect(throws([] { CarpetX::get_interpolator("ddf-eno", 3, {1, 1, 1}); }),
           "unknown prolongation type aborts (driver.cxx:620)");
    expect(throws([] { CarpetX::get_interpolator("eno", 1, {1, 1, 1}); }),
           "unsupported order aborts (driver.cxx:623)");
  }

  // one adapter per operator; group tags override the parameters
  {
    const std::array<int, 3> it{1, 0, 1};
    cctk_shim::prolongation_type() = "natural";
    cctk_shim::prolongation_order() = 1;
    amrex::Interpolater *a = CarpetX::get_interpolator("eno", 3, it);
    expect(a == CarpetX::get_interpolator("eno", 3, it), "same key, same adapter");
    expect(a == CarpetX::get_group_interpolator(
                    "prolongation_type=\"eno\" prolongation_order=3 parities={+1 +1 +1}", it),
           "group tags set type and order");
    expect(CarpetX::get_group_interpolator(nullptr, it) ==
               CarpetX::get_interpolator("natural", 1, it),
           "no tags: the driver parameters");
    expect(CarpetX::get_group_interpolator("checkpoint=\"no\" Prolongation_Order=5", it) ==
               CarpetX::get_interpolator("natural", 5, it),
           "tags set the order only (case-insensitive key)");
    expect(throws([&] { CarpetX::get_group_interpolator("prolongation_order=\"x\"", it); }),
           "a non-integer prolongation_order aborts");
  }
  (void)hipStreamDestroy(s);
  if (failures) {
    std::printf("interpolater check FAILED (%d)\n", failures);
    return 1;
  }
  std::printf("interpolater check ok\n");
  return 0;
}
