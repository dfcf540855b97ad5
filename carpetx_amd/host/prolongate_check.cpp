// prolongate_check -- the one-box configuration through the C++ host API,
// checked the way CarpetX's TestProlongate checks its operators
// (TestProlongate/src/test.cxx:43-73): data that the operator must
// reproduce exactly (polynomials of the operator's order per direction,
// cell averages in conservative directions) is prolongated, and the fine
// values are compared with the same function on the fine grid.
//
//   1. VC order-3 polynomial interpolation, coarse 32^3 cells + 3 ghosts ->
//      fine [0,63]^3, a seeded random tensor cubic: exact to rounding.
//   2. Restriction back: injection of the prolongation of uniform[-1,1]
//      data reproduces the coarse values bit for bit.
//   3. The same prolongation through a work list (persistent kernel) and a
//      replayed graph: bit-identical to the per-box call.
//   4. Every registered operator of the families without limiter branches
//      or fallback (interpolate, conservative, ddf, hermite, natural,
//      minmod) in all 8 centrings: exact for polynomials of its
//      per-direction order.
//   5. A level ghost fill with reflection symmetry on the lower faces: the
//      coarse FAB holds NaN outside the domain, the fill computes those
//      points by reflection, and data even (component 0) or odd in x
//      (component 1, parity -1) about the faces is reproduced exactly.
//   6. The regrid-time fills on the same data: a new level filled whole
//      and then given its own reflection boundaries, and a remake of it
//      onto a shifted box (old data where the old box covers, prolongation
//      elsewhere), both reproduce the data at every point of the FABs.
//   7. Box-in-box tags of 64^3 cells of [0, 1]^3 for a cube of radius 1/4
//      about the centre, then the threshold test: exactly the cells
//      [17, 46]^3 are tagged (|i + 1/2 - 32| <= 16 - 1 on this dyadic
//      grid), counted and boxed on the device, and the tag bytes agree.
//
// Exit status 0 when every check passes.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "cxp.hpp"

namespace {

int failures = 0;

void expect(bool ok, const char *what) {
  std::printf("%s: %s\n", ok ? "ok  " : "FAIL", what);
  std::fflush(stdout);
  if (!ok)
    ++failures;
}

struct DevFab {
  cxp::Fab fab;
  explicit DevFab(const cxp::Box &b, int ncomp = 1) {
    fab.box = b;
    fab.ncomp = ncomp;
    cxp::check(hipMalloc(&fab.data, sizeof(double) * b.numPts() * ncomp) ==
                       hipSuccess
                   ? CXP_OK
                   : CXP_ERR_HIP,
               "hipMalloc");
  }
  ~DevFab() { (void)hipFree(fab.data); }
  void put(const std::vector<double> &h) {
    (void)hipMemcpy(fab.data, h.data(), h.size() * sizeof(double),
                    hipMemcpyHostToDevice);
  }
  std::vector<double> get() const {
    std::vector<double> h(fab.box.numPts() * fab.ncomp);
    (void)hipMemcpy(h.data(), fab.data, h.size() * sizeof(double),
                    hipMemcpyDeviceToHost);
    return h;
  }
};

// A 1-D polynomial sum_q a[q] x^q and its average over [x - w/2, x + w/2].
struct Poly1 {
  std::vector<double> a;
  double at(double x) const {
    double s = 0, p = 1;
    for (double c : a) {
      s += c * p;
      p *= x;
    }
    return s;
  }
  double avg(double x, double w) const {
    double s = 0;
    const double lo = x - w / 2, hi = x + w / 2;
    for (size_t q = 0; q < a.size(); ++q)
      s += a[q] * (std::pow(hi, q + 1) - std::pow(lo, q + 1)) / (q + 1);
    return s / w;
  }
};

// Values of prod_d p_d on a box of spacing h: vertex-centred directions at
// i h, cell-centred ones at (i + 1/2) h, averaged over the cell where `avg`.
std::vector<double> sample(const cxp::Box &b, double h, const cxp::IntVect &cc,
                           const std::array<bool, 3> &avg,
                           const std::array<Poly1, 3> &p) {
  std::vector<double> v[3];
  for (int d = 0; d < 3; ++d)
    for (int i = b.lo[d]; i <= b.hi[d]; ++i) {
      const double x = (i + 0.5 * cc[d]) * h;
      v[d].push_back(avg[d] ? p[d].avg(x, h) : p[d].at(x));
    }
  std::vector<double> out;
  out.reserve(b.numPts());
  for (double z : v[2])
    for (double y : v[1])
      for (double x : v[0])
        out.push_back(x * y * z);
  return out;
}

double max_rel_err(const std::vector<double> &got,
                   const std::vector<double> &want) {
  double m = 0, scale = 0;
  for (size_t i = 0; i < want.size(); ++i)
    scale = std::fmax(scale, std::fabs(want[i]));
  for (size_t i = 0; i < want.size(); ++i)
    m = std::fmax(m, std::fabs(got[i] - want[i]));
  return m / (scale > 0 ? scale : 1);
}

std::array<Poly1, 3> random_polys(std::mt19937_64 &rng, const cxp::IntVect &deg) {
  std::uniform_real_distribution<double> u(-1, 1);
  std::array<Poly1, 3> p;
  for (int d = 0; d < 3; ++d)
    for (int q = 0; q <= deg[d]; ++q)
      p[d].a.push_back(u(rng));
  return p;
}

} // namespace

int main() {
  std::mt19937_64 rng(2);
  const int n = 32;  // coarse cells
  const double hc = 1.0 / n, hf = hc / 2;
  const cxp::Box region({0, 0, 0}, {2 * n - 1, 2 * n - 1, 2 * n - 1});
  hipStream_t s;
  (void)hipStreamCreate(&s);

  // 1. VC order-3 polynomial interpolation of a random tensor cubic
  const auto &P = cxp::get_interpolator("interpolate", 3, {0, 0, 0});
  const cxp::Box cbox({-3, -3, -3}, {n + 3, n + 3, n + 3});
  expect(cbox.contains(P.CoarseBox(region)), "CoarseBox([0,63]^3) inside coarse [-3,35]^3");
  const auto cubic = random_polys(rng, {3, 3, 3});
  DevFab crse(cbox), fine(region);
  crse.put(sample(cbox, hc, {0, 0, 0}, {false, false, false}, cubic));
  P.interp(crse.fab, 0, fine.fab, 0, 1, region, 2, false, s);
  (void)hipStreamSynchronize(s);
  const auto f1 = fine.get();
  const double e1 = max_rel_err(f1, sample(region, hf, {0, 0, 0},
                                           {false, false, false}, cubic));
  std::printf("      VC o3 cubic: max rel error %.3e\n", e1);
  expect(e1 < 1e-13, "VC order-3 prolongation reproduces a tensor cubic");

  // 2. injection round trip on uniform[-1,1] data, bit for bit
  {
    std::uniform_real_distribution<double> u(-1, 1);
    std::vector<double> h(cbox.numPts());
    for (double &x : h)
      x = u(rng);
    crse.put(h);
    P.interp(crse.fab, 0, fine.fab, 0, 1, region, 2, false, s);
    const cxp::Box creg({0, 0, 0}, {n - 1, n - 1, n - 1});
    DevFab back(creg);
    cxp::restrict_fab({0, 0, 0}, fine.fab, 0, back.fab, 0, 1, creg, s);
    (void)hipStreamSynchronize(s);
    const auto b = back.get();
    bool same = true;
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          const double want = h[(i + 3) + (j + 3) * (n + 7) + (k + 3) * (n + 7) * (n + 7)];
          same = same && std::memcmp(&want, &b[i + j * n + k * n * n], 8) == 0;
        }
    expect(same, "restrict(prolongate(u)) == u bit for bit (VC injection)");
  }

  // 3. the same call as a work list and as a replayed graph
  {
    const auto ref = fine.get();
    DevFab f2(region);
    cxp::WorkList wl;
    wl.append(P, crse.fab, 0, f2.fab, 0, 1, region);
    wl.upload(s);
    wl.run(s);
    (void)hipStreamSynchronize(s);
    expect(f2.get() == ref, "work list result equals the per-box call");
    (void)hipMemsetAsync(f2.fab.data, 0, sizeof(double) * region.numPts(), s);
    cxp::Graph g([&] { wl.run(s); }, s);
    for (int r = 0; r < 3; ++r)
      g.launch(s);
    (void)hipStreamSynchronize(s);
    expect(f2.get() == ref, "graph replays equal the per-box call");
  }

  // 4. every operator of the families without limiters or fallback
  const char *fams[] = {"interpolate", "conservative", "ddf", "hermite",
                        "natural", "minmod"};
  int checked = 0, exact = 0;
  for (const char *fam : fams)
    for (int order = 0; order <= 7; ++order)
      for (int c = 0; c < 8; ++c) {
        const cxp::IntVect it{(c >> 2) & 1, (c >> 1) & 1, c & 1};
        cxp_op h = nullptr;
        if (cxp_op_lookup(fam, order, it.data(), &h) != CXP_OK)
          continue;
        const auto &op = cxp::get_interpolator(fam, order, it);
        cxp::IntVect cent, intp, ord, gh;
        int fb = 0;
        op.describe(cent, intp, ord, gh, fb);
        bool skip = false;
        cxp::IntVect deg;
        std::array<bool, 3> avg;
        for (int d = 0; d < 3; ++d) {
          // VC CONS order 0 returns 0 at odd points (reference quirk)
          skip = skip || (cent[d] == CXP_VC && intp[d] == CXP_CONS && ord[d] == 0);
          deg[d] = intp[d] == CXP_MINMOD ? 1 : ord[d];
          avg[d] = cent[d] == CXP_CC && (intp[d] == CXP_CONS || intp[d] == CXP_MINMOD);
        }
        if (skip)
          continue;
        const cxp::Box reg({-3, 1, -2}, {20, 17, 13});
        const cxp::Box cb = op.CoarseBox(reg);
        const auto p = random_polys(rng, deg);
        DevFab cr(cb), fi(reg);
        cr.put(sample(cb, hc, it, avg, p));
        op.interp(cr.fab, 0, fi.fab, 0, 1, reg, 2, false, s);
        (void)hipStreamSynchronize(s);
        const double e = max_rel_err(fi.get(), sample(reg, hf, it, avg, p));
        ++checked;
        if (e < 1e-12)
          ++exact;
        else
          std::printf("      %s o%d it=(%d,%d,%d): max rel error %.3e\n", fam,
                      order, it[0], it[1], it[2], e);
      }
  std::printf("      %d of %d operators exact for polynomials of their order\n",
              exact, checked);
  expect(checked > 100 && exact == checked,
         "every operator reproduces polynomials of its order");

  // 5. level fill with reflection on the lower faces
  {
    const cxp::Box cdom({0, 0, 0}, {n, n, n});
    const cxp::Box cfab = cdom.grow(3);
    const cxp::Box fvalid({0, 0, 0}, {15, 15, 15});
    const cxp::Box ffab = fvalid.grow(3);
    // component 0: (1 + x^2/2)(2 - y^2)(1 + z^2); component 1: (x - x^3) y^2 (3 + z^2)
    auto f = [](int c, double x, double y, double z) {
      return c == 0 ? (1 + 0.5 * x * x) * (2 - y * y) * (1 + z * z)
                    : (x - x * x * x) * (y * y) * (3 + z * z);
    };
    auto fill = [&](const cxp::Box &b, double h, bool nan_outside) {
      std::vector<double> v;
      for (int c = 0; c < 2; ++c)
        for (int k = b.lo[2]; k <= b.hi[2]; ++k)
          for (int j = b.lo[1]; j <= b.hi[1]; ++j)
            for (int i = b.lo[0]; i <= b.hi[0]; ++i) {
              const bool in = i >= 0 && j >= 0 && k >= 0 && i <= n && j <= n && k <= n;
              v.push_back(nan_outside && !in ? std::nan("") : f(c, i * h, j * h, k * h));
            }
      return v;
    };
    DevFab cr(cfab, 2), fi(ffab, 2);
    cr.put(fill(cfab, hc, true));
    std::vector<double> init(ffab.numPts() * 2, 0.0);
    fi.put(init);
    cxp::Level cl{{cr.fab}, {cdom}, cdom};
    cxp::Level fl{{fi.fab}, {fvalid}, cxp::Box({0, 0, 0}, {2 * n, 2 * n, 2 * n})};
    cxp::BoundaryConditions bc;
    for (int d = 0; d < 3; ++d)
      bc.type[0][d] = CXP_BC_REFLECTION;
    bc.parities = {{1, 1, 1}, {-1, 1, 1}};
    const auto &P3 = cxp::get_interpolator("interpolate", 3, {0, 0, 0});
    cxp::WorkList wl;
    wl.append_level(P3, cl, 0, fl, 0, 2, &bc);
    wl.upload(s);
    wl.run(s);
    (void)hipStreamSynchronize(s);
    const auto got = fi.get();
    const auto want = fill(ffab, hf, false);
    double m = 0;
    size_t q = 0, ghosts = 0;
    for (int c = 0; c < 2; ++c)
      for (int k = ffab.lo[2]; k <= ffab.hi[2]; ++k)
        for (int j = ffab.lo[1]; j <= ffab.hi[1]; ++j)
          for (int i = ffab.lo[0]; i <= ffab.hi[0]; ++i, ++q) {
            const bool valid = i >= 0 && j >= 0 && k >= 0 && i <= 15 && j <= 15 && k <= 15;
            if (valid) {
              m = std::fmax(m, std::fabs(got[q]));  // left alone: 0
            } else {
              ++ghosts;
              m = std::fmax(m, std::fabs(got[q] - want[q]));
            }
          }
    std::printf("      level fill, %zu ghost points: max error %.3e\n", ghosts, m);
    expect(ghosts > 0 && m < 1e-12,
           "level fill with reflection reproduces even and odd data");

    // 6. new level (whole box, then its own boundaries) and a remake
    auto max_err = [&](const DevFab &d, const cxp::Box &b) {
      const auto g = d.get();
      const auto w = fill(b, hf, false);
      double e = 0;
      for (size_t i = 0; i < g.size(); ++i)
        e = std::fmax(e, std::fabs(g[i] - w[i]));
      return e;
    };
    const cxp::Box fdom({0, 0, 0}, {2 * n, 2 * n, 2 * n});
    DevFab nf(ffab, 2);
    nf.put(std::vector<double>(ffab.numPts() * 2, std::nan("")));
    cxp::Level nl{{nf.fab}, {fvalid}, fdom};
    cxp::WorkList wn;
    wn.append_new_level(P3, cl, 0, nl, 0, 2, &bc);
    wn.append_boundaries(nl, {0, 0, 0}, 0, 2, bc);
    wn.upload(s);
    wn.run(s);
    (void)hipStreamSynchronize(s);
    const double en = max_err(nf, ffab);
    std::printf("      new level, %lld points: max error %.3e\n",
                (long long)ffab.numPts() * 2, en);
    expect(en < 1e-12, "new-level fill with boundaries reproduces the data everywhere");

    const cxp::Box rvalid({4, 0, 2}, {19, 13, 15});
    const cxp::Box rfab = rvalid.grow(3);
    DevFab rf(rfab, 2);
    rf.put(std::vector<double>(rfab.numPts() * 2, std::nan("")));
    cxp::Level rl{{rf.fab}, {rvalid}, fdom};
    cxp::WorkList wr;
    wr.append_remake_level(P3, cl, 0, nl, rl, 0, 2, &bc);
    wr.append_boundaries(rl, {0, 0, 0}, 0, 2, bc);
    wr.upload(s);
    wr.run(s);
    (void)hipStreamSynchronize(s);
    const double er = max_err(rf, rfab);
    std::printf("      remake, %lld points: max error %.3e\n",
                (long long)rfab.numPts() * 2, er);
    expect(er < 1e-12, "remade-level fill with boundaries reproduces the data everywhere");
  }

  // 7. box-in-box tags and the threshold test
  {
    const int nc = 64;
    const cxp::Box cells({0, 0, 0}, {nc - 1, nc - 1, nc - 1});
    DevFab err(cells);
    cxp::BoxInBoxRegion cube;
    cube.num_levels = 2;
    cube.position = {0.5, 0.5, 0.5};
    cube.radius = {0.25, 0.25, 0.25};
    const double h = 1.0 / nc;
    cxp::boxinbox_setup({cube}, 1, {h / 2, h / 2, h / 2}, {h, h, h}, err.fab,
                        cells, s);
    signed char *tags = nullptr;
    int *words = nullptr;
    (void)hipMalloc(&tags, cells.numPts());
    (void)hipMalloc(&words, 7 * sizeof(int));
    cxp::error_est(err.fab, cells, 1.0, tags, words, s);
    int hw[7];
    std::vector<signed char> ht(cells.numPts());
    (void)hipMemcpyAsync(hw, words, sizeof hw, hipMemcpyDeviceToHost, s);
    (void)hipMemcpyAsync(ht.data(), tags, ht.size(), hipMemcpyDeviceToHost, s);
    (void)hipStreamSynchronize(s);
    const cxp::TagSummary sum(hw);
    const auto e = err.get();
    int64_t nt = 0, ne = 0;
    bool agree = true;
    for (size_t i = 0; i < ht.size(); ++i) {
      nt += ht[i] == CXP_TAG_SET;
      ne += e[i] == 1.0;
      agree = agree && (ht[i] == CXP_TAG_SET || ht[i] == CXP_TAG_CLEAR) &&
              (ht[i] == CXP_TAG_SET) == (e[i] == 1.0) && (e[i] == 0.0 || e[i] == 1.0);
    }
    std::printf("      tags: %lld set, box [%d,%d]^3\n", (long long)sum.count,
                sum.box.lo[0], sum.box.hi[0]);
    expect(sum.count == 30 * 30 * 30 && nt == sum.count && ne == sum.count && agree &&
               sum.box == cxp::Box({17, 17, 17}, {46, 46, 46}),
           "box-in-box cube tags exactly [17,46]^3, counted and boxed");
    (void)hipFree(tags);
    (void)hipFree(words);
  }

  (void)hipStreamDestroy(s);
  std::printf("%s\n", failures ? "FAILED" : "all checks passed");
  return failures ? 1 : 0;
}
