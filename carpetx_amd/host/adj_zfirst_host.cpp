// Host build of the z-first linear adjoint (carpetx_amd/csrc/adj_zfirst.hpp):
// the workgroups of k_adj_zf one after another, the threads of each phase
// one after another between its barriers.  Every fine read is checked
// against the fine window of its tile and the region, every LDS read
// against the plane whose phase wrote it, every coarse write against the
// output window, and every output point is written exactly once.  The CPU
// check of the kernel's indexing and arithmetic (tests/test_adj_zfirst_host.py),
// not a product path.
#include <cstdio>
#include <vector>

#include "../csrc/adj_zfirst.hpp"

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

struct Check {
  long long bad = 0;
  void report(const char *what, long long a, long long b) {
    if (bad++ < 8)
      std::fprintf(stderr, "adj zfirst host check: %s %lld %lld\n", what, a, b);
  }
};

template <int S>
long long run(const double *fbar, const int rlo[3], const int rhi[3],
              double *cbar, const int clo[3], const int chi[3], int kchunk) {
  using G = AzGeom<S>;
  Check ck;
  const int kNone = -(1 << 30);
  int olo[3], ohi[3];
  for (int d = 0; d < 3; ++d) {
    olo[d] = (rlo[d] >> 1) + G::wlo;
    ohi[d] = (rhi[d] >> 1) + G::whi;
    if (olo[d] < clo[d] || ohi[d] > chi[d]) {
      ck.report("output window outside cbar", d, olo[d]);
      return ck.bad;
    }
  }
  const int ntx = (ohi[0] - olo[0] + AZ_T) / AZ_T;
  const int nty = (ohi[1] - olo[1] + AZ_T) / AZ_T;
  const int nk = ohi[2] - olo[2] + 1;
  const int nkc = (nk + kchunk - 1) / kchunk;
  std::vector<int> writes((size_t)at(olo, ohi, ohi[0], ohi[1], ohi[2]) + 1, 0);
  for (int b = 0; b < ntx * nty * nkc; ++b) {
    int i0, j0, kca;
    az_tile(olo, b, ntx, nty, kchunk, i0, j0, kca);
    const int kcb = kca + kchunk - 1 < ohi[2] ? kca + kchunk - 1 : ohi[2];
    const int fxa = 2 * (i0 - G::whi), fya = 2 * (j0 - G::whi);
    // the tile's fine window in x and y, all z
    auto fb = [&](int c, int fz) -> double {
      if (c < 0 || c >= G::NZ) {
        ck.report("column outside the window", c, fz);
        return 0.0;
      }
      const int gx = fxa + c % G::W, gy = fya + c / G::W;
      return inside(rlo, rhi, gx, gy, fz) ? fbar[at(rlo, rhi, gx, gy, fz)] : 0.0;
    };
    std::vector<double> z((size_t)AZ_NT * G::R * (G::span + 1), 0.0);
    std::vector<double> zs(G::NZ, 0.0), ys(G::NY, 0.0);
    std::vector<int> zs_ph(G::NZ, kNone), ys_ph(G::NY, kNone);
    for (int pz = kca - G::whi; pz <= kcb - G::wlo; ++pz) {
      for (int tid = 0; tid < AZ_NT; ++tid)
        for (int r = 0; r < G::R; ++r) {
          const int c = tid + AZ_NT * r;
          double w[G::span + 1];
          double *zz = &z[((size_t)tid * G::R + r) * (G::span + 1)];
          for (int q = 0; q <= G::span; ++q)
            w[q] = zz[q];
          const double f0 = c < G::NZ ? fb(c, 2 * pz) : 0.0;
          const double f1 = c < G::NZ ? fb(c, 2 * pz + 1) : 0.0;
          az_zstep<S>(w, f0, f1);
          for (int q = 0; q <= G::span; ++q)
            zz[q] = w[q];
        }
      const int k = pz + G::wlo;
      if (k >= kca) {
        for (int tid = 0; tid < AZ_NT; ++tid)
          for (int r = 0; r < G::R; ++r) {
            const int c = tid + AZ_NT * r;
            if (c < G::NZ) {
              zs[c] = z[((size_t)tid * G::R + r) * (G::span + 1)];
              zs_ph[c] = k;
            }
          }
        for (int tid = 0; tid < AZ_NT; ++tid)
          for (int e = tid; e < G::NY; e += AZ_NT) {
            ys[e] = az_y<S>(e, [&](int o) {
              if (o < 0 || o >= G::NZ || zs_ph[o] != k) {
                ck.report("Zs read outside or stale", o, k);
                return 0.0;
              }
              return zs[o];
            });
            ys_ph[e] = k;
          }
        for (int tid = 0; tid < AZ_NT; ++tid) {
          const int i = tid % AZ_T, j = tid / AZ_T;
          const double x = az_x<S>(i, j, [&](int o) {
            if (o < 0 || o >= G::NY || ys_ph[o] != k) {
              ck.report("Ys read outside or stale", o, k);
              return 0.0;
            }
            return ys[o];
          });
          if (i0 + i <= ohi[0] && j0 + j <= ohi[1]) {
            if (!inside(olo, ohi, i0 + i, j0 + j, k)) {
              ck.report("output outside the window", i0 + i, k);
              continue;
            }
            ++writes[(size_t)at(olo, ohi, i0 + i, j0 + j, k)];
            cbar[at(clo, chi, i0 + i, j0 + j, k)] += x;
          }
        }
      }
      for (int tid = 0; tid < AZ_NT; ++tid)
        for (int r = 0; r < G::R; ++r) {
          double *zz = &z[((size_t)tid * G::R + r) * (G::span + 1)];
          for (int q = 0; q < G::span; ++q)
            zz[q] = zz[q + 1];
          zz[G::span] = 0.0;
        }
    }
  }
  for (size_t m = 0; m < writes.size(); ++m)
    if (writes[m] != 1)
      ck.report("output point not written exactly once", (long long)m, writes[m]);
  return ck.bad;
}

} // namespace

extern "C" {

// cbar (box [clo, chi]) += P^T fbar (the region's box) for the operator
// whose three directions use stencil sid, z-chunks of kchunk planes.
// Returns the number of failed checks, or -1 for a stencil without the
// z-first kernel.
long long cxp_host_adj_zfirst(int sid, const double *fbar, const int rlo[3],
                              const int rhi[3], double *cbar, const int clo[3],
                              const int chi[3], int kchunk) {
  switch (sid) {
#define CXP_HZ(S)                                                              \
  case S:                                                                      \
    return run<S>(fbar, rlo, rhi, cbar, clo, chi, kchunk);
    CXP_HZ(SID_VPOLY1) CXP_HZ(SID_VPOLY3) CXP_HZ(SID_VPOLY5) CXP_HZ(SID_VPOLY7)
    CXP_HZ(SID_CPOLY0) CXP_HZ(SID_CPOLY1) CXP_HZ(SID_CPOLY2) CXP_HZ(SID_CPOLY3)
    CXP_HZ(SID_CPOLY4) CXP_HZ(SID_CPOLY5) CXP_HZ(SID_CPOLY6) CXP_HZ(SID_CPOLY7)
    CXP_HZ(SID_VCONS0) CXP_HZ(SID_VCONS1) CXP_HZ(SID_CCONS0) CXP_HZ(SID_CCONS1)
    CXP_HZ(SID_CCONS2) CXP_HZ(SID_CCONS3) CXP_HZ(SID_CCONS4) CXP_HZ(SID_CCONS5)
    CXP_HZ(SID_CCONS6) CXP_HZ(SID_CCONS7)
#undef CXP_HZ
  }
  return -1;
}

} // extern "C"
