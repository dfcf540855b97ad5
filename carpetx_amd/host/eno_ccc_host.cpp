// Host build of the CCC ENO routines (carpetx_amd/csrc/eno_ccc.hpp): the
// workgroups of k_eno_ccc one after another, the threads of each phase one
// after another between its barriers, every window, |D2| and Mall access
// checked against its array and against the phase that wrote it, every
// coarse read against the coarse box and the window [parents - 2,
// parents + 2], every fine write against the region (each fine point
// written exactly once); the adjoint's per-coarse-point sum with its reads
// checked; and the tiled adjoint's workgroups (k_eno_ccc_adj_tiled) with
// every LDS, code and output access checked.  The CPU check of the kernels'
// indexing and arithmetic (tests/test_eno_ccc_host.py), not a product path.
#include <cstdio>
#include <vector>

#include "../csrc/eno_ccc.hpp"

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
      std::fprintf(stderr, "eno ccc host check: %s %lld %lld\n", what, a, b);
  }
};

// An LDS array: values with the phase that wrote them.
struct Arr {
  std::vector<double> v;
  std::vector<int> ph;
  explicit Arr(size_t n) : v(n, 0.0), ph(n, -1) {}
};

} // namespace

extern "C" {

// The CCC ENO forward over the region [rlo, rhi] from crse (box
// [clo, chi]): fine values into fine (the region's box) and each fine
// point's branch code into codes (the region's points).  Returns the number
// of failed checks.
long long cxp_host_eno_ccc_forward(const double *crse, const int clo[3],
                                   const int chi[3], double *fine,
                                   unsigned char *codes, const int rlo[3],
                                   const int rhi[3]) {
  Check ck;
  int plo[3], phi[3], olo[3], ohi[3];
  for (int d = 0; d < 3; ++d) {
    plo[d] = rlo[d] >> 1;
    phi[d] = rhi[d] >> 1;
    olo[d] = plo[d] - 2;
    ohi[d] = phi[d] + 2;
  }
  const int ntx = (phi[0] - plo[0] + EC_PX) / EC_PX;
  const int nty = (phi[1] - plo[1] + EC_PY) / EC_PY;
  const int ntz = (phi[2] - plo[2] + EC_PZ) / EC_PZ;
  long long npts = 1;
  for (int d = 0; d < 3; ++d)
    npts *= rhi[d] - rlo[d] + 1;
  std::vector<int> writes((size_t)npts, 0);
  auto cv = [&](int x, int y, int z) {
    if (!inside(clo, chi, x, y, z) || !inside(olo, ohi, x, y, z)) {
      ck.report("coarse read outside the box or window", x, z);
      return 0.0;
    }
    return crse[at(clo, chi, x, y, z)];
  };
  for (int b = 0; b < ntx * nty * ntz; ++b) {
    int p0[3];
    ec_tile(plo, b, ntx, nty, p0);
    Arr cw(EC_NC), d2(EC_ND), mall(EC_NM);
    int phase = 0;
    auto ld = [&](Arr &a, const char *name, int q) {
      if (q < 0 || q >= (int)a.v.size()) {
        ck.report(name, q, -1);
        return 0.0;
      }
      if (a.ph[q] < 0 || a.ph[q] >= phase)
        ck.report(name, q, a.ph[q]);
      return a.v[q];
    };
    auto st = [&](Arr &a, const char *name, int q, double v) {
      if (q < 0 || q >= (int)a.v.size()) {
        ck.report(name, q, -2);
        return;
      }
      if (a.ph[q] == phase)
        ck.report(name, q, -3);
      a.v[q] = v;
      a.ph[q] = phase;
    };
    for (int tid = 0; tid < EC_T; ++tid)
      for (int q = tid; q < EC_NC; q += EC_T)
        st(cw, "window store", q, ec_coarse(olo, ohi, p0, q, cv));
    ++phase;
    for (int tid = 0; tid < EC_T; ++tid)
      for (int e = tid; e < EC_ND; e += EC_T)
        st(d2, "d2 store", e,
           ec_d2(e, [&](int q) { return ld(cw, "window load (d2)", q); }));
    ++phase;
    for (int tid = 0; tid < EC_T; ++tid)
      for (int q = tid; q < EC_NM; q += EC_T)
        st(mall, "mall store", q,
           ec_mall(q, [&](int e) { return ld(d2, "d2 load", e); }));
    ++phase;
    for (int tid = 0; tid < EC_T; ++tid) {
      const int l[3] = {tid % EC_PX, (tid / EC_PX) % EC_PY,
                        tid / (EC_PX * EC_PY)};
      const int p[3] = {p0[0] + l[0], p0[1] + l[1], p0[2] + l[2]};
      if (p[0] > phi[0] || p[1] > phi[1] || p[2] > phi[2])
        continue;
      const int code =
          ec_search(l, [&](int q) { return ld(mall, "mall load", q); });
      double v[2][2][2];
      ec_children(l, code, [&](int q) { return ld(cw, "window load", q); },
                  v);
      for (int oz = 0; oz < 2; ++oz)
        for (int oy = 0; oy < 2; ++oy)
          for (int ox = 0; ox < 2; ++ox) {
            const int f[3] = {2 * p[0] + ox, 2 * p[1] + oy, 2 * p[2] + oz};
            if (!inside(rlo, rhi, f[0], f[1], f[2]))
              continue;
            const long long m = at(rlo, rhi, f[0], f[1], f[2]);
            ++writes[(size_t)m];
            fine[m] = v[oz][oy][ox];
            codes[m] = (unsigned char)code;
          }
    }
  }
  for (long long m = 0; m < npts; ++m)
    if (writes[(size_t)m] != 1)
      ck.report("fine point not written exactly once", m, writes[(size_t)m]);
  return ck.bad;
}

// cbar (box [clo, chi]) += J^T fbar (the region's box) from codes; returns
// the number of failed checks.
long long cxp_host_eno_ccc_adjoint(const unsigned char *codes,
                                   const double *fbar, const int rlo[3],
                                   const int rhi[3], double *cbar,
                                   const int clo[3], const int chi[3]) {
  Check ck;
  int plo[3], phi[3], olo[3], ohi[3];
  for (int d = 0; d < 3; ++d) {
    plo[d] = rlo[d] >> 1;
    phi[d] = rhi[d] >> 1;
    olo[d] = plo[d] - 2;
    ohi[d] = phi[d] + 2;
  }
  auto code = [&](int px, int py, int pz) -> int {
    if (px < plo[0] || px > phi[0] || py < plo[1] || py > phi[1] ||
        pz < plo[2] || pz > phi[2]) {
      ck.report("code of a parent outside the region", px, pz);
      return 13;
    }
    const int fx = 2 * px > rlo[0] ? 2 * px : rlo[0];
    const int fy = 2 * py > rlo[1] ? 2 * py : rlo[1];
    const int fz = 2 * pz > rlo[2] ? 2 * pz : rlo[2];
    if (!inside(rlo, rhi, fx, fy, fz)) {
      ck.report("code read outside the region", fx, fz);
      return 13;
    }
    return codes[at(rlo, rhi, fx, fy, fz)];
  };
  auto fb = [&](int x, int y, int z) {
    return inside(rlo, rhi, x, y, z) ? fbar[at(rlo, rhi, x, y, z)] : 0.0;
  };
  for (int k = olo[2]; k <= ohi[2]; ++k)
    for (int j = olo[1]; j <= ohi[1]; ++j)
      for (int i = olo[0]; i <= ohi[0]; ++i) {
        if (!inside(clo, chi, i, j, k)) {
          ck.report("coarse write outside cbar", i, k);
          continue;
        }
        const int q[3] = {i, j, k};
        cbar[at(clo, chi, i, j, k)] += ec_adj_one(q, plo, phi, code, fb);
      }
  return ck.bad;
}

// The tiled adjoint's workgroups (k_eno_ccc_adj_tiled) one after another,
// z-chunks of zc planes: cbar (box [clo, chi]) += J^T fbar from codes.
// Every LDS read is checked against the plane whose phase L wrote it, every
// code read against the region, and every output point is written exactly
// once.  Returns the number of failed checks.
long long cxp_host_eno_ccc_adjoint_tiled(const unsigned char *codes,
                                         const double *fbar, const int rlo[3],
                                         const int rhi[3], double *cbar,
                                         const int clo[3], const int chi[3],
                                         int zc) {
  Check ck;
  const int kNone = -(1 << 30);  // plane stamp of a slot never written
  int plo[3], phi[3], olo[3], ohi[3];
  for (int d = 0; d < 3; ++d) {
    plo[d] = rlo[d] >> 1;
    phi[d] = rhi[d] >> 1;
    olo[d] = plo[d] - 2;
    ohi[d] = phi[d] + 2;
  }
  const int ntx = (ohi[0] - olo[0] + EA_TX) / EA_TX;
  const int nty = (ohi[1] - olo[1] + EA_TY) / EA_TY;
  const int ntz = (ohi[2] - olo[2] + zc) / zc;
  std::vector<int> writes((size_t)at(olo, ohi, ohi[0], ohi[1], ohi[2]) + 1, 0);
  auto code = [&](int px, int py, int pz) -> int {
    if (px < plo[0] || px > phi[0] || py < plo[1] || py > phi[1] ||
        pz < plo[2] || pz > phi[2]) {
      ck.report("code of a parent outside the region", px, pz);
      return 13;
    }
    const int fx = 2 * px > rlo[0] ? 2 * px : rlo[0];
    const int fy = 2 * py > rlo[1] ? 2 * py : rlo[1];
    const int fz = 2 * pz > rlo[2] ? 2 * pz : rlo[2];
    if (!inside(rlo, rhi, fx, fy, fz)) {
      ck.report("code read outside the region", fx, fz);
      return 13;
    }
    return codes[at(rlo, rhi, fx, fy, fz)];
  };
  auto kids = [&](int px, int py, int pz, double (&f)[2][2][2]) {
    if (px < plo[0] || px > phi[0] || py < plo[1] || py > phi[1] ||
        pz < plo[2] || pz > phi[2])
      ck.report("children of a parent outside the region", px, pz);
    for (int oz = 0; oz < 2; ++oz)
      for (int oy = 0; oy < 2; ++oy)
        for (int ox = 0; ox < 2; ++ox) {
          const int x = 2 * px + ox, y = 2 * py + oy, z = 2 * pz + oz;
          f[oz][oy][ox] =
              inside(rlo, rhi, x, y, z) ? fbar[at(rlo, rhi, x, y, z)] : 0.0;
        }
  };
  for (int b = 0; b < ntx * nty * ntz; ++b) {
    int q0[3];
    ea_tile(olo, b, ntx, nty, zc, q0);
    const int qz1 = q0[2] + zc - 1 < ohi[2] ? q0[2] + zc - 1 : ohi[2];
    std::vector<unsigned> cs(EA_NP, 0x5A5A5A5Au);
    std::vector<int> cs_ph(EA_NP, kNone);
    std::vector<double> hs(EA_NH * (EA_NP + 1), 0.0);
    std::vector<int> hs_ph(9 * (EA_NP + 1), kNone);
    std::vector<double> acc((size_t)EA_NT * 5, 0.0);
    for (int pz = q0[2] - 2; pz <= qz1 + 2; ++pz) {
      for (int tid = 0; tid < EA_NT; ++tid)
        for (int n = tid; n < EA_NP; n += EA_NT) {
          double h[EA_NH], f[2][2][2];
          const int code_n = ea_load(q0, n, pz, plo, phi, code, kids, f);
          const unsigned c = ea_sums(code_n, f, ec_w, h);
          if (cs_ph[n] == pz)
            ck.report("parent slot written twice in a plane", n, pz);
          cs[n] = c;
          cs_ph[n] = pz;
          if (c != 0u)
            for (int k = 0; k < 9; ++k) {
              hs[2 * (k * (EA_NP + 1) + n)] = h[k];
              hs[2 * (k * (EA_NP + 1) + n) + 1] = h[9 + k];
              hs_ph[k * (EA_NP + 1) + n] = pz;
            }
        }
      for (int tid = 0; tid < EA_NT; ++tid) {
        const int lx = tid % EA_TX, ly = tid / EA_TX;
        double a5[5];
        for (int u = 0; u < 5; ++u)
          a5[u] = acc[tid * 5 + u];
        ea_gather(
            lx, ly,
            [&](int n) {
              if (n < 0 || n >= EA_NP || cs_ph[n] != pz) {
                ck.report("info slot read outside the tile or stale", n, pz);
                return 0u;
              }
              return cs[n];
            },
            [&](int b, double &h0, double &h1) {
              h0 = h1 = 0.0;
              if (b % 16 != 0) {
                ck.report("H pair offset not a pair", b, pz);
                return;
              }
              const int i = b / 16;
              if (i < 0 || i >= 9 * (EA_NP + 1)) {
                ck.report("H pair read outside the array", i, pz);
                return;
              }
              // pair EA_NP: the zero pair, never written
              if (i != EA_NP && hs_ph[i] != pz) {
                ck.report("stale H pair read", i, pz);
                return;
              }
              h0 = hs[2 * i];
              h1 = hs[2 * i + 1];
            },
            a5);
        const int q[3] = {q0[0] + lx, q0[1] + ly, pz - 2};
        if (q[0] <= ohi[0] && q[1] <= ohi[1] && q[2] >= q0[2]) {
          if (!inside(olo, ohi, q[0], q[1], q[2]) ||
              !inside(clo, chi, q[0], q[1], q[2])) {
            ck.report("coarse write outside the window or cbar", q[0], q[2]);
          } else {
            ++writes[(size_t)at(olo, ohi, q[0], q[1], q[2])];
            cbar[at(clo, chi, q[0], q[1], q[2])] += a5[0];
          }
        }
        acc[tid * 5 + 0] = a5[1];
        acc[tid * 5 + 1] = a5[2];
        acc[tid * 5 + 2] = a5[3];
        acc[tid * 5 + 3] = a5[4];
        acc[tid * 5 + 4] = 0.0;
      }
    }
  }
  for (size_t m = 0; m < writes.size(); ++m)
    if (writes[m] != 1)
      ck.report("output point not written exactly once", (long long)m, writes[m]);
  return ck.bad;
}

} // extern "C"
