// CCC ENO kernels (routines and their description in eno_ccc.hpp) and
// their launches (eno_launch.hpp):
//  * k_eno_ccc<CODES, PARENT>  forward: values (and, given a code array,
//                      each fine point's branch code beside them), or with
//                      CODES the branch codes alone, per fine point or (with
//                      PARENT, for the adjoint that recomputes them) one per
//                      coarse parent;
//  * k_eno_ccc_adj_tiled<PARENT>  crse_bar += J^T fine_bar from the branch
//                      codes of either layout.
// One component per grid row y.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <type_traits>

#include "eno_ccc.hpp"
#include "eno_launch.hpp"
#include "xcd.hpp"

namespace cxp {

namespace {

struct EcArgs {
  const double *crse;
  double *fine;
  long long c_sy, c_sz, f_sy, f_sz, ccs, fcs;
  int c_lo[3], f_lo[3];
  int rlo[3], rhi[3], plo[3], phi[3], olo[3], ohi[3];
  int ntx, nty;
  long long npts;  // fine points of the region (codes per component)
  long long npar;  // coarse parents of the region (compact codes per component)
};

// CODES: branch codes alone (no values); PARENT: one code per coarse
// parent, parent-linear over [plo, phi] (else one per fine point of the
// region).
template <bool CODES, bool PARENT = false>
__global__ void __launch_bounds__(EcGeom<PARENT ? EC_PZ_CODES : EC_PZ>::T)
    k_eno_ccc(EcArgs a, unsigned char *__restrict__ codes) {
  constexpr int PZ = PARENT ? EC_PZ_CODES : EC_PZ;
  using Gm = EcGeom<PZ>;
  // With PARENT (no children to compute) Mall is built one direction at a
  // time through one |D2| array of the largest direction's size: 33 KB of
  // LDS instead of 47.5 KB, four workgroups per compute unit instead of
  // three.
  constexpr int NDMAX = Gm::NDX > Gm::NDY
                            ? (Gm::NDX > Gm::NDZ ? Gm::NDX : Gm::NDZ)
                            : (Gm::NDY > Gm::NDZ ? Gm::NDY : Gm::NDZ);
  __shared__ double lds[Gm::NC + (PARENT ? NDMAX : Gm::ND) + Gm::NM];
  double *const cw = lds;
  double *const d2 = lds + Gm::NC;
  double *const mall = lds + Gm::NC + (PARENT ? NDMAX : Gm::ND);
  const int tid = threadIdx.x;
  const int comp = blockIdx.y;
  int p0[3];
  ec_tile<PZ>(a.plo, xcd_block(blockIdx.x, gridDim.x), a.ntx, a.nty, p0);
  const double *cp = a.crse + comp * a.ccs;
  auto cv = [&](int x, int y, int z) {
    return cp[(x - a.c_lo[0]) + (y - a.c_lo[1]) * a.c_sy +
              (z - a.c_lo[2]) * a.c_sz];
  };
  auto cl = [&](int q) { return cw[q]; };
  auto dl = [&](int e) { return d2[e]; };
  auto ml = [&](int q) { return mall[q]; };
  // A
  {
    double v[(Gm::NC + Gm::T - 1) / Gm::T];
#pragma unroll
    for (int n = 0; n < (Gm::NC + Gm::T - 1) / Gm::T; ++n) {
      const int q = tid + Gm::T * n;
      v[n] = q < Gm::NC ? ec_coarse(a.olo, a.ohi, p0, q, cv) : 0.0;
    }
#pragma unroll
    for (int n = 0; n < (Gm::NC + Gm::T - 1) / Gm::T; ++n) {
      const int q = tid + Gm::T * n;
      if (q < Gm::NC)
        cw[q] = v[n];
    }
  }
  __syncthreads();
  if constexpr (PARENT) {
    // B, C per direction: Mall = fmax(fmax(Mx, My), Mz), each thread
    // updating its own Mall points (the chain ec_mall takes)
    auto pass = [&](auto dirc, int off, int nd) {
      constexpr int DIR = decltype(dirc)::value;
      for (int e = tid; e < nd; e += Gm::T)
        d2[e] = ec_d2<PZ>(off + e, cl);
      __syncthreads();
      for (int q = tid; q < Gm::NM; q += Gm::T) {
        const double m = ec_mall_dir<DIR>(q, dl);
        mall[q] = DIR == 0 ? m : fmax(mall[q], m);
      }
      __syncthreads();
    };
    pass(std::integral_constant<int, 0>{}, 0, Gm::NDX);
    pass(std::integral_constant<int, 1>{}, Gm::NDX, Gm::NDY);
    pass(std::integral_constant<int, 2>{}, Gm::NDX + Gm::NDY, Gm::NDZ);
  } else {
    // B
    for (int e = tid; e < Gm::ND; e += Gm::T)
      d2[e] = ec_d2<PZ>(e, cl);
    __syncthreads();
    // C
    for (int q = tid; q < Gm::NM; q += Gm::T)
      mall[q] = ec_mall<PZ>(q, dl);
    __syncthreads();
  }
  // D, E: one parent per thread
  const int l[3] = {tid % EC_PX, (tid / EC_PX) % EC_PY, tid / (EC_PX * EC_PY)};
  const int p[3] = {p0[0] + l[0], p0[1] + l[1], p0[2] + l[2]};
  if (p[0] > a.phi[0] || p[1] > a.phi[1] || p[2] > a.phi[2])
    return;
  const int code = ec_search(l, ml);
  if constexpr (PARENT) {
    const long long npx = a.phi[0] - a.plo[0] + 1, npy = a.phi[1] - a.plo[1] + 1;
    codes[comp * a.npar + (p[0] - a.plo[0]) + (p[1] - a.plo[1]) * npx +
          (p[2] - a.plo[2]) * npx * npy] = (unsigned char)code;
    return;
  }
  double *fp = a.fine + comp * a.fcs;
  unsigned char *cd = codes ? codes + comp * a.npts : nullptr;
  const long long rnx = a.rhi[0] - a.rlo[0] + 1;
  const long long rny = a.rhi[1] - a.rlo[1] + 1;
  // child (ox, oy, oz): its value (unless CODES) and, given codes, its code
  auto put = [&](int ox, int oy, int oz, double val) {
    const int f[3] = {2 * p[0] + ox, 2 * p[1] + oy, 2 * p[2] + oz};
    if (f[0] < a.rlo[0] || f[0] > a.rhi[0] || f[1] < a.rlo[1] ||
        f[1] > a.rhi[1] || f[2] < a.rlo[2] || f[2] > a.rhi[2])
      return;
    if constexpr (!CODES)
      fp[(f[0] - a.f_lo[0]) + (f[1] - a.f_lo[1]) * a.f_sy +
         (f[2] - a.f_lo[2]) * a.f_sz] = val;
    if (cd)
      cd[(f[0] - a.rlo[0]) + (f[1] - a.rlo[1]) * rnx +
         (f[2] - a.rlo[2]) * rnx * rny] = (unsigned char)code;
  };
  if constexpr (CODES) {
    for (int oz = 0; oz < 2; ++oz)
      for (int oy = 0; oy < 2; ++oy)
        for (int ox = 0; ox < 2; ++ox)
          put(ox, oy, oz, 0.0);
  } else {
    ec_children_each(l, code, cl, put);
  }
}

// The tiled adjoint (eno_ccc.hpp, ea_*): crse_bar += J^T fine_bar over the
// output window [olo, ohi], tiles numbered so each XCD runs a contiguous
// range; zc output planes per workgroup.  Thread t loads parent slots t and
// t + EA_NT (EA_NP <= 2 EA_NT); the next plane's loads are issued before
// the gather of the current one, so their latency overlaps it.
template <bool PARENT>
__global__ void __launch_bounds__(EA_NT, 2)
    k_eno_ccc_adj_tiled(EcArgs a, const unsigned char *__restrict__ codes,
                        int ntx, int nty, int zc) {
  static_assert(EA_NP <= 2 * EA_NT, "two parent slots per thread");
  // H pairs: pair k (EA_NP + 1) + n = (H[0][k], H[1][k]) of slot n; pair
  // EA_NP, zeros, is read for the classes a slot does not add to
  __shared__ double hs[EA_NH * (EA_NP + 1)];
  __shared__ unsigned cs[EA_NP];
  __shared__ double ws[18];
  const int tid = threadIdx.x;
  const int comp = blockIdx.y;
  int q0[3];
  ea_tile(a.olo, xcd_block(blockIdx.x, gridDim.x), ntx, nty, zc, q0);
  if (tid < 18)
    ws[tid] = ec_w(tid / 6 - 1, (tid / 3) % 2, tid % 3 - 1);
  if (tid < 2)
    hs[2 * EA_NP + tid] = 0.0;
  const unsigned char *cd = codes + comp * (PARENT ? a.npar : a.npts);
  const double *fp = a.fine + comp * a.fcs;
  const long long rnx = a.rhi[0] - a.rlo[0] + 1;
  const long long rny = a.rhi[1] - a.rlo[1] + 1;
  const long long npx = a.phi[0] - a.plo[0] + 1, npy = a.phi[1] - a.plo[1] + 1;
  // per fine point: a parent's code sits at each of its fine points in the
  // region, read the first; per parent: parent-linear
  auto code = [&](int px, int py, int pz) -> int {
    if constexpr (PARENT)
      return cd[(px - a.plo[0]) + (py - a.plo[1]) * npx +
                (pz - a.plo[2]) * npx * npy];
    const int fx = max(2 * px, a.rlo[0]), fy = max(2 * py, a.rlo[1]),
              fz = max(2 * pz, a.rlo[2]);
    return cd[(fx - a.rlo[0]) + (fy - a.rlo[1]) * rnx +
              (fz - a.rlo[2]) * rnx * rny];
  };
  // the eight children of a parent in the region: one base offset, the
  // children outside the region masked to 0
  auto kids = [&](int px, int py, int pz, double (&f)[2][2][2]) {
    const int x = 2 * px, y = 2 * py, z = 2 * pz;
    const bool xo[2] = {x >= a.rlo[0], x + 1 <= a.rhi[0]};
    const bool yo[2] = {y >= a.rlo[1], y + 1 <= a.rhi[1]};
    const bool zo[2] = {z >= a.rlo[2], z + 1 <= a.rhi[2]};
    const double *b = fp + (x - a.f_lo[0]) + (y - a.f_lo[1]) * a.f_sy +
                      (z - a.f_lo[2]) * a.f_sz;
#pragma unroll
    for (int oz = 0; oz < 2; ++oz)
#pragma unroll
      for (int oy = 0; oy < 2; ++oy)
#pragma unroll
        for (int ox = 0; ox < 2; ++ox)
          f[oz][oy][ox] = (xo[ox] && yo[oy] && zo[oz])
                              ? b[ox + oy * a.f_sy + oz * a.f_sz]
                              : 0.0;
  };
  auto w = [&](int sh, int o, int t) { return ws[(sh + 1) * 6 + o * 3 + t + 1]; };
  const int lx = tid % EA_TX, ly = tid / EA_TX;
  const int qx = q0[0] + lx, qy = q0[1] + ly;
  const int qz1 = min(q0[2] + zc - 1, a.ohi[2]);
  const bool own = qx <= a.ohi[0] && qy <= a.ohi[1];
  const bool two = tid + EA_NT < EA_NP;
  double *bp = const_cast<double *>(a.crse) + comp * a.ccs +
               (qx - a.c_lo[0]) + (qy - a.c_lo[1]) * a.c_sy;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  double f0[2][2][2], f1[2][2][2];
  int c0 = ea_load(q0, tid, q0[2] - 2, a.plo, a.phi, code, kids, f0);
  int c1 = two ? ea_load(q0, tid + EA_NT, q0[2] - 2, a.plo, a.phi, code, kids, f1)
               : -1;
  for (int pz = q0[2] - 2; pz <= qz1 + 2; ++pz) {
    {
      double h[EA_NH];
      unsigned info = ea_sums(c0, f0, w, h);
      cs[tid] = info;
      if (info) {
#pragma unroll
        for (int k = 0; k < 9; ++k) {
          hs[2 * (k * (EA_NP + 1) + tid)] = h[k];
          hs[2 * (k * (EA_NP + 1) + tid) + 1] = h[9 + k];
        }
      }
      if (two) {
        info = ea_sums(c1, f1, w, h);
        cs[tid + EA_NT] = info;
        if (info) {
#pragma unroll
          for (int k = 0; k < 9; ++k) {
            hs[2 * (k * (EA_NP + 1) + tid + EA_NT)] = h[k];
            hs[2 * (k * (EA_NP + 1) + tid + EA_NT) + 1] = h[9 + k];
          }
        }
      }
    }
    __syncthreads();
    if (pz < qz1 + 2) {
      c0 = ea_load(q0, tid, pz + 1, a.plo, a.phi, code, kids, f0);
      c1 = two ? ea_load(q0, tid + EA_NT, pz + 1, a.plo, a.phi, code, kids, f1)
               : -1;
    }
    ea_gather(
        lx, ly, [&](int n) { return cs[n]; },
        [&](int b, double &h0, double &h1) {
          const double *q = (const double *)((const char *)hs + b);
          h0 = q[0];
          h1 = q[1];
        },
        acc);
    const int qz = pz - 2;
    if (own && qz >= q0[2])
      bp[(long long)(qz - a.c_lo[2]) * a.c_sz] += acc[0];
    acc[0] = acc[1];
    acc[1] = acc[2];
    acc[2] = acc[3];
    acc[3] = acc[4];
    acc[4] = 0.0;
    __syncthreads();  // the next plane rewrites cs and hs
  }
}

EcArgs ec_args(const cxp_fab *crse, int crse_comp, const cxp_fab *fine,
               int fine_comp, const int rlo[3], const int rhi[3]) {
  EcArgs a{};
  a.crse = crse->data + (long long)crse_comp * crse->comp_stride;
  a.fine = fine->data + (long long)fine_comp * fine->comp_stride;
  a.c_sy = crse->hi[0] - crse->lo[0] + 1;
  a.c_sz = a.c_sy * (crse->hi[1] - crse->lo[1] + 1);
  a.f_sy = fine->hi[0] - fine->lo[0] + 1;
  a.f_sz = a.f_sy * (fine->hi[1] - fine->lo[1] + 1);
  a.ccs = crse->comp_stride;
  a.fcs = fine->comp_stride;
  a.npts = 1;
  for (int d = 0; d < 3; ++d) {
    a.c_lo[d] = crse->lo[d];
    a.f_lo[d] = fine->lo[d];
    a.rlo[d] = rlo[d];
    a.rhi[d] = rhi[d];
    a.plo[d] = rlo[d] >> 1;
    a.phi[d] = rhi[d] >> 1;
    a.olo[d] = a.plo[d] - 2;  // CoarseBox: the parents +- the ENO reach 2
    a.ohi[d] = a.phi[d] + 2;
    a.npts *= rhi[d] - rlo[d] + 1;
  }
  a.npar = (long long)(a.phi[0] - a.plo[0] + 1) * (a.phi[1] - a.plo[1] + 1) *
           (a.phi[2] - a.plo[2] + 1);
  a.ntx = (a.phi[0] - a.plo[0] + EC_PX) / EC_PX;
  a.nty = (a.phi[1] - a.plo[1] + EC_PY) / EC_PY;
  return a;
}

// Workgroups of k_eno_ccc_adj_tiled resident at once on the current
// device (CUs x workgroups per CU), cached per device.
long long ea_slots() {
  static std::mutex mu;
  static std::map<int, long long> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess)
    return 512;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(dev);
  if (it != cache.end())
    return it->second;
  int cus = 0, per_cu = 0;
  long long n = 512;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                            dev) == hipSuccess &&
      hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per_cu, k_eno_ccc_adj_tiled<false>, EA_NT, 0) == hipSuccess &&
      cus > 0 && per_cu > 0)
    n = (long long)cus * per_cu;
  cache[dev] = n;
  return n;
}

} // namespace

bool eno_ccc_fits(const OpDesc &op, int per_group) {
  return !per_group && op.sid[0] == SID_ENO2 && op.sid[1] == SID_ENO2 &&
         op.sid[2] == SID_ENO2 && !op_fallback_per_var(op);
}

int eno_ccc_launch(const cxp_fab *crse, int crse_comp, const cxp_fab *fine,
                   int fine_comp, int ncomp, const int rlo[3],
                   const int rhi[3], unsigned char *codes, bool codes_only,
                   hipStream_t s, bool parent_codes) {
  const EcArgs a = ec_args(crse, crse_comp, fine ? fine : crse,
                           fine ? fine_comp : 0, rlo, rhi);
  if (codes_only && parent_codes) {
    const int ntz = (a.phi[2] - a.plo[2] + EC_PZ_CODES) / EC_PZ_CODES;
    const dim3 grid((unsigned)(a.ntx * a.nty * ntz), ncomp);
    hipLaunchKernelGGL((k_eno_ccc<true, true>), grid,
                       dim3(EcGeom<EC_PZ_CODES>::T), 0, s, a, codes);
    return hipGetLastError() == hipSuccess ? CXP_OK : CXP_ERR_HIP;
  }
  const int ntz = (a.phi[2] - a.plo[2] + EC_PZ) / EC_PZ;
  const dim3 grid((unsigned)(a.ntx * a.nty * ntz), ncomp);
  if (codes_only)
    hipLaunchKernelGGL(k_eno_ccc<true>, grid, dim3(EC_T), 0, s, a, codes);
  else
    hipLaunchKernelGGL(k_eno_ccc<false>, grid, dim3(EC_T), 0, s, a, codes);
  return hipGetLastError() == hipSuccess ? CXP_OK : CXP_ERR_HIP;
}

int eno_ccc_adj_launch(const cxp_fab *crse_bar, int crse_comp,
                       const cxp_fab *fine_bar, int fine_comp, int ncomp,
                       const int rlo[3], const int rhi[3],
                       const unsigned char *codes, hipStream_t s,
                       bool parent_codes) {
  const EcArgs a =
      ec_args(crse_bar, crse_comp, fine_bar, fine_comp, rlo, rhi);
  const int ntx = (a.ohi[0] - a.olo[0] + EA_TX) / EA_TX;
  const int nty = (a.ohi[1] - a.olo[1] + EA_TY) / EA_TY;
  // As many z-chunks as fill the device's resident workgroups once (each
  // chunk marches 4 parent planes more than its outputs), at least 4 planes
  // per chunk.
  const long long nk = a.ohi[2] - a.olo[2] + 1;
  const long long cols = (long long)ntx * nty * ncomp;
  long long nkc = ea_slots() / cols;
  nkc = nkc < 1 ? 1 : (nkc > nk ? nk : nkc);
  long long zc = (nk + nkc - 1) / nkc;
  zc = zc < 4 ? (nk < 4 ? nk : 4) : zc;
  const long long ntz = (nk + zc - 1) / zc;
  const dim3 grid((unsigned)(ntx * nty * ntz), ncomp);
  if (parent_codes)
    hipLaunchKernelGGL(k_eno_ccc_adj_tiled<true>, grid, dim3(EA_NT), 0, s, a,
                       codes, ntx, nty, (int)zc);
  else
    hipLaunchKernelGGL(k_eno_ccc_adj_tiled<false>, grid, dim3(EA_NT), 0, s, a,
                       codes, ntx, nty, (int)zc);
  return hipGetLastError() == hipSuccess ? CXP_OK : CXP_ERR_HIP;
}

} // namespace cxp
