// CCC ENO kernels (routines and their description in eno_ccc.hpp) and
// their launches (eno_launch.hpp):
//  * k_eno_ccc<CODES>  forward: values (and, given a code array, each fine
//                      point's branch code beside them), or with CODES the
//                      branch codes alone;
//  * k_eno_ccc_adj     crse_bar += J^T fine_bar from the branch codes.
// One component per grid row y.
#include <hip/hip_runtime.h>

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
};

template <bool CODES>
__global__ void __launch_bounds__(EC_T)
    k_eno_ccc(EcArgs a, unsigned char *__restrict__ codes) {
  __shared__ double cw[EC_NC];
  __shared__ double d2[EC_ND];
  __shared__ double mall[EC_NM];
  const int tid = threadIdx.x;
  const int comp = blockIdx.y;
  int p0[3];
  ec_tile(a.plo, xcd_block(blockIdx.x, gridDim.x), a.ntx, a.nty, p0);
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
    double v[(EC_NC + EC_T - 1) / EC_T];
#pragma unroll
    for (int n = 0; n < (EC_NC + EC_T - 1) / EC_T; ++n) {
      const int q = tid + EC_T * n;
      v[n] = q < EC_NC ? ec_coarse(a.olo, a.ohi, p0, q, cv) : 0.0;
    }
#pragma unroll
    for (int n = 0; n < (EC_NC + EC_T - 1) / EC_T; ++n) {
      const int q = tid + EC_T * n;
      if (q < EC_NC)
        cw[q] = v[n];
    }
  }
  __syncthreads();
  // B
  for (int e = tid; e < EC_ND; e += EC_T)
    d2[e] = ec_d2(e, cl);
  __syncthreads();
  // C
  for (int q = tid; q < EC_NM; q += EC_T)
    mall[q] = ec_mall(q, dl);
  __syncthreads();
  // D, E: one parent per thread
  const int l[3] = {tid % EC_PX, (tid / EC_PX) % EC_PY, tid / (EC_PX * EC_PY)};
  const int p[3] = {p0[0] + l[0], p0[1] + l[1], p0[2] + l[2]};
  if (p[0] > a.phi[0] || p[1] > a.phi[1] || p[2] > a.phi[2])
    return;
  const int code = ec_search(l, ml);
  double v[2][2][2];
  if constexpr (!CODES)
    ec_children(l, code, cl, v);
  double *fp = a.fine + comp * a.fcs;
  unsigned char *cd = codes ? codes + comp * a.npts : nullptr;
  const long long rnx = a.rhi[0] - a.rlo[0] + 1;
  const long long rny = a.rhi[1] - a.rlo[1] + 1;
  for (int oz = 0; oz < 2; ++oz)
    for (int oy = 0; oy < 2; ++oy)
      for (int ox = 0; ox < 2; ++ox) {
        const int f[3] = {2 * p[0] + ox, 2 * p[1] + oy, 2 * p[2] + oz};
        if (f[0] < a.rlo[0] || f[0] > a.rhi[0] || f[1] < a.rlo[1] ||
            f[1] > a.rhi[1] || f[2] < a.rlo[2] || f[2] > a.rhi[2])
          continue;
        if constexpr (!CODES)
          fp[(f[0] - a.f_lo[0]) + (f[1] - a.f_lo[1]) * a.f_sy +
             (f[2] - a.f_lo[2]) * a.f_sz] = v[oz][oy][ox];
        if (cd)
          cd[(f[0] - a.rlo[0]) + (f[1] - a.rlo[1]) * rnx +
             (f[2] - a.rlo[2]) * rnx * rny] = (unsigned char)code;
      }
}

// One thread per coarse point of the output window [olo, ohi], x fastest.
__global__ void __launch_bounds__(256)
    k_eno_ccc_adj(EcArgs a, const unsigned char *__restrict__ codes) {
  const int nx = a.ohi[0] - a.olo[0] + 1, ny = a.ohi[1] - a.olo[1] + 1,
            nz = a.ohi[2] - a.olo[2] + 1;
  const long long n = (long long)blockIdx.x * 256 + threadIdx.x;
  if (n >= (long long)nx * ny * nz)
    return;
  const int comp = blockIdx.y;
  const int q[3] = {a.olo[0] + (int)(n % nx), a.olo[1] + (int)((n / nx) % ny),
                    a.olo[2] + (int)(n / ((long long)nx * ny))};
  const unsigned char *cd = codes + comp * a.npts;
  const double *fp = a.fine + comp * a.fcs;
  const long long rnx = a.rhi[0] - a.rlo[0] + 1;
  const long long rny = a.rhi[1] - a.rlo[1] + 1;
  // a parent's code sits at each of its fine points in the region; read
  // the first
  auto code = [&](int px, int py, int pz) -> int {
    const int fx = max(2 * px, a.rlo[0]), fy = max(2 * py, a.rlo[1]),
              fz = max(2 * pz, a.rlo[2]);
    return cd[(fx - a.rlo[0]) + (fy - a.rlo[1]) * rnx +
              (fz - a.rlo[2]) * rnx * rny];
  };
  auto fb = [&](int x, int y, int z) {
    if (x < a.rlo[0] || x > a.rhi[0] || y < a.rlo[1] || y > a.rhi[1] ||
        z < a.rlo[2] || z > a.rhi[2])
      return 0.0;
    return fp[(x - a.f_lo[0]) + (y - a.f_lo[1]) * a.f_sy +
              (z - a.f_lo[2]) * a.f_sz];
  };
  const double s = ec_adj_one(q, a.plo, a.phi, code, fb);
  double *bp = const_cast<double *>(a.crse) + comp * a.ccs;
  bp[(q[0] - a.c_lo[0]) + (q[1] - a.c_lo[1]) * a.c_sy +
     (q[2] - a.c_lo[2]) * a.c_sz] += s;
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
  a.ntx = (a.phi[0] - a.plo[0] + EC_PX) / EC_PX;
  a.nty = (a.phi[1] - a.plo[1] + EC_PY) / EC_PY;
  return a;
}

} // namespace

bool eno_ccc_fits(const OpDesc &op, int per_group) {
  return !per_group && op.sid[0] == SID_ENO2 && op.sid[1] == SID_ENO2 &&
         op.sid[2] == SID_ENO2 && !op_fallback_per_var(op);
}

int eno_ccc_launch(const cxp_fab *crse, int crse_comp, const cxp_fab *fine,
                   int fine_comp, int ncomp, const int rlo[3],
                   const int rhi[3], unsigned char *codes, bool codes_only,
                   hipStream_t s) {
  const EcArgs a = ec_args(crse, crse_comp, fine ? fine : crse,
                           fine ? fine_comp : 0, rlo, rhi);
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
                       const unsigned char *codes, hipStream_t s) {
  const EcArgs a =
      ec_args(crse_bar, crse_comp, fine_bar, fine_comp, rlo, rhi);
  long long nout = 1;
  for (int d = 0; d < 3; ++d)
    nout *= a.ohi[d] - a.olo[d] + 1;
  hipLaunchKernelGGL(k_eno_ccc_adj,
                     dim3((unsigned)((nout + 255) / 256), ncomp), dim3(256), 0,
                     s, a, codes);
  return hipGetLastError() == hipSuccess ? CXP_OK : CXP_ERR_HIP;
}

} // namespace cxp
