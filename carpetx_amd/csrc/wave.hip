// One RK4 stage of the two-level wave harness (u_t = rho, rho_t = Laplace u)
// on one box, fused: the right-hand side at every interior point and the
// stage's update.  This is test-harness code for the differentiable
// evolution, not part of the transfer path's interface.
//
// Arrays hold 2 components (u, rho), x fastest, component stride
// nx * ny * nz.  The interior is [g, n - g) per direction; with periodic,
// g = 0 and neighbour indices wrap.  The 4th-order Laplacian is evaluated
// in the order of the torch expression it replaces, so the results are
// bit-identical with it (no FMA contraction):
//   out = -90 u;  per direction z, y, x:
//   out = (out + 16 (u[-1] + u[+1])) - (u[-2] + u[+2]);   lap = c out.
// Stage s (k = (rho, lap)):
//   s = 1:    acc = k;              dst = base + a k
//   s = 2, 3: acc = acc + 2 k;      dst = base + a k
//   s = 4:                          dst = base + a (acc + k)
#include <hip/hip_runtime.h>

#include "../../include/cxp_wave.h"

namespace {

__global__ void __launch_bounds__(256)
    k_wave_stage(int stage, const double *__restrict__ y, double *dst,
                 const double *base, double *acc, int nx, int ny, int nz,
                 int g, int periodic, double c, double a) {
  const int ix = blockIdx.x * 64 + (threadIdx.x & 63) + g;
  const int iy = blockIdx.y * 4 + (threadIdx.x >> 6) + g;
  const int iz = blockIdx.z + g;
  if (ix >= nx - g || iy >= ny - g || iz >= nz - g)
    return;
  const long long sy = nx, sz = (long long)nx * ny, cs = sz * nz;
  const long long p = ix + iy * sy + iz * sz;
  auto at = [&](int dx, int dy, int dz) {
    int x = ix + dx, yy = iy + dy, z = iz + dz;
    if (periodic) {
      x = (x + nx) % nx;
      yy = (yy + ny) % ny;
      z = (z + nz) % nz;
    }
    return y[x + yy * sy + z * sz];
  };
  double out = -90.0 * y[p];
  out = (out + 16.0 * (at(0, 0, -1) + at(0, 0, 1))) - (at(0, 0, -2) + at(0, 0, 2));
  out = (out + 16.0 * (at(0, -1, 0) + at(0, 1, 0))) - (at(0, -2, 0) + at(0, 2, 0));
  out = (out + 16.0 * (at(-1, 0, 0) + at(1, 0, 0))) - (at(-2, 0, 0) + at(2, 0, 0));
  const double k[2] = {y[cs + p], c * out};
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    const long long q = m * cs + p;
    if (stage == 4) {
      dst[q] = base[q] + a * (acc[q] + k[m]);
    } else {
      acc[q] = stage == 1 ? k[m] : acc[q] + 2.0 * k[m];
      dst[q] = base[q] + a * k[m];
    }
  }
}

} // namespace

extern "C" int cxp_wave_rk4_stage(int stage, const double *y, double *dst,
                                  const double *base, double *acc, int nx,
                                  int ny, int nz, int ghosts, int periodic,
                                  double c, double a, void *stream) {
  if (stage < 1 || stage > 4 || !y || !dst || !base || !acc || ghosts < 0 ||
      (periodic && ghosts != 0) || nx - 2 * ghosts < 1 ||
      ny - 2 * ghosts < 1 || nz - 2 * ghosts < 1 ||
      (periodic && (nx < 3 || ny < 3 || nz < 3)) || y == dst)
    return CXP_ERR_ARG;
  const int ix = nx - 2 * ghosts, iy = ny - 2 * ghosts, iz = nz - 2 * ghosts;
  const dim3 grid((ix + 63) / 64, (iy + 3) / 4, iz);
  hipLaunchKernelGGL(k_wave_stage, grid, dim3(256), 0, (hipStream_t)stream,
                     stage, y, dst, base, acc, nx, ny, nz, ghosts, periodic, c,
                     a);
  return hipGetLastError() == hipSuccess ? CXP_OK : CXP_ERR_HIP;
}
