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

// Periodic index of i + o, |o| <= 2 < n (no integer division).
__device__ __forceinline__ int wrap(int i, int n) {
  return i < 0 ? i + n : (i >= n ? i - n : i);
}

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
      if (dx)
        x = wrap(x, nx);
      if (dy)
        yy = wrap(yy, ny);
      if (dz)
        z = wrap(z, nz);
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

// Adjoint of one stage, in gather form (no atomics).  Given the gradients
// dst_bar and, for stages 1-3, acc_out_bar, with k_bar = a dst_bar + w
// acc_out_bar (w = 1 for stage 1, 2 for stages 2-3; stage 4: k_bar =
// a dst_bar) at interior points:
//   y_bar[rho](p) += k_bar[u](p)               (interior p)
//   y_bar[u](q)   += c sum_p L(p, q) k_bar[rho](p)   (every q a stencil reaches)
// L is symmetric in its offsets: -90 at q = p, 16 at distance 1 and -1 at
// distance 2 along an axis.  One thread per point q of y (ghosts
// included; with periodic, the wrapped neighbours).  The pointwise parts
// (base_bar, acc_in_bar) are done by the caller.
//
// A wave covers 64 consecutive x of one (y, z) row, from two points before
// its 60 outputs to two after: each lane forms k_bar[rho] of its own point
// once, and the x-direction taps take it from the lanes beside it; the y
// and z taps read the gradients themselves.
#define WA_OUT 60  // outputs per wave (lanes 2 .. 61)

__global__ void __launch_bounds__(256)
    k_wave_stage_adjoint(int stage, const double *__restrict__ dst_bar,
                         const double *__restrict__ acc_bar,
                         double *__restrict__ y_bar, int nx, int ny, int nz,
                         int g, int periodic, double c, double a) {
  const int lane = threadIdx.x & 63;
  const int qx = blockIdx.x * WA_OUT - 2 + lane;
  const int qy = blockIdx.y * 4 + (threadIdx.x >> 6);
  const int qz = blockIdx.z;
  if (qy >= ny)
    return;  // whole waves: the shuffles below see only live lanes
  const long long sy = nx, sz = (long long)nx * ny, cs = sz * nz;
  const double w = stage == 1 ? 1.0 : 2.0;
  auto inside = [&](int x, int yy, int z) {
    return x >= g && x < nx - g && yy >= g && yy < ny - g && z >= g &&
           z < nz - g;
  };
  // k_bar of component m at interior point (x, yy, z)
  auto kbar = [&](int m, int x, int yy, int z) {
    const long long p = m * cs + x + yy * sy + z * sz;
    return stage == 4 ? a * dst_bar[p] : a * dst_bar[p] + w * acc_bar[p];
  };
  // this lane's point, wrapped (periodic) or as is; k_bar[rho] there, 0
  // where it is not interior or lies beyond the reach of any output
  int ox = qx;
  bool reach = qx >= -2 && qx <= nx + 1;
  if (periodic && reach)
    ox = wrap(qx, nx);
  const double kb_self =
      (reach && inside(ox, qy, qz)) ? kbar(1, ox, qy, qz) : 0.0;
  const int off[4] = {-2, -1, 1, 2};
  const double wt[4] = {-1.0, 16.0, 16.0, -1.0};
  double kx[4];
#pragma unroll
  for (int m = 0; m < 4; ++m)
    kx[m] = __shfl(kb_self, lane + off[m] < 0 ? 0 : (lane + off[m] > 63 ? 63 : lane + off[m]));
  if (lane < 2 || lane > 61 || qx >= nx)
    return;
  const long long q = qx + qy * sy + qz * sz;
  double su = 0.0;
  // the point itself
  if (inside(qx, qy, qz))
    su += -90.0 * kb_self;
#pragma unroll
  for (int dir = 0; dir < 3; ++dir)
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      int x = qx, yy = qy, z = qz;
      if (dir == 0)
        x += off[m];
      else if (dir == 1)
        yy += off[m];
      else
        z += off[m];
      if (periodic) {
        if (dir == 0)
          x = wrap(x, nx);
        else if (dir == 1)
          yy = wrap(yy, ny);
        else
          z = wrap(z, nz);
      }
      if (inside(x, yy, z))
        su += wt[m] * (dir == 0 ? kx[m] : kbar(1, x, yy, z));
    }
  y_bar[q] += c * su;
  if (inside(qx, qy, qz))
    y_bar[cs + q] += kbar(0, qx, qy, qz);
}

} // namespace

extern "C" int cxp_wave_rk4_stage_adjoint(int stage, const double *dst_bar,
                                          const double *acc_bar, double *y_bar,
                                          int nx, int ny, int nz, int ghosts,
                                          int periodic, double c, double a,
                                          void *stream) {
  if (stage < 1 || stage > 4 || !dst_bar || !y_bar ||
      (stage < 4 && !acc_bar) || ghosts < 0 || (periodic && ghosts != 0) ||
      (!periodic && ghosts < 2) ||
      nx - 2 * ghosts < 1 || ny - 2 * ghosts < 1 || nz - 2 * ghosts < 1 ||
      (periodic && (nx < 3 || ny < 3 || nz < 3)))
    return CXP_ERR_ARG;
  const dim3 grid((nx + WA_OUT - 1) / WA_OUT, (ny + 3) / 4, nz);
  hipLaunchKernelGGL(k_wave_stage_adjoint, grid, dim3(256), 0,
                     (hipStream_t)stream, stage, dst_bar, acc_bar, y_bar, nx,
                     ny, nz, ghosts, periodic, c, a);
  return hipGetLastError() == hipSuccess ? CXP_OK : CXP_ERR_HIP;
}

extern "C" int cxp_wave_rk4_stage(int stage, const double *y, double *dst,
                                  const double *base, double *acc, int nx,
                                  int ny, int nz, int ghosts, int periodic,
                                  double c, double a, void *stream) {
  if (stage < 1 || stage > 4 || !y || !dst || !base || !acc || ghosts < 0 ||
      (periodic && ghosts != 0) || (!periodic && ghosts < 2) ||
      nx - 2 * ghosts < 1 ||
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
