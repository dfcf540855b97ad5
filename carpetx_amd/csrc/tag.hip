// Refinement tags of box-in-box regions (BoxInBox/src/boxinbox.cxx:102-174)
// and ErrorEst's threshold test (CarpetX/src/driver.cxx:1025-1080), on the
// device.  Both are one pass over a box's cells: the tag kernel writes one
// double per cell and reads nothing, the threshold kernel reads one double
// and writes at most one byte per cell, so both are bound by HBM.
//
// The membership test keeps the reference's operation order (cell centre
// x0 + i dx, then (X - position) / (radius - dx), then the squared sum from
// zero or the NaN-propagating maximum from -inf); the library is built
// without FMA contraction, so the decisions equal an IEEE evaluation of
// those expressions bit for bit.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>

#include "../../include/cxp_tag.h"

namespace {

struct BibArgs {
  int n;  // regions that apply at this level
  int shape[CXP_BIB_MAX_REGIONS];
  double pos[CXP_BIB_MAX_REGIONS][3];
  double rmax[CXP_BIB_MAX_REGIONS][3];
  double x0[3], dx[3];
};

// The box [lo, lo + n) inside a FAB whose first point is at element offset
// `base` of the array: element of cell (ix, iy, iz) of the box.
struct BoxMap {
  long long base;
  long long sy, sz;
  int lo[3];
  int n[3];
};

// max1 of Arith (Arith/src/defs.hxx:224-227): NaN in either argument wins.
__device__ __forceinline__ double max1(double x, double y) {
  return x != x ? x : (y != y ? y : (x < y ? y : x));
}

// One thread per two (x, y) columns of cells, 64 apart in x (so each of
// the wave's stores covers 64 consecutive cells), marching over a chunk of
// zc planes: 4 rows of y per block.  Each region's x and y quotients, and
// the first two terms of its sum or maximum, are computed once per column,
// and its z quotient once per plane for both columns; the order of the
// operations is the reference's, so the hoisting changes no result.
__global__ void __launch_bounds__(256)
    k_bib_tag(BibArgs a, double *__restrict__ err, BoxMap m, int zc) {
  const int ix = blockIdx.x * 128 + (threadIdx.x & 63);
  const int iy = blockIdx.y * 4 + (threadIdx.x >> 6);
  const int z0 = blockIdx.z * zc;
  const int z1 = min(z0 + zc, m.n[2]);
  if (ix >= m.n[0] || iy >= m.n[1])
    return;
  const bool two = ix + 64 < m.n[0];
  const double Y = a.x0[1] + (double)(m.lo[1] + iy) * a.dx[1];
  double part[2][CXP_BIB_MAX_REGIONS];  // sphere: (0 + qx qx) + qy qy; cube: max1
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const double X = a.x0[0] + (double)(m.lo[0] + ix + 64 * c) * a.dx[0];
#pragma unroll
    for (int r = 0; r < CXP_BIB_MAX_REGIONS; ++r) {
      if (r < a.n) {
        const double qx = (X - a.pos[r][0]) / a.rmax[r][0];
        const double qy = (Y - a.pos[r][1]) / a.rmax[r][1];
        part[c][r] = a.shape[r] == CXP_BIB_SPHERE
                         ? (0.0 + qx * qx) + qy * qy
                         : max1(max1(-INFINITY, fabs(qx)), fabs(qy));
      }
    }
  }
  double *p = err + m.base + ix + iy * m.sy + z0 * m.sz;
  for (int iz = z0; iz < z1; ++iz, p += m.sz) {
    const double Z = a.x0[2] + (double)(m.lo[2] + iz) * a.dx[2];
    bool refine0 = false, refine1 = false;
#pragma unroll
    for (int r = 0; r < CXP_BIB_MAX_REGIONS; ++r) {
      if (r < a.n) {
        const double qz = (Z - a.pos[r][2]) / a.rmax[r][2];
        if (a.shape[r] == CXP_BIB_SPHERE) {
          const double q2 = qz * qz;
          refine0 |= part[0][r] + q2 <= 1.0;
          refine1 |= part[1][r] + q2 <= 1.0;
        } else {
          const double aq = fabs(qz);
          refine0 |= max1(part[0][r], aq) <= 1.0;
          refine1 |= max1(part[1][r], aq) <= 1.0;
        }
      }
    }
    p[0] = refine0 ? 1.0 : 0.0;
    if (two)
      p[64] = refine1 ? 1.0 : 0.0;
  }
}

__global__ void k_box_init(int *box) {
  const int t = threadIdx.x;
  if (t < 7)
    box[t] = t == 0 ? 0 : (t < 4 ? INT_MAX : INT_MIN);
}

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
    v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
    v = min(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
    v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// Threshold test.  Block (by, bz) of 16 waves takes rows 16 by .. 16 by +
// 15 of planes [bz zc, bz zc + zc) and walks them 64 cells at a time.  A
// block contributes one set of atomics (count, bounding box) to the 7
// result words; device-scope atomics on one line from every XCD
// serialise, so the launch is sized to about one block per compute unit.
#define CXP_EST_WAVES 16
__global__ void __launch_bounds__(64 * CXP_EST_WAVES)
    k_error_est(const double *__restrict__ err, BoxMap m, double threshold,
                signed char *__restrict__ tags, int *box, int zc) {
  const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
  const int iy = blockIdx.x * CXP_EST_WAVES + ly;
  const int z0 = blockIdx.y * zc;
  const int z1 = min(z0 + zc, m.n[2]);
  int cnt = 0;
  int bl[3] = {INT_MAX, INT_MAX, INT_MAX};
  int bh[3] = {INT_MIN, INT_MIN, INT_MIN};
  if (iy < m.n[1]) {
    for (int iz = z0; iz < z1; ++iz) {
      const long long row = m.base + iy * m.sy + iz * m.sz;
      // four loads in flight per lane before the first compare
      for (int x0 = 0; x0 < m.n[0]; x0 += 256) {
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int ix = x0 + 64 * u + lx;
          v[u] = ix < m.n[0] ? err[row + ix] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int ix = x0 + 64 * u + lx;
          if (ix < m.n[0]) {
            const bool set = v[u] >= threshold;
            if (tags)
              tags[row + ix] = set ? CXP_TAG_SET : CXP_TAG_CLEAR;
            if (set) {
              ++cnt;
              bl[0] = min(bl[0], ix);
              bh[0] = max(bh[0], ix);
              bl[1] = min(bl[1], iy);
              bh[1] = max(bh[1], iy);
              bl[2] = min(bl[2], iz);
              bh[2] = max(bh[2], iz);
            }
          }
        }
      }
    }
  }
  if (!box)
    return;  // uniform: box is a kernel argument
  __shared__ int part[CXP_EST_WAVES][7];
  int v[7] = {wave_sum(cnt)};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    v[1 + d] = wave_min(bl[d]);
    v[4 + d] = wave_max(bh[d]);
  }
  if (lx == 0)
#pragma unroll
    for (int k = 0; k < 7; ++k)
      part[ly][k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    int c = 0, lo3[3] = {INT_MAX, INT_MAX, INT_MAX},
        hi3[3] = {INT_MIN, INT_MIN, INT_MIN};
    for (int w = 0; w < CXP_EST_WAVES; ++w) {
      c += part[w][0];
      for (int d = 0; d < 3; ++d) {
        lo3[d] = min(lo3[d], part[w][1 + d]);
        hi3[d] = max(hi3[d], part[w][4 + d]);
      }
    }
    if (c > 0) {  // box-relative indices -> global ones
      atomicAdd(box, c);
      for (int d = 0; d < 3; ++d) {
        atomicMin(box + 1 + d, m.lo[d] + lo3[d]);
        atomicMax(box + 4 + d, m.lo[d] + hi3[d]);
      }
    }
  }
}

// Planes per block so that a launch has about `target` blocks.
int z_chunk(long long blocks_per_plane, int nz, long long target) {
  long long zc = (blocks_per_plane * nz + target - 1) / target;
  if (zc < 1)
    zc = 1;
  if (zc > nz)
    zc = nz;
  return (int)zc;
}

// [lo, hi] nonempty and inside [flo, fhi]: the map of the box into the FAB.
// The launches put y rows (4 per block) on grid.y, so the box's y extent is
// at most 4 x 65535 rows.
int box_map(const int flo[3], const int fhi[3], const int lo[3],
            const int hi[3], BoxMap &m) {
  for (int d = 0; d < 3; ++d)
    if (lo[d] > hi[d] || lo[d] < flo[d] || hi[d] > fhi[d] || flo[d] > fhi[d])
      return CXP_ERR_REGION;
  if ((long long)hi[1] - lo[1] + 1 > 4LL * 65535)
    return CXP_ERR_ARG;
  const long long nx = (long long)fhi[0] - flo[0] + 1;
  const long long ny = (long long)fhi[1] - flo[1] + 1;
  m.sy = nx;
  m.sz = nx * ny;
  m.base = (lo[0] - flo[0]) + (lo[1] - flo[1]) * m.sy + (lo[2] - flo[2]) * m.sz;
  for (int d = 0; d < 3; ++d) {
    m.lo[d] = lo[d];
    m.n[d] = hi[d] - lo[d] + 1;
  }
  return CXP_OK;
}

} // namespace

extern "C" int cxp_boxinbox_tag(const cxp_bib_region *regions, int nregions,
                                int level, const double x0[3],
                                const double dx[3], double *err,
                                const int err_lo[3], const int err_hi[3],
                                const int lo[3], const int hi[3],
                                void *stream) {
  if (nregions < 0 || nregions > CXP_BIB_MAX_REGIONS ||
      (nregions > 0 && !regions) || !x0 || !dx || !err || !err_lo ||
      !err_hi || !lo || !hi)
    return CXP_ERR_ARG;
  BibArgs a = {};
  for (int r = 0; r < nregions; ++r) {
    const cxp_bib_region &g = regions[r];
    if (g.shape != CXP_BIB_SPHERE && g.shape != CXP_BIB_CUBE)
      return CXP_ERR_ARG;
    if (!g.active || !(level < g.num_levels))
      continue;
    // boxinbox.cxx:143-151: every radius of an applying region positive
    for (int d = 0; d < 3; ++d)
      if (g.radius[d] <= 0.0)  // a NaN radius passes, as there
        return CXP_ERR_ARG;
    a.shape[a.n] = g.shape;
    for (int d = 0; d < 3; ++d) {
      a.pos[a.n][d] = g.position[d];
      a.rmax[a.n][d] = g.radius[d] - dx[d];  // boxinbox.cxx:166
    }
    ++a.n;
  }
  for (int d = 0; d < 3; ++d) {
    a.x0[d] = x0[d];
    a.dx[d] = dx[d];
  }
  BoxMap m;
  const int e = box_map(err_lo, err_hi, lo, hi, m);
  if (e != CXP_OK)
    return e;
  const int nbx = (m.n[0] + 127) / 128, nby = (m.n[1] + 3) / 4;
  const int zc = z_chunk((long long)nbx * nby, m.n[2], 8192);
  const dim3 grid(nbx, nby, (m.n[2] + zc - 1) / zc);
  hipLaunchKernelGGL(k_bib_tag, grid, dim3(256), 0, (hipStream_t)stream, a,
                     err, m, zc);
  return hipGetLastError() == hipSuccess ? CXP_OK : CXP_ERR_HIP;
}

extern "C" int cxp_error_est(const double *err, const int err_lo[3],
                             const int err_hi[3], const int lo[3],
                             const int hi[3], double threshold,
                             signed char *tags, int *box_out, void *stream) {
  if (!err || !err_lo || !err_hi || !lo || !hi || (!tags && !box_out))
    return CXP_ERR_ARG;
  BoxMap m;
  const int e = box_map(err_lo, err_hi, lo, hi, m);
  if (e != CXP_OK)
    return e;
  const hipStream_t s = (hipStream_t)stream;
  if (box_out)
    hipLaunchKernelGGL(k_box_init, dim3(1), dim3(64), 0, s, box_out);
  const int nby = (m.n[1] + CXP_EST_WAVES - 1) / CXP_EST_WAVES;
  const int zc = z_chunk(nby, m.n[2], 256);
  hipLaunchKernelGGL(k_error_est, dim3(nby, (m.n[2] + zc - 1) / zc),
                     dim3(64 * CXP_EST_WAVES), 0, s, err, m, threshold, tags,
                     box_out, zc);
  return hipGetLastError() == hipSuccess ? CXP_OK : CXP_ERR_HIP;
}
