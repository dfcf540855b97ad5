// Adjoint (transpose) of the rf=2 prolongation: crse_bar += P^T fine_bar.
//
// The reference has no adjoints; these kernels are written from the forward
// operators.  All of them gather over coarse points, so no atomics are
// needed:
//  * linear operators (POLY, HERMITE, CONS): P = Pz (x) Py (x) Px, so
//    P^T = Px^T (x) Py^T (x) Pz^T and the three transposes commute.  A
//    workgroup owns a coarse (x, y) tile, marches through fine z-planes,
//    applies x^T and y^T inside each plane (LDS), and accumulates z^T into a
//    9-slot register window of coarse planes that is flushed and rotated
//    once per coarse plane.
//  * MINMOD: the Jacobian Jz Jy Jx has data-dependent weights, so the
//    transposes are applied in order z^T, y^T, x^T by three kernels through
//    two scratch arrays; each pass recomputes the limiter branch of its
//    forward application from the forward input.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/cxp.h"
#include "kernels.hpp"
#include "ops.hpp"

namespace cxp {

// --------------------------------------------------------------------------
// Weight tables of the linear operators: w[sid][off][q] is the weight of
// coarse offset q + wlo(sid) in the forward formula for fine offset off.

struct DeltaAcc {
  int o;
  __host__ __device__ constexpr double operator()(int k) const {
    return k == o ? 1.0 : 0.0;
  }
};

template <int S> constexpr double lin_weight(int off, int o) {
  return Op1<S>::apply(DeltaAcc{o}, off);
}

constexpr double lin_weight_rt(int sid, int off, int o) {
  switch (sid) {
#define CXP_LW(S)                                                              \
  case S:                                                                      \
    return lin_weight<S>(off, o);
    CXP_LW(SID_VPOLY1) CXP_LW(SID_VPOLY3) CXP_LW(SID_VPOLY5) CXP_LW(SID_VPOLY7)
    CXP_LW(SID_CPOLY0) CXP_LW(SID_CPOLY1) CXP_LW(SID_CPOLY2) CXP_LW(SID_CPOLY3)
    CXP_LW(SID_CPOLY4) CXP_LW(SID_CPOLY5) CXP_LW(SID_CPOLY6) CXP_LW(SID_CPOLY7)
    CXP_LW(SID_VCONS0) CXP_LW(SID_VCONS1) CXP_LW(SID_CCONS0) CXP_LW(SID_CCONS1)
    CXP_LW(SID_CCONS2) CXP_LW(SID_CCONS3) CXP_LW(SID_CCONS4) CXP_LW(SID_CCONS5)
    CXP_LW(SID_CCONS6) CXP_LW(SID_CCONS7)
#undef CXP_LW
  }
  return 0.0;
}

struct AdjW {
  double w[SID_COUNT][2][9];
};

constexpr AdjW make_adjw() {
  AdjW t{};
  for (int s = 0; s < SID_COUNT; ++s)
    for (int off = 0; off < 2; ++off)
      for (int q = 0; q < 9; ++q) {
        const SidInfo si = sid_info(s);
        const int o = q + sid_wlo(s);
        const bool in = o >= si.lo[off] && o <= si.hi[off];
        t.w[s][off][q] = in ? lin_weight_rt(s, off, o) : 0.0;
      }
  return t;
}

__constant__ AdjW kAdjW = make_adjw();

// --------------------------------------------------------------------------
// Linear adjoint

struct AdjTask {
  DFab fbar;  // fine adjoint (input)
  DFab cbar;  // coarse adjoint (accumulated)
  int rlo[3], rhi[3];
  int sid[3];
  int wlo[3], whi[3];
  int olo[3], ohi[3];  // coarse output range
  long long fcs, ccs;
};

#define CXP_ADJ_LX 5   // coarse tile x = 32
#define CXP_ADJ_LY 4   // coarse tile y = 16
#define CXP_ADJ_FXS 84 // fine window row capacity: 2*32 + 2*9 + 2
#define CXP_ADJ_FYS 52 // fine window rows: 2*16 + 2*9 + 2

__global__ void __launch_bounds__(256)
    k_adj_linear(AdjTask t, int ntx, int nty, int kchunk) {
  constexpr int CTX = 1 << CXP_ADJ_LX;
  constexpr int CTY = 1 << CXP_ADJ_LY;
  constexpr int NP = CTX * CTY / 256;
  __shared__ double Fs[CXP_ADJ_FYS * CXP_ADJ_FXS];
  __shared__ double Qs[CXP_ADJ_FYS * CTX];
  const int tid = threadIdx.x;
  const int b = blockIdx.x;
  const int tx = b % ntx, ty = (b / ntx) % nty, tz = b / (ntx * nty);
  const long long comp = blockIdx.y;
  const int ci0 = t.olo[0] + tx * CTX;
  const int cj0 = t.olo[1] + ty * CTY;
  const int kca = t.olo[2] + tz * kchunk;
  const int kcb = min(t.ohi[2], kca + kchunk - 1);
  const int CTXa = min(CTX, t.ohi[0] - ci0 + 1);
  const int CTYa = min(CTY, t.ohi[1] - cj0 + 1);
  const int fxa = 2 * (ci0 - t.whi[0]);
  const int fxw = 2 * (ci0 + CTXa - 1 - t.wlo[0]) + 1 - fxa + 1;
  const int fya = 2 * (cj0 - t.whi[1]);
  const int fyw = 2 * (cj0 + CTYa - 1 - t.wlo[1]) + 1 - fya + 1;
  const int nqx = t.whi[0] - t.wlo[0] + 1;
  const int nqy = t.whi[1] - t.wlo[1] + 1;
  DFab fb = t.fbar;
  fb.p += comp * t.fcs;
  DFab cb = t.cbar;
  cb.p += comp * t.ccs;
  const double(*wx)[9] = kAdjW.w[t.sid[0]];
  const double(*wy)[9] = kAdjW.w[t.sid[1]];
  const double(*wz)[9] = kAdjW.w[t.sid[2]];

  double acc[NP][9];
#pragma unroll
  for (int n = 0; n < NP; ++n)
#pragma unroll
    for (int q = 0; q < 9; ++q)
      acc[n][q] = 0.0;

  for (int m = kca - t.whi[2]; m <= kcb - t.wlo[2]; ++m) {
#pragma unroll
    for (int off = 0; off < 2; ++off) {
      const int fz = 2 * m + off;
      if (fz < t.rlo[2] || fz > t.rhi[2])
        continue;
      // fine plane window -> LDS, zero outside the region
      for (int idx = tid; idx < fxw * fyw; idx += 256) {
        const int fy = idx / fxw;
        const int fx = idx - fy * fxw;
        const int gx = fxa + fx, gy = fya + fy;
        double v = 0.0;
        if (gx >= t.rlo[0] && gx <= t.rhi[0] && gy >= t.rlo[1] && gy <= t.rhi[1])
          v = gld(fb.p, didx(fb, gx, gy, fz));
        Fs[fy * CXP_ADJ_FXS + fx] = v;
      }
      __syncthreads();
      // x^T: fine x -> coarse x, every window row
      for (int idx = tid; idx < (fyw << CXP_ADJ_LX); idx += 256) {
        const int fy = idx >> CXP_ADJ_LX;
        const int cil = idx & (CTX - 1);
        if (cil < CTXa) {
          const int ci = ci0 + cil;
          const double *row = Fs + fy * CXP_ADJ_FXS;
          double s = 0.0;
          for (int q = 0; q < nqx; ++q) {
            const int base = 2 * (ci - (q + t.wlo[0])) - fxa;
            s += wx[0][q] * row[base] + wx[1][q] * row[base + 1];
          }
          Qs[idx] = s;
        }
      }
      __syncthreads();
      // y^T, then z^T into the coarse-plane window
#pragma unroll
      for (int n = 0; n < NP; ++n) {
        const int p = tid + n * 256;
        const int cil = p & (CTX - 1);
        const int cjl = p >> CXP_ADJ_LX;
        if (cil < CTXa && cjl < CTYa) {
          const int cj = cj0 + cjl;
          double s = 0.0;
          for (int q = 0; q < nqy; ++q) {
            const int base = 2 * (cj - (q + t.wlo[1])) - fya;
            s += wy[0][q] * Qs[(base << CXP_ADJ_LX) + cil] +
                 wy[1][q] * Qs[((base + 1) << CXP_ADJ_LX) + cil];
          }
#pragma unroll
          for (int q = 0; q < 9; ++q)
            acc[n][q] += wz[off][q] * s;
        }
      }
    }
    // coarse plane m + wlo is complete
    const int kc = m + t.wlo[2];
    if (kc >= kca && kc <= kcb) {
#pragma unroll
      for (int n = 0; n < NP; ++n) {
        const int p = tid + n * 256;
        const int cil = p & (CTX - 1);
        const int cjl = p >> CXP_ADJ_LX;
        if (cil < CTXa && cjl < CTYa) {
          const long long di = didx(cb, ci0 + cil, cj0 + cjl, kc);
          gst(cb.p, di, gld(cb.p, di) + acc[n][0]);
        }
      }
    }
#pragma unroll
    for (int n = 0; n < NP; ++n) {
#pragma unroll
      for (int q = 0; q < 8; ++q)
        acc[n][q] = acc[n][q + 1];
      acc[n][8] = 0.0;
    }
  }
}

// --------------------------------------------------------------------------
// MINMOD adjoint (three passes)

// Linearised minmod (impl.hxx:857-877): y = c + s * k * d with k = 1/4 when
// the two one-sided differences share a sign, d the smaller of them; returns
// the weights of (c(-1), c(0), c(+1)).
__device__ __forceinline__ void minmod_lin(double cm, double c0, double cp,
                                           int off, double w[3]) {
  const double dm = c0 - cm;
  const double dp = cp - c0;
  const double k = copysign(0.125, dm) + copysign(0.125, dp);  // +-1/4 or 0
  const double s = off == 0 ? -1.0 : 1.0;
  w[0] = 0.0;
  w[1] = 1.0;
  w[2] = 0.0;
  if (k != 0.0) {
    if (fabs(dm) <= fabs(dp)) {  // delta = k * dm
      w[0] = -s * k;
      w[1] += s * k;
    } else {  // delta = k * dp
      w[1] -= s * k;
      w[2] = s * k;
    }
  }
}

// 1-D operator value (linear or minmod) at fine offset off from 3+ values.
struct MmTask {
  DFab crse;   // forward input
  DFab fbar;   // fine adjoint
  DFab cbar;   // coarse adjoint (accumulated)
  double *ybar;  // scratch: fine x, fine y, coarse z
  double *xbar;  // scratch: fine x, coarse y, coarse z
  int rlo[3], rhi[3];
  int sid[3];
  int olo[3], ohi[3];  // coarse output range (window of the region)
  long long ccs, fcs;
};

// value of the x-pass at fine x f0, coarse (jc, kc), per sid of x
__device__ inline double mm_xval(const MmTask &t, const DFab &c, int f0, int jc,
                                 int kc) {
  const int ic = f0 >> 1, off = f0 & 1;
  const double *row = c.p + didx(c, ic, jc, kc);
  return apply1d(t.sid[0], [&](int o) { return row[o]; }, 0, off);
}
__device__ inline double mm_yval(const MmTask &t, const DFab &c, int f0, int f1,
                                 int kc) {
  const int jc = f1 >> 1, off = f1 & 1;
  return apply1d(t.sid[1], [&](int o) { return mm_xval(t, c, f0, jc + o, kc); },
                 0, off);
}

// Weight of coarse offset o in the forward 1-D application with sid at fine
// offset off; for MINMOD the branch comes from the three input values.
__device__ inline double mm_weight(int sid, int off, int o, double vm, double v0,
                                   double vp) {
  if (sid == SID_MINMOD) {
    double w[3];
    minmod_lin(vm, v0, vp, off, w);
    return (o >= -1 && o <= 1) ? w[o + 1] : 0.0;
  }
  const int q = o - sid_wlo(sid);
  return (q >= 0 && q < 9) ? kAdjW.w[sid][off][q] : 0.0;
}

// Pass 1: ybar(f0, f1, kc) = sum_fz dZ(fz)/dY(kc) * fbar(f0, f1, fz)
__global__ void __launch_bounds__(256) k_adj_mm_z(MmTask t, long long comp) {
  const int nx = t.rhi[0] - t.rlo[0] + 1, ny = t.rhi[1] - t.rlo[1] + 1;
  const int nk = t.ohi[2] - t.olo[2] + 1;
  const long long n = (long long)nx * ny * nk;
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n)
    return;
  const int f0 = t.rlo[0] + (int)(g % nx);
  const int f1 = t.rlo[1] + (int)((g / nx) % ny);
  const int kc = t.olo[2] + (int)(g / ((long long)nx * ny));
  DFab c = t.crse;
  c.p += comp * t.ccs;
  DFab fb = t.fbar;
  fb.p += comp * t.fcs;
  const int sz = t.sid[2];
  double s = 0.0;
  for (int o = sid_wlo(sz); o <= sid_whi(sz); ++o) {
    const int kz = kc - o;
    double vm = 0, v0 = 0, vp = 0;
    bool have = false;
    for (int off = 0; off < 2; ++off) {
      const int fz = 2 * kz + off;
      if (fz < t.rlo[2] || fz > t.rhi[2])
        continue;
      if (sz == SID_MINMOD && !have) {
        vm = mm_yval(t, c, f0, f1, kz - 1);
        v0 = mm_yval(t, c, f0, f1, kz);
        vp = mm_yval(t, c, f0, f1, kz + 1);
        have = true;
      }
      s += mm_weight(sz, off, o, vm, v0, vp) * fb.p[didx(fb, f0, f1, fz)];
    }
  }
  t.ybar[g] = s;
}

// Pass 2: xbar(f0, jc, kc) = sum_f1 dY(f1)/dX(jc) * ybar(f0, f1, kc)
__global__ void __launch_bounds__(256) k_adj_mm_y(MmTask t, long long comp) {
  const int nx = t.rhi[0] - t.rlo[0] + 1, ny = t.rhi[1] - t.rlo[1] + 1;
  const int nj = t.ohi[1] - t.olo[1] + 1, nk = t.ohi[2] - t.olo[2] + 1;
  const long long n = (long long)nx * nj * nk;
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n)
    return;
  const int f0 = t.rlo[0] + (int)(g % nx);
  const int jc = t.olo[1] + (int)((g / nx) % nj);
  const int kcl = (int)(g / ((long long)nx * nj));
  const int kc = t.olo[2] + kcl;
  DFab c = t.crse;
  c.p += comp * t.ccs;
  const int sy = t.sid[1];
  double s = 0.0;
  for (int o = sid_wlo(sy); o <= sid_whi(sy); ++o) {
    const int jy = jc - o;
    double vm = 0, v0 = 0, vp = 0;
    bool have = false;
    for (int off = 0; off < 2; ++off) {
      const int f1 = 2 * jy + off;
      if (f1 < t.rlo[1] || f1 > t.rhi[1])
        continue;
      if (sy == SID_MINMOD && !have) {
        vm = mm_xval(t, c, f0, jy - 1, kc);
        v0 = mm_xval(t, c, f0, jy, kc);
        vp = mm_xval(t, c, f0, jy + 1, kc);
        have = true;
      }
      const long long yi = (long long)(f0 - t.rlo[0]) +
                           (long long)nx * ((f1 - t.rlo[1]) + (long long)ny * kcl);
      s += mm_weight(sy, off, o, vm, v0, vp) * t.ybar[yi];
    }
  }
  t.xbar[g] = s;
}

// Pass 3: cbar(ic, jc, kc) += sum_f0 dX(f0)/dC(ic) * xbar(f0, jc, kc)
__global__ void __launch_bounds__(256) k_adj_mm_x(MmTask t, long long comp) {
  const int nx = t.rhi[0] - t.rlo[0] + 1;
  const int ni = t.ohi[0] - t.olo[0] + 1, nj = t.ohi[1] - t.olo[1] + 1,
            nk = t.ohi[2] - t.olo[2] + 1;
  const long long n = (long long)ni * nj * nk;
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n)
    return;
  const int ic = t.olo[0] + (int)(g % ni);
  const int jcl = (int)((g / ni) % nj);
  const int kcl = (int)(g / ((long long)ni * nj));
  const int jc = t.olo[1] + jcl, kc = t.olo[2] + kcl;
  DFab c = t.crse;
  c.p += comp * t.ccs;
  DFab cb = t.cbar;
  cb.p += comp * t.ccs;
  const int sx = t.sid[0];
  double s = 0.0;
  for (int o = sid_wlo(sx); o <= sid_whi(sx); ++o) {
    const int ix = ic - o;
    // only coarse x owning a fine point of the region is read: its
    // neighbours ix-1, ix+1 then lie inside CoarseBox(region)
    if (2 * ix + 1 < t.rlo[0] || 2 * ix > t.rhi[0])
      continue;
    double vm = 0, v0 = 0, vp = 0;
    if (sx == SID_MINMOD) {
      const double *row = c.p + didx(c, ix, jc, kc);
      vm = row[-1];
      v0 = row[0];
      vp = row[1];
    }
    for (int off = 0; off < 2; ++off) {
      const int f0 = 2 * ix + off;
      if (f0 < t.rlo[0] || f0 > t.rhi[0])
        continue;
      const long long xi = (long long)(f0 - t.rlo[0]) +
                           (long long)nx * (jcl + (long long)nj * kcl);
      s += mm_weight(sx, off, o, vm, v0, vp) * t.xbar[xi];
    }
  }
  cb.p[didx(cb, ic, jc, kc)] += s;
}

// --------------------------------------------------------------------------

inline int adj_scratch(size_t bytes, double **out) {
  static std::mutex mu;
  static double *buf = nullptr;
  static size_t cap = 0;
  std::lock_guard<std::mutex> lk(mu);
  if (bytes > cap) {
    if (buf)
      (void)hipFree(buf);
    buf = nullptr;
    cap = 0;
    if (hipMalloc(&buf, bytes) != hipSuccess)
      return CXP_ERR_HIP;
    cap = bytes;
  }
  *out = buf;
  return CXP_OK;
}

inline DFab adj_dfab(const cxp_fab *f, int comp) {
  DFab v;
  v.p = f->data + (int64_t)comp * f->comp_stride;
  for (int d = 0; d < 3; ++d)
    v.lo[d] = f->lo[d];
  v.sy = f->hi[0] - f->lo[0] + 1;
  v.sz = (long long)v.sy * (f->hi[1] - f->lo[1] + 1);
  return v;
}

inline int adjoint_launch(const OpDesc &op, const cxp_fab *crse_in,
                          const cxp_fab *crse_bar, int crse_comp,
                          const cxp_fab *fine_bar, int fine_comp, int ncomp,
                          const int rlo[3], const int rhi[3], int per_group,
                          hipStream_t s) {
  (void)per_group;
  if (op_has_eno(op) || op_fallback_per_var(op) || op_fallback_per_group(op))
    return CXP_ERR_UNSUPPORTED;
  bool minmod = false;
  for (int d = 0; d < 3; ++d)
    minmod = minmod || op.sid[d] == SID_MINMOD;
  int olo[3], ohi[3];
  for (int d = 0; d < 3; ++d) {
    olo[d] = (rlo[d] >> 1) + sid_wlo(op.sid[d]);
    ohi[d] = (rhi[d] >> 1) + sid_whi(op.sid[d]);
  }
  if (!minmod) {
    AdjTask t;
    t.fbar = adj_dfab(fine_bar, fine_comp);
    t.cbar = adj_dfab(crse_bar, crse_comp);
    for (int d = 0; d < 3; ++d) {
      t.rlo[d] = rlo[d];
      t.rhi[d] = rhi[d];
      t.sid[d] = op.sid[d];
      t.wlo[d] = sid_wlo(op.sid[d]);
      t.whi[d] = sid_whi(op.sid[d]);
      t.olo[d] = olo[d];
      t.ohi[d] = ohi[d];
    }
    t.fcs = fine_bar->comp_stride;
    t.ccs = crse_bar->comp_stride;
    const int ntx = (ohi[0] - olo[0] + 1 + 31) / 32;
    const int nty = (ohi[1] - olo[1] + 1 + 15) / 16;
    const int nk = ohi[2] - olo[2] + 1;
    int kchunk = nk;
    while (kchunk > 4 && (long long)ntx * nty * ((nk + kchunk - 1) / kchunk) *
                                 ncomp < 1024)
      kchunk = (kchunk + 1) / 2;
    const int nkc = (nk + kchunk - 1) / kchunk;
    hipLaunchKernelGGL(k_adj_linear, dim3(ntx * nty * nkc, ncomp), dim3(256), 0,
                       s, t, ntx, nty, kchunk);
    return hipGetLastError() == hipSuccess ? CXP_OK : CXP_ERR_HIP;
  }
  if (!crse_in)
    return CXP_ERR_ARG;
  // the forward input must cover the same coarse box
  int clo[3], chi[3];
  op_coarse_box(op, rlo, rhi, clo, chi);
  for (int d = 0; d < 3; ++d)
    if (clo[d] < crse_in->lo[d] || chi[d] > crse_in->hi[d])
      return CXP_ERR_COARSE;
  if (crse_in->comp_stride != crse_bar->comp_stride && ncomp > 1)
    return CXP_ERR_ARG;
  MmTask t;
  t.crse = adj_dfab(crse_in, crse_comp);
  t.fbar = adj_dfab(fine_bar, fine_comp);
  t.cbar = adj_dfab(crse_bar, crse_comp);
  for (int d = 0; d < 3; ++d) {
    t.rlo[d] = rlo[d];
    t.rhi[d] = rhi[d];
    t.sid[d] = op.sid[d];
    t.olo[d] = olo[d];
    t.ohi[d] = ohi[d];
  }
  t.ccs = crse_in->comp_stride;
  t.fcs = fine_bar->comp_stride;
  const long long nx = rhi[0] - rlo[0] + 1, ny = rhi[1] - rlo[1] + 1;
  const long long ni = ohi[0] - olo[0] + 1, nj = ohi[1] - olo[1] + 1,
                  nk = ohi[2] - olo[2] + 1;
  const long long ny_elems = nx * ny * nk, nx_elems = nx * nj * nk;
  double *scr;
  const int e = adj_scratch((size_t)(ny_elems + nx_elems) * sizeof(double), &scr);
  if (e)
    return e;
  t.ybar = scr;
  t.xbar = scr + ny_elems;
  for (int comp = 0; comp < ncomp; ++comp) {
    hipLaunchKernelGGL(k_adj_mm_z, dim3((unsigned)((ny_elems + 255) / 256)),
                       dim3(256), 0, s, t, (long long)comp);
    hipLaunchKernelGGL(k_adj_mm_y, dim3((unsigned)((nx_elems + 255) / 256)),
                       dim3(256), 0, s, t, (long long)comp);
    hipLaunchKernelGGL(k_adj_mm_x, dim3((unsigned)((ni * nj * nk + 255) / 256)),
                       dim3(256), 0, s, t, (long long)comp);
  }
  return hipGetLastError() == hipSuccess ? CXP_OK : CXP_ERR_HIP;
}

} // namespace cxp
