// MINMOD adjoint kernels and their launches (routines in
// minmod_adjoint.hpp; launches declared in minmod_launch.hpp):
//  * k_mm_record_col   branch codes from the forward input, by columns;
//  * k_mm_gather       gather from codes, one thread per coarse point (any
//                      centring of the minmod family);
//  * k_mm_gather_tiled gather from codes, CCC, tiled;
//  * k_mm_fused        CCC adjoint at crse_in in one launch, codes
//                      recomputed per workgroup.
// One component per grid row y.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <utility>

#include "minmod_adjoint.hpp"
#include "minmod_launch.hpp"
#include "xcd.hpp"

namespace cxp {

namespace {

constexpr int kMmRecordChunk = 12;  // codes per column piece

// z-chunks of the tiled CCC kernels: at least this many planes (a chunk
// recomputes 4 planes of X and XY and re-reads 2 parents of fine data).
constexpr int kMtMinDepth = 2;

struct MmData {
  DFabC crse;  // forward input (record only)
  DFabC fbar;
  DFabC cbar;
  unsigned char *codes;
  long long ccs_in, fcs, ccs_bar;  // component strides
};

// Arguments of the tiled CCC kernels: one component's FABs (x fastest),
// component strides, the code arrays' boxes (gather only), the region and
// the output window, the tiling.
struct MtArgs {
  const double *crse;
  const double *fbar;
  double *cbar;
  const unsigned char *codes;
  long long c_sy, c_sz, f_sy, f_sz, b_sy, b_sz;
  long long ccs, fcs, bcs, ncodes;
  long long boff[3];
  int c_lo[3], f_lo[3], b_lo[3];
  int blo[3][3], bn[3][3];
  int rlo[3], rhi[3], olo[3], ohi[3];
  int ntx, nty, kc;
};

__global__ void __launch_bounds__(256) k_mm_gather(MmGeom g, MmData t) {
  const int ni = g.ohi[0] - g.olo[0] + 1, nj = g.ohi[1] - g.olo[1] + 1,
            nk = g.ohi[2] - g.olo[2] + 1;
  const long long n = (long long)blockIdx.x * 256 + threadIdx.x;
  if (n >= (long long)ni * nj * nk)
    return;
  const int comp = blockIdx.y;
  const int i = g.olo[0] + (int)(n % ni);
  const int j = g.olo[1] + (int)((n / ni) % nj);
  const int k = g.olo[2] + (int)(n / ((long long)ni * nj));
  const unsigned char *cd = t.codes + comp * g.ncodes;
  const double *fp = t.fbar.p + comp * t.fcs;
  const DFabC fb = t.fbar;
  auto codes = [&](int w, int x, int y, int z) {
    return cd[mm_code_index(g, w, x, y, z)];
  };
  auto fv = [&](int x, int y, int z) { return fp[dcidx(fb, x, y, z)]; };
  const double s = mm_gather_one(g, codes, fv, i, j, k);
  double *cb = t.cbar.p + comp * t.ccs_bar;
  cb[dcidx(t.cbar, i, j, k)] += s;
}

// Column record: one thread per column piece (mm_record_column), pieces of
// zc codes; a Bz piece evaluates zc + 2 XY values instead of 3 zc.
__global__ void __launch_bounds__(256)
    k_mm_record_col(MmGeom g, MmData t, long long pieces, int zc) {
  const long long n = (long long)blockIdx.x * 256 + threadIdx.x;
  if (n >= pieces)
    return;
  const int comp = blockIdx.y;
  const double *cp = t.crse.p + comp * t.ccs_in;
  const DFabC c = t.crse;
  unsigned char *cd = t.codes + comp * g.ncodes;
  auto cv = [&](int i, int j, int k) { return cp[dcidx(c, i, j, k)]; };
  auto store = [&](int w, int x, int y, int z, unsigned char v) {
    cd[mm_code_index(g, w, x, y, z)] = v;
  };
  mm_record_column(g, cv, store, n, zc);
}

// The footprint positions of thread tid: fine offsets in a plane of fbar
// (0 for positions outside the region, whose loads are masked) and the
// in-region flags.
struct MtPos {
  int off[MT_NPOS];  // the launch checks that a plane has < 2^31 points
  bool in[MT_NPOS];
};

__device__ inline MtPos mt_positions(const MtArgs &a, const MtTile &t,
                                     int tid) {
  MtPos p;
  const int fx = 2 * t.i0 - 2 + (tid & (MT_FP - 1));
  const bool xin = fx >= a.rlo[0] && fx <= a.rhi[0];
#pragma unroll
  for (int m = 0; m < MT_NPOS; ++m) {
    const int fy = 2 * t.j0 - 2 + tid / MT_FP + (MT_T / MT_FP) * m;
    p.in[m] = xin && fy >= a.rlo[1] && fy <= a.rhi[1];
    p.off[m] = p.in[m] ? (int)((fx - a.f_lo[0]) + (fy - a.f_lo[1]) * a.f_sy)
                       : 0;
  }
  return p;
}

// x^T outputs of plane k: crse_bar(i0 + li, j0 + lj, k) += mt_xt(...).
template <class BL, class BC>
__device__ inline void mt_store_xt(const MtArgs &a, const MtTile &t,
                                   double *bp, int tid, int k,
                                   const BL &xb_ld, const BC &bxc) {
  if (tid >= MT_NX)
    return;
  const int li = tid % MT_O, lj = tid / MT_O;
  const int i = t.i0 + li, j = t.j0 + lj;
  const double v = mt_xt(li, lj, xb_ld, bxc);
  if (i <= a.ohi[0] && j <= a.ohi[1]) {
    double *q = bp + (i - a.b_lo[0]) + (j - a.b_lo[1]) * a.b_sy +
                (k - a.b_lo[2]) * a.b_sz;
    *q += v;
  }
}

// Gather from recorded codes: parents pz = k0 - 1 .. k1 + 1 (z^T with the
// Bz codes and fine loads), then y^T and x^T of plane pz - 1 through LDS.
__global__ void __launch_bounds__(MT_T) k_mm_gather_tiled(MtArgs a) {
  __shared__ double xy[MT_FP * MT_FP];
  __shared__ double xb[MT_NY];
  const int tid = threadIdx.x;
  const int comp = blockIdx.y;
  const MtTile t = mt_tile(a.olo, a.ohi, xcd_block(blockIdx.x, gridDim.x),
                           a.ntx, a.nty, a.kc);
  const double *fp = a.fbar + comp * a.fcs;
  double *bp = a.cbar + comp * a.bcs;
  const unsigned char *cd = a.codes + comp * a.ncodes;
  const MtPos pos = mt_positions(a, t, tid);
  const int fx = 2 * t.i0 - 2 + (tid & (MT_FP - 1));
  const bool xin = fx >= a.rlo[0] && fx <= a.rhi[0];
  const int plo[3] = {a.rlo[0] >> 1, a.rlo[1] >> 1, a.rlo[2] >> 1};
  const int phi[3] = {a.rhi[0] >> 1, a.rhi[1] >> 1, a.rhi[2] >> 1};
  // in-plane part of each position's Bz index (its plane has < 2^31 codes:
  // fewer than the fine plane's points)
  int bzo[MT_NPOS];
#pragma unroll
  for (int m = 0; m < MT_NPOS; ++m) {
    const int fy = 2 * t.j0 - 2 + tid / MT_FP + (MT_T / MT_FP) * m;
    bzo[m] = pos.in[m] ? (fy - a.blo[2][1]) * a.bn[2][0] + (fx - a.blo[2][0])
                       : 0;
  }
  const long long bz_plane = (long long)a.bn[2][1] * a.bn[2][0];
  auto xy_ld = [&](int q) { return xy[q]; };
  auto xb_ld = [&](int q) { return xb[q]; };
  MtWin win;
  mt_win_zero(win);
  for (int pz = t.k0 - 1; pz <= t.k1 + 1; ++pz) {
    const bool zin = pz >= plo[2] && pz <= phi[2];
    const int fz0 = 2 * pz;
    const bool in0 = zin && fz0 >= a.rlo[2], in1 = zin && fz0 + 1 <= a.rhi[2];
    const double *pl0 = fp + (fz0 - a.f_lo[2]) * a.f_sz;
    const double *pl1 = pl0 + a.f_sz;
    const int k = pz - 1;
    const bool out = k >= t.k0;
    const unsigned char *bzp = cd + a.boff[2] + (pz - a.blo[2][2]) * bz_plane;
#pragma unroll
    for (int m = 0; m < MT_NPOS; ++m) {
      double g0 = 0.0, g1 = 0.0;
      unsigned char code = MM_ZERO;
      if (zin && pos.in[m]) {
        code = bzp[bzo[m]];
        if (in0)
          g0 = pl0[pos.off[m]];
        if (in1)
          g1 = pl1[pos.off[m]];
      }
      const double done = mt_zt(win, m, code, g0, g1);
      if (out)
        xy[tid + MT_T * m] = done;
    }
    if (!out)
      continue;
    __syncthreads();
    auto byc = [&](int lx, int lp) -> unsigned char {
      const int py = t.j0 - 1 + lp;
      if (!xin || py < plo[1] || py > phi[1])
        return MM_ZERO;
      return cd[a.boff[1] +
                ((long long)(k - a.blo[1][2]) * a.bn[1][1] +
                 (py - a.blo[1][1])) *
                    a.bn[1][0] +
                (2 * t.i0 - 2 + lx - a.blo[1][0])];
    };
    mt_yt_all(tid, xy_ld, byc, [&](int e, double v) { xb[e] = v; });
    __syncthreads();
    auto bxc = [&](int lp, int lj) -> unsigned char {
      const int px = t.i0 - 1 + lp, j = t.j0 + lj;
      if (px < plo[0] || px > phi[0] || j > a.ohi[1])
        return MM_ZERO;
      return cd[a.boff[0] +
                ((long long)(k - a.blo[0][2]) * a.bn[0][1] +
                 (j - a.blo[0][1])) *
                    a.bn[0][0] +
                (px - a.blo[0][0])];
    };
    mt_store_xt(a, t, bp, tid, k, xb_ld, bxc);
  }
}

// Tiled record of a CCC operator's codes (minmod_adjoint.hpp): each tile
// stores the codes it owns.  g carries the code arrays' boxes.
__global__ void __launch_bounds__(MT_T) k_mm_record_tiled(MtArgs a, MmGeom g) {
  __shared__ double cpl[MT_CW * MT_CW];
  __shared__ double xpl[MT_FP * MT_CW];
  const int tid = threadIdx.x;
  const int comp = blockIdx.y;
  const MtTile t = mt_tile(a.olo, a.ohi, xcd_block(blockIdx.x, gridDim.x),
                           a.ntx, a.nty, a.kc);
  const double *cp = a.crse + comp * a.ccs;
  unsigned char *cd = const_cast<unsigned char *>(a.codes) + comp * a.ncodes;
  const int plo2 = a.rlo[2] >> 1, phi2 = a.rhi[2] >> 1;
  auto cv = [&](int x, int y, int z) {
    return cp[(x - a.c_lo[0]) + (y - a.c_lo[1]) * a.c_sy +
              (z - a.c_lo[2]) * a.c_sz];
  };
  auto cpl_ld = [&](int q) { return cpl[q]; };
  auto xpl_st = [&](int q, double v) { xpl[q] = v; };
  auto xpl_ld = [&](int q) { return xpl[q]; };
  MtXY xyv;
  mt_xy_zero(xyv);
  for (int s = t.k0 - 1; s <= t.k1 + 1; ++s) {
    const bool own = s >= t.k0 && s <= t.k1;
    double cl[MT_NCL];
#pragma unroll
    for (int n = 0; n < MT_NCL; ++n) {
      const int q = tid + MT_T * n;
      cl[n] = q < MT_CW * MT_CW ? mt_coarse(a.olo, a.ohi, t, q, s, cv) : 0.0;
    }
#pragma unroll
    for (int n = 0; n < MT_NCL; ++n) {
      const int q = tid + MT_T * n;
      if (q < MT_CW * MT_CW)
        cpl[q] = cl[n];
    }
    __syncthreads();
    mt_x(tid, cpl_ld, xpl_st, own, [&](int q, unsigned char c) {
      int px, j;
      if (mt_owns_bx(g, t, q, px, j))
        cd[mm_code_index(g, 0, px, j, s)] = c;
    });
    __syncthreads();
    if (own)
      mt_by(tid, xpl_ld, [&](int q, unsigned char c) {
        int fx, py;
        if (mt_owns_by(g, t, q, fx, py))
          cd[mm_code_index(g, 1, fx, py, s)] = c;
      });
    unsigned char bz[MT_NPOS];
    mt_xy(tid, xpl_ld, xyv, bz);
    const int pz = s - 1;
    if (pz >= t.k0 && pz <= t.k1 && pz >= plo2 && pz <= phi2)
#pragma unroll
      for (int m = 0; m < MT_NPOS; ++m) {
        int fx, fy;
        if (mt_owns_bz(g, t, tid, m, fx, fy))
          cd[mm_code_index(g, 2, fx, fy, pz)] = bz[m];
      }
  }
}

// Fused CCC adjoint at crse_in (the march of minmod_adjoint.hpp: phases A,
// B, C, two barriers per coarse plane).  Block-uniform control flow: every
// thread reaches every barrier.
__global__ void __launch_bounds__(MT_T)
k_mm_fused(MtArgs a) {
  __shared__ double cpl[MT_CW * MT_CW];
  __shared__ double xpl[MT_FP * MT_CW];
  __shared__ double xy[MT_FP * MT_FP];
  __shared__ double xb[MT_NY];
  __shared__ unsigned char bxr[MT_RING][MT_PB * MT_O];
  __shared__ unsigned char byr[MT_RING][MT_FP * MT_PB];
  const int tid = threadIdx.x;
  const int comp = blockIdx.y;
  const MtTile t = mt_tile(a.olo, a.ohi, xcd_block(blockIdx.x, gridDim.x),
                           a.ntx, a.nty, a.kc);
  const double *cp = a.crse + comp * a.ccs;
  const double *fp = a.fbar + comp * a.fcs;
  double *bp = a.cbar + comp * a.bcs;
  const MtPos pos = mt_positions(a, t, tid);
  const int plo2 = a.rlo[2] >> 1, phi2 = a.rhi[2] >> 1;
  auto cv = [&](int x, int y, int z) {
    return cp[(x - a.c_lo[0]) + (y - a.c_lo[1]) * a.c_sy +
              (z - a.c_lo[2]) * a.c_sz];
  };
  auto cpl_ld = [&](int q) { return cpl[q]; };
  auto xpl_st = [&](int q, double v) { xpl[q] = v; };
  auto xpl_ld = [&](int q) { return xpl[q]; };
  auto xy_ld = [&](int q) { return xy[q]; };
  auto xb_ld = [&](int q) { return xb[q]; };
  MtWin win;
  MtXY xyv;
  mt_win_zero(win);
  mt_xy_zero(xyv);
  // plane s's coarse values and the fine adjoints of parent s - 1, loaded
  // one step ahead so their latency overlaps the previous step
  double cl[MT_NCL], g0[MT_NPOS], g1[MT_NPOS];
  auto load_step = [&](int s) {
    if (s <= t.k1 + 2)
#pragma unroll
      for (int n = 0; n < MT_NCL; ++n) {
        const int q = tid + MT_T * n;
        cl[n] = q < MT_CW * MT_CW ? mt_coarse(a.olo, a.ohi, t, q, s, cv) : 0.0;
      }
    const int pz = s - 1;
    const bool zt =
        pz >= t.k0 - 1 && pz <= t.k1 + 1 && pz >= plo2 && pz <= phi2;
    const int fz0 = 2 * pz;
    const bool in0 = zt && fz0 >= a.rlo[2], in1 = zt && fz0 + 1 <= a.rhi[2];
    const double *pl0 = fp + (fz0 - a.f_lo[2]) * a.f_sz;
    const double *pl1 = pl0 + a.f_sz;
#pragma unroll
    for (int m = 0; m < MT_NPOS; ++m) {
      g0[m] = pos.in[m] && in0 ? pl0[pos.off[m]] : 0.0;
      g1[m] = pos.in[m] && in1 ? pl1[pos.off[m]] : 0.0;
    }
  };
  load_step(t.k0 - 2);
  for (int s = t.k0 - 2; s <= t.k1 + 3; ++s) {
    const bool produce = s <= t.k1 + 2;
    const bool own = s >= t.k0 && s <= t.k1;
    const int k = s - 3;  // plane whose y^T, x^T run in this step
    unsigned char *bxs = bxr[s & (MT_RING - 1)];
    unsigned char *bys = byr[s & (MT_RING - 1)];
    const unsigned char *bxk = bxr[k & (MT_RING - 1)];
    const unsigned char *byk = byr[k & (MT_RING - 1)];
    // A: this plane's coarse values to LDS; the fine adjoints of parent
    // s - 1 are used in C
    double gc0[MT_NPOS], gc1[MT_NPOS];
#pragma unroll
    for (int m = 0; m < MT_NPOS; ++m) {
      gc0[m] = g0[m];
      gc1[m] = g1[m];
    }
    if (produce)
#pragma unroll
      for (int n = 0; n < MT_NCL; ++n) {
        const int q = tid + MT_T * n;
        if (q < MT_CW * MT_CW)
          cpl[q] = cl[n];
      }
    if (s < t.k1 + 3)
      load_step(s + 1);
    __syncthreads();
    if (produce)
      mt_x(tid, cpl_ld, xpl_st, own,
           [&](int q, unsigned char c) { bxs[q] = c; });
    if (k >= t.k0) {
      auto byc = [&](int lx, int lp) { return byk[lp * MT_FP + lx]; };
      mt_yt_all(tid, xy_ld, byc, [&](int e, double v) { xb[e] = v; });
    }
    __syncthreads();
    unsigned char bz[MT_NPOS] = {};
    if (produce) {
      if (own)
        mt_by(tid, xpl_ld, [&](int q, unsigned char c) { bys[q] = c; });
      mt_xy(tid, xpl_ld, xyv, bz);
    }
    if (k >= t.k0)
      mt_store_xt(a, t, bp, tid, k, xb_ld,
                  [&](int lp, int lj) { return bxk[lj * MT_PB + lp]; });
    const bool store = s - 2 >= t.k0 && s - 2 <= t.k1;
#pragma unroll
    for (int m = 0; m < MT_NPOS; ++m) {
      const double done = mt_zt(win, m, bz[m], gc0[m], gc1[m]);
      if (store)
        xy[tid + MT_T * m] = done;
    }
  }
}

DFabC mm_dfab(const cxp_fab *f, int comp) {
  DFabC v;
  v.p = f->data + (int64_t)comp * f->comp_stride;
  for (int d = 0; d < 3; ++d)
    v.lo[d] = f->lo[d];
  v.sy = f->hi[0] - f->lo[0] + 1;
  v.sz = v.sy * (f->hi[1] - f->lo[1] + 1);
  return v;
}

// Workgroups of a tiled kernel resident at once on the current device
// (CUs x blocks per CU), cached per device and kernel.
template <class K> long long mt_slots(K kernel, int which) {
  static std::mutex mu;
  static std::map<std::pair<int, int>, long long> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess)
    return 1024;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find({dev, which});
  if (it != cache.end())
    return it->second;
  int cus = 0, per_cu = 0;
  long long n = 1024;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                            dev) == hipSuccess &&
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, MT_T, 0) ==
          hipSuccess &&
      cus > 0 && per_cu > 0)
    n = (long long)cus * per_cu;
  cache[{dev, which}] = n;
  return n;
}

// The tiled kernels' arguments and grid for a CCC operator.
MtArgs mt_args(const MmGeom &g, const cxp_fab *crse_in, int crse_comp,
               const cxp_fab *crse_bar, int bar_comp, const cxp_fab *fine_bar,
               int fine_comp, int ncomp, const unsigned char *codes,
               long long slots, dim3 &grid) {
  MtArgs a{};
  if (crse_in) {
    const DFabC c = mm_dfab(crse_in, crse_comp);
    a.crse = c.p;
    a.c_sy = c.sy;
    a.c_sz = c.sz;
    a.ccs = crse_in->comp_stride;
    for (int d = 0; d < 3; ++d)
      a.c_lo[d] = c.lo[d];
  }
  const DFabC f = mm_dfab(fine_bar, fine_comp), b = mm_dfab(crse_bar, bar_comp);
  a.fbar = f.p;
  a.f_sy = f.sy;
  a.f_sz = f.sz;
  a.fcs = fine_bar->comp_stride;
  a.cbar = b.p;
  a.b_sy = b.sy;
  a.b_sz = b.sz;
  a.bcs = crse_bar->comp_stride;
  a.codes = codes;
  a.ncodes = g.ncodes;
  for (int d = 0; d < 3; ++d) {
    a.f_lo[d] = f.lo[d];
    a.b_lo[d] = b.lo[d];
    a.rlo[d] = g.rlo[d];
    a.rhi[d] = g.rhi[d];
    a.olo[d] = g.olo[d];
    a.ohi[d] = g.ohi[d];
    a.boff[d] = g.boff[d];
    for (int e = 0; e < 3; ++e) {
      a.blo[d][e] = g.blo[d][e];
      a.bn[d][e] = g.bn[d][e];
    }
  }
  a.ntx = (g.ohi[0] - g.olo[0] + MT_O) / MT_O;
  a.nty = (g.ohi[1] - g.olo[1] + MT_O) / MT_O;
  // As many z-chunks as fill the device's resident workgroups once: longer
  // marches are cheaper per plane, a second partial round costs a whole
  // chunk's time.
  const long long nk = g.ohi[2] - g.olo[2] + 1;
  const long long cols = (long long)a.ntx * a.nty * ncomp;
  long long nkc = slots / cols;
  nkc = nkc < 1 ? 1 : (nkc > nk ? nk : nkc);
  long long kc = (nk + nkc - 1) / nkc;
  kc = kc < kMtMinDepth ? kMtMinDepth : kc;
  a.kc = (int)kc;
  grid = dim3((unsigned)(a.ntx * a.nty * ((nk + kc - 1) / kc)), ncomp);
  return a;
}

} // namespace

long long mm_code_bytes(const OpDesc &op, const int rlo[3], const int rhi[3],
                        int ncomp) {
  MmGeom g;
  mm_setup(g, op.sid, rlo, rhi);
  return g.ncodes * ncomp;
}

int mm_record_launch(const OpDesc &op, const cxp_fab *crse_in, int crse_comp,
                     int ncomp, const int rlo[3], const int rhi[3],
                     unsigned char *codes, hipStream_t s) {
  MmGeom g;
  mm_setup(g, op.sid, rlo, rhi);
  if (g.ncodes == 0)
    return CXP_OK;
  if (op_all_minmod(op)) {
    dim3 grid;
    // the record reads no fine data; the fine box is the region's
    cxp_fab region{};
    for (int d = 0; d < 3; ++d) {
      region.lo[d] = rlo[d];
      region.hi[d] = rhi[d];
    }
    region.comp_stride = 1;
    const MtArgs a =
        mt_args(g, crse_in, crse_comp, &region, 0, &region, 0, ncomp, codes,
                mt_slots(k_mm_record_tiled, 2), grid);
    hipLaunchKernelGGL(k_mm_record_tiled, grid, dim3(MT_T), 0, s, a, g);
    return hipGetLastError() == hipSuccess ? CXP_OK : CXP_ERR_HIP;
  }
  MmData t{};
  t.crse = mm_dfab(crse_in, crse_comp);
  t.ccs_in = crse_in->comp_stride;
  t.codes = codes;
  long long pieces = 0;
  for (int w = 0; w < 3; ++w)
    pieces += mm_columns(g, w, kMmRecordChunk);
  hipLaunchKernelGGL(k_mm_record_col,
                     dim3((unsigned)((pieces + 255) / 256), ncomp), dim3(256),
                     0, s, g, t, pieces, kMmRecordChunk);
  return hipGetLastError() == hipSuccess ? CXP_OK : CXP_ERR_HIP;
}

int mm_gather_launch(const OpDesc &op, const cxp_fab *crse_bar, int crse_comp,
                     const cxp_fab *fine_bar, int fine_comp, int ncomp,
                     const int rlo[3], const int rhi[3],
                     const unsigned char *codes, hipStream_t s) {
  MmGeom g;
  mm_setup(g, op.sid, rlo, rhi);
  if (op_all_minmod(op) && mm_tiled_fits(fine_bar)) {
    dim3 grid;
    const MtArgs a =
        mt_args(g, nullptr, 0, crse_bar, crse_comp, fine_bar, fine_comp, ncomp,
                codes, mt_slots(k_mm_gather_tiled, 0), grid);
    hipLaunchKernelGGL(k_mm_gather_tiled, grid, dim3(MT_T), 0, s, a);
    return hipGetLastError() == hipSuccess ? CXP_OK : CXP_ERR_HIP;
  }
  long long nout = 1;
  for (int d = 0; d < 3; ++d)
    nout *= g.ohi[d] - g.olo[d] + 1;
  MmData t{};
  t.fbar = mm_dfab(fine_bar, fine_comp);
  t.cbar = mm_dfab(crse_bar, crse_comp);
  t.fcs = fine_bar->comp_stride;
  t.ccs_bar = crse_bar->comp_stride;
  t.codes = const_cast<unsigned char *>(codes);
  hipLaunchKernelGGL(k_mm_gather, dim3((unsigned)((nout + 255) / 256), ncomp),
                     dim3(256), 0, s, g, t);
  return hipGetLastError() == hipSuccess ? CXP_OK : CXP_ERR_HIP;
}

int mm_fused_launch(const OpDesc &op, const cxp_fab *crse_in,
                    const cxp_fab *crse_bar, int crse_comp,
                    const cxp_fab *fine_bar, int fine_comp, int ncomp,
                    const int rlo[3], const int rhi[3], hipStream_t s) {
  if (!op_all_minmod(op) || !mm_tiled_fits(fine_bar))
    return CXP_ERR_UNSUPPORTED;
  MmGeom g;
  mm_setup(g, op.sid, rlo, rhi);
  dim3 grid;
  const MtArgs a =
      mt_args(g, crse_in, crse_comp, crse_bar, crse_comp, fine_bar, fine_comp,
              ncomp, nullptr, mt_slots(k_mm_fused, 1), grid);
  hipLaunchKernelGGL(k_mm_fused, grid, dim3(MT_T), 0, s, a);
  return hipGetLastError() == hipSuccess ? CXP_OK : CXP_ERR_HIP;
}

} // namespace cxp
