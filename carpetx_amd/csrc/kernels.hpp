// Device kernels of the rf=2 transfer path (gfx950, FP64).
//
//  * sep_tile: separable prolongation of one fine tile.  The workgroup owns
//    a TX x TY fine (x, y) tile and marches through coarse z-planes:
//      coarse plane tile + halo  -> LDS                  (one HBM read each)
//      x-pass (fine x, coarse y) -> LDS
//      y-pass (fine x, fine y)   -> LDS ring of the last R planes
//      z-pass from the ring      -> fine FAB, stores coalesced along x.
//    Each pass evaluates the reference 1-D operator with its own summation
//    form, in the order call_stencil_3d uses (x innermost), so results are
//    bit-identical to the reference's per-point tensor product when the
//    library is compiled without FMA contraction.
//  * point kernels: ENO shift search and linear fallback need the whole
//    3-D stencil per fine point (prolongate_3d_rf2_impl.hxx:1393-1671 and
//    :1836-2148), so those operators run one thread per fine point.
//  * restriction: injection in VC directions, mean of the two children in CC
//    directions (AMReX average_down family called by Restrict,
//    CarpetX/src/schedule.cxx:2791-2816).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "stencil1d.hpp"
#include "xcd.hpp"
#include "../../include/cxp.h"

namespace cxp {

// One component of a FAB, addressed in the level's global index space.
struct DFab {
  double *p;     // element (lo) of this component
  int lo[3];
  int sy;        // x-row stride
  long long sz;  // plane stride
};

// Global-address-space accessors.  Pointers that come out of task structs
// are generic; dereferenced as such the compiler emits flat instructions,
// which also count against lgkmcnt, so every LDS barrier would wait for all
// outstanding HBM traffic.  Casting to address space 1 gives global_load /
// global_store, tracked by vmcnt only.
typedef __attribute__((address_space(1))) double gdouble;
typedef __attribute__((address_space(1))) const double gcdouble;
__device__ __forceinline__ double gld(const double *p, long long i) {
  return ((gcdouble *)p)[i];
}
__device__ __forceinline__ void gst(double *p, long long i, double v) {
  ((gdouble *)p)[i] = v;
}

// Raw buffer access.  Out-of-range offsets are dropped by the hardware
// (stores) or read as 0 (loads), so masked lanes need no branch and every
// iteration of a loop issues the same number of memory instructions; the
// compiler can then wait for the prefetch loads alone (vmcnt(N)) instead of
// also draining the stores issued after them.
typedef unsigned int cxp_u32x2 __attribute__((ext_vector_type(2)));
#define CXP_OOB 0x7ffffff0
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const double *p,
                                                            int nbytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, nbytes,
                                           0x00020000);
}
__device__ __forceinline__ double bld(__amdgpu_buffer_rsrc_t r, int boff) {
  return __builtin_bit_cast(double,
                            __builtin_amdgcn_raw_buffer_load_b64(r, boff, 0, 0));
}
// Load at voff + soff with a uniform soff.  Callers keep voff = CXP_OOB for
// lanes that must read 0 and voff + soff inside the buffer for the others,
// so the result does not depend on whether the range check includes soff.
__device__ __forceinline__ double bld_s(__amdgpu_buffer_rsrc_t r, int voff,
                                        int soff) {
  return __builtin_bit_cast(
      double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ __forceinline__ void bst(__amdgpu_buffer_rsrc_t r, int boff,
                                    double v) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(cxp_u32x2, v), r,
                                        boff, 0, 0);
}

__host__ __device__ __forceinline__ long long didx(const DFab &f, int i, int j,
                                                  int k) {
  return (long long)(i - f.lo[0]) + (long long)(j - f.lo[1]) * f.sy +
         (long long)(k - f.lo[2]) * f.sz;
}

// A separable prolongation task: one (patch, component).
struct SepTask {
  DFab crse;
  DFab fine;
  int rlo[3], rhi[3];
  int sid[3];
  int wlo[3], whi[3];  // union stencil window per direction
  long long ccs, fcs;  // component strides (per-box launches)
  int cbytes, fbytes;  // bytes of one component of each FAB (< 2 GiB)
};

// A work tile of the batched list.
struct SepTile {
  int task;
  int comp;  // component offset applied with ccs/fcs
  int x0, y0;
  int z0, z1;
  int cfg;  // tile shape: log2 of the fine tile width (+ 16: two row pairs
            // per thread); -1: thin tile
  int pad;
};

#define CXP_LINEAR_SIDS(X)                                                     \
  X(SID_VPOLY1) X(SID_VPOLY3) X(SID_VPOLY5) X(SID_VPOLY7) X(SID_CPOLY0)        \
  X(SID_CPOLY1) X(SID_CPOLY2) X(SID_CPOLY3) X(SID_CPOLY4) X(SID_CPOLY5)        \
  X(SID_CPOLY6) X(SID_CPOLY7) X(SID_VCONS0) X(SID_VCONS1) X(SID_CCONS0)        \
  X(SID_CCONS1) X(SID_CCONS2) X(SID_CCONS3) X(SID_CCONS4) X(SID_CCONS5)        \
  X(SID_CCONS6) X(SID_CCONS7) X(SID_MINMOD)

// Runtime tile geometry.  A workgroup of 128 threads owns a TX x TY fine
// (x, y) tile with TX * TY = 256 RP: thread t owns fine x = x0 + (t % TX)
// and the 2 RP fine rows y0 + 2 RP (t / TX) + {0 .. 2 RP - 1} (RP pairs of
// rows, each pair sharing one coarse y).  Tiles start at even x and y.
// RP = 2 gives each thread twice the points per coarse plane for the same
// barriers and loop overhead.
struct Geom {
  int lx;        // log2 TX
  int tx, ty;    // fine tile extents
  int cxs, cys;  // LDS capacity of the coarse tile (row stride, rows)
  int xs_off;    // start of the x-pass buffer (16-byte aligned)
};

__host__ __device__ inline Geom make_geom(int lx, int rp = 1) {
  Geom g;
  g.lx = lx;
  g.tx = 1 << lx;
  g.ty = (256 * rp) >> lx;
  g.cxs = g.tx / 2 + 10;  // tile / 2 + 1 + 8 halo, padded
  g.cys = g.ty / 2 + 10;
  g.xs_off = (g.cys * g.cxs + 1) & ~1;
  return g;
}
__host__ __device__ inline int geom_lds_doubles(const Geom &g) {
  return g.xs_off + g.cys * g.tx;
}

// LDS budget per workgroup (doubles): the largest shape in use, TX = 64
// with RP = 2 (TY = 8), needs 14 x 42 + 14 x 64 = 1484.
constexpr int kLdsDoubles = 1488;
#define CXP_NT 128
// Coarse values per thread and plane: at most (TX/2 + 9) x (TY/2 + 9),
// <= 41 x 11 = 451 <= 4 * 128 for the RP = 1 shapes (4 <= TX <= 64) and
// <= 41 x 13 = 533 <= 5 * 128 for the RP = 2 ones (8 <= TX <= 64).
#define CXP_PRE 5
__host__ __device__ constexpr int sep_pre(int rp) { return rp == 1 ? 4 : 5; }
// Register window of coarse z-planes: the widest z stencil spans 9 planes.
#define CXP_ZW 9

// x-pass: X(fine x pair, coarse y row) from the coarse plane tile.  One work
// item computes both fine offsets of a coarse x and writes them as a pair;
// lanes run over the TX/2 coarse x of a row, rows step by 256 / TX.
template <int SID>
__device__ __forceinline__ void sep_xpass(const Geom &g,
                                          const double *__restrict__ Cs,
                                          double *__restrict__ Xs, int CY,
                                          int npx, int x0, int ci0, int tid) {
  const int q = tid & ((g.tx >> 1) - 1);
  if (q >= npx)
    return;
  const int ic0 = x0 >> 1;
  for (int cy = tid >> (g.lx - 1); cy < CY; cy += (256 >> g.lx)) {
    const double *row = Cs + cy * g.cxs + (ic0 + q - ci0);
    const auto c = [&](int o) { return row[o]; };
    double2 v;
    v.x = Op1<SID>::apply(c, 0);
    v.y = Op1<SID>::apply(c, 1);
    *reinterpret_cast<double2 *>(Xs + cy * g.tx + 2 * q) = v;
  }
}

// y-pass: both fine offsets of the coarse y row0 (local row of the x-pass
// buffer).
template <int SID>
__device__ __forceinline__ void sep_ypass(const Geom &g,
                                          const double *__restrict__ Xs,
                                          int row0, int fi, double &y0v,
                                          double &y1v) {
  const double *col = Xs + (row0 << g.lx) + fi;
  const auto c = [&](int o) { return col[o << g.lx]; };
  y0v = Op1<SID>::apply(c, 0);
  y1v = Op1<SID>::apply(c, 1);
}

// z-pass from the register window: w[ZW-1] holds coarse plane kc + whi.
template <int SID>
__device__ __forceinline__ double sep_zval(const double (&w)[CXP_ZW],
                                           int off) {
  constexpr int whi = sid_whi(SID);
  return Op1<SID>::apply([&](int o) { return w[CXP_ZW - 1 - whi + o]; },
                         off);
}

// Process one tile [x0, x0+TX) x [y0, y0+TY) x [z0, z1] of a task, clipped
// to the task's fine region (x0, y0 even).  All threads of the block call it.
// SU >= 0: every direction uses stencil SU, compiled without dispatch;
// SU < 0: per-pass runtime dispatch on the task's stencils.  RP: row pairs
// per thread (the geometry's).
template <int SU = -1, int RP = 1>
__device__ inline void sep_tile(const Geom &g, const SepTask &t,
                                long long comp, int x0, int y0, int z0, int z1,
                                double *smem) {
  constexpr int NR = 2 * RP;  // fine rows per thread
  double *Cs = smem;
  double *Xs = Cs + g.xs_off;
  const int tid = threadIdx.x;
  const int xe = min(x0 + g.tx - 1, t.rhi[0]);  // last fine x of the tile
  const int ye = min(y0 + g.ty - 1, t.rhi[1]);
  const int ci0 = (x0 >> 1) + t.wlo[0];
  const int ci1 = (xe >> 1) + t.whi[0];
  const int cj0 = (y0 >> 1) + t.wlo[1];
  const int cj1 = (ye >> 1) + t.whi[1];
  const int CX = ci1 - ci0 + 1;
  const int CY = cj1 - cj0 + 1;
  const int npx = (xe >> 1) - (x0 >> 1) + 1;  // coarse x of the tile
  const int kf = (z0 >> 1) + t.wlo[2];
  const int kl = (z1 >> 1) + t.whi[2];
  const int kc0 = z0 >> 1;
  const int kc1 = z1 >> 1;

  DFab crse = t.crse;
  crse.p += comp * t.ccs;
  DFab fine = t.fine;
  fine.p += comp * t.fcs;

  const int sx = t.sid[0], sy = t.sid[1], sz = t.sid[2];
  const int ncl = CX * CY;

  // owned fine points: x = x0 + fi, y = y0 + NR fjq + r, r < NR
  const int fi = tid & (g.tx - 1);
  const int fjq = tid >> g.lx;
  const int fx = x0 + fi;
  const int fy = y0 + NR * fjq;
  const bool xin = fx >= t.rlo[0] && fx <= xe;
  bool yin[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r)
    yin[r] = fy + r >= t.rlo[1] && fy + r <= ye;
  const bool own = xin && fy <= ye;  // the rows overlap the tile's rows
  const int row0 = (fy >> 1) - cj0;

  // Each thread moves up to CXP_PRE coarse values per plane; the next
  // plane's loads are issued before the current plane is processed so their
  // latency overlaps the passes.
  constexpr int PRE = sep_pre(RP);
  const __amdgpu_buffer_rsrc_t rc = make_rsrc(crse.p, t.cbytes);
  const __amdgpu_buffer_rsrc_t rf = make_rsrc(fine.p, t.fbytes);
  int soff[PRE], loff[PRE];
#pragma unroll
  for (int r = 0; r < PRE; ++r) {
    const int idx = tid + r * CXP_NT;
    const int cy = idx / CX;
    const int cx = idx - cy * CX;
    soff[r] = idx < ncl ? cy * g.cxs + cx : -1;
    loff[r] = 8 * (cy * crse.sy + cx);
  }
  const int pstride = 8 * (int)crse.sz;  // bytes per coarse plane
  int pbase = 8 * (int)didx(crse, ci0, cj0, kf);
  double pre[PRE];
#pragma unroll
  for (int r = 0; r < PRE; ++r)
    pre[r] = bld(rc, soff[r] >= 0 ? pbase + loff[r] : CXP_OOB);
  // dummy stores: every path into the plane loop then has the same pending
  // loads-then-stores pattern
#pragma unroll
  for (int r = 0; r < 2 * NR; ++r)
    bst(rf, CXP_OOB, 0.0);

  double w[NR][CXP_ZW];
#pragma unroll
  for (int r = 0; r < NR; ++r)
#pragma unroll
    for (int q = 0; q < CXP_ZW; ++q)
      w[r][q] = 0.0;
  // byte offset of the first owned row (x/y part) in the fine component
  const int fo0 = 8 * ((fx - fine.lo[0]) + (fy - fine.lo[1]) * fine.sy);
  const int frow = 8 * fine.sy;

  for (int kp = kf; kp <= kl; ++kp) {
    // coarse plane tile -> LDS, then prefetch the next plane (out of range
    // after the last plane)
#pragma unroll
    for (int r = 0; r < PRE; ++r)
      if (soff[r] >= 0)
        Cs[soff[r]] = pre[r];
    pbase += pstride;
    const bool more = kp < kl;
#pragma unroll
    for (int r = 0; r < PRE; ++r)
      pre[r] = bld(rc, (more && soff[r] >= 0) ? pbase + loff[r] : CXP_OOB);
    __syncthreads();
    if constexpr (SU >= 0) {
      sep_xpass<SU>(g, Cs, Xs, CY, npx, x0, ci0, tid);
    } else {
      switch (sx) {
#define CXP_X(S)                                                               \
  case S:                                                                      \
    sep_xpass<S>(g, Cs, Xs, CY, npx, x0, ci0, tid);                            \
    break;
        CXP_LINEAR_SIDS(CXP_X)
#undef CXP_X
      }
    }
    __syncthreads();
    // y-pass into the register window (rotated by one coarse plane)
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
      for (int q = 0; q < CXP_ZW - 1; ++q)
        w[r][q] = w[r][q + 1];
    double y[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r)
      y[r] = 0.0;
    if (own) {
#pragma unroll
      for (int p = 0; p < RP; ++p) {
        if constexpr (SU >= 0) {
          sep_ypass<SU>(g, Xs, row0 + p, fi, y[2 * p], y[2 * p + 1]);
        } else {
          switch (sy) {
#define CXP_Y(S)                                                               \
  case S:                                                                      \
    sep_ypass<S>(g, Xs, row0 + p, fi, y[2 * p], y[2 * p + 1]);                 \
    break;
            CXP_LINEAR_SIDS(CXP_Y)
#undef CXP_Y
          }
        }
      }
    }
#pragma unroll
    for (int r = 0; r < NR; ++r)
      w[r][CXP_ZW - 1] = y[r];
    // z-pass: both fine z of coarse plane kc = kp - whi; always 2 NR
    // stores, the ones outside the tile go out of range
    const int kc = kp - t.whi[2];
    const bool zact = kc >= kc0 && kc <= kc1;
    double v[NR][2];
#pragma unroll
    for (int r = 0; r < NR; ++r)
      v[r][0] = v[r][1] = 0.0;
    if (zact) {
      if constexpr (SU >= 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          v[r][0] = sep_zval<SU>(w[r], 0);
          v[r][1] = sep_zval<SU>(w[r], 1);
        }
      } else {
        switch (sz) {
#define CXP_Z(S)                                                               \
  case S:                                                                      \
    for (int r = 0; r < NR; ++r) {                                             \
      v[r][0] = sep_zval<S>(w[r], 0);                                          \
      v[r][1] = sep_zval<S>(w[r], 1);                                          \
    }                                                                          \
    break;
          CXP_LINEAR_SIDS(CXP_Z)
#undef CXP_Z
        }
      }
    }
    const int fz0 = 2 * kc, fz1 = 2 * kc + 1;
    const bool z0ok = zact && fz0 >= z0 && fz0 <= z1;
    const bool z1ok = zact && fz1 >= z0 && fz1 <= z1;
    const int zb0 = 8 * (int)((long long)(fz0 - fine.lo[2]) * fine.sz);
    const int zb1 = zb0 + 8 * (int)fine.sz;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      bst(rf, (own && yin[r] && z0ok) ? fo0 + r * frow + zb0 : CXP_OOB,
          v[r][0]);
      bst(rf, (own && yin[r] && z1ok) ? fo0 + r * frow + zb1 : CXP_OOB,
          v[r][1]);
    }
  }
}

// Per-box launch: grid.x = tiles (x fastest, then y, then z-chunks),
// grid.y = component.  The tile shape is uniform over the launch.
template <int SU, int RP>
__global__ void __launch_bounds__(CXP_NT)
    k_sep_box(SepTask t, int lx, int ntx, int nty, int zchunk) {
  __shared__ double smem[kLdsDoubles];
  const Geom g = make_geom(lx, RP);
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int tx = b % ntx;
  const int ty = (b / ntx) % nty;
  const int tz = b / (ntx * nty);
  const int x0 = (t.rlo[0] & ~1) + tx * g.tx;
  const int y0 = (t.rlo[1] & ~1) + ty * g.ty;
  const int z0 = t.rlo[2] + tz * zchunk;
  const int z1 = min(t.rhi[2], z0 + zchunk - 1);
  sep_tile<SU, RP>(g, t, blockIdx.y, x0, y0, z0, z1, smem);
}

// ---------------------------------------------------------------------------
// Poisoning: one value into every fine point of a list's patches, as
// CarpetX's poison_invalid_gf writes the poison pattern into invalid ghost
// points before they are filled (CarpetX/src/valid.cxx:157-215).  A run is
// one (patch, component) region, rows of nx points; a chunk is a run of
// rows one workgroup writes, about 2048 points.  Stores only, one per point,
// x fastest, so its time is the floor of the list's write pattern.

struct PoisonRun {
  double *p;     // the region's first point (lo corner) in its component
  int sy;        // x-row stride
  long long sz;  // plane stride
  int nx, ny;    // region extents in x and y
};
struct PoisonChunk {
  int run;
  int row0, nrows;  // rows (y fastest, then z) of the region
};

__global__ void __launch_bounds__(256)
    k_poison(const PoisonRun *__restrict__ runs,
             const PoisonChunk *__restrict__ chunks, double value) {
  const PoisonChunk ch = chunks[blockIdx.x];
  const PoisonRun r = runs[ch.run];
  const int npts = ch.nrows * r.nx;
  for (int i = threadIdx.x; i < npts; i += 256) {
    const int rr = i / r.nx;
    const int x = i - rr * r.nx;
    const int row = ch.row0 + rr;
    const int z = row / r.ny;
    const int y = row - z * r.ny;
    gst(r.p, x + (long long)y * r.sy + z * r.sz, value);
  }
}

// ---------------------------------------------------------------------------
// Coarse-patch gather: copies of coarse-level regions into the coarse
// patches a level fill reads when no single coarse FAB covers a patch's
// CoarseBox (the reference's ParallelCopy into its coarse patch,
// CarpetX/src/fillpatch.cxx:86-93).  A run is one (region, component): rows
// of nx points from a source FAB to an offset in the list's patch buffer; a
// chunk is about 2048 points of one run.

struct GatherRun {
  const double *src;  // the region's first point in the source component
  double *dstp;       // destination component, or null: the patch buffer
  long long dst;      // the region's first point (doubles) from dstp
  int src_sy;
  long long src_sz;
  int dst_sy;
  long long dst_sz;
  int nx, ny;
};
struct GatherChunk {
  int run;
  int row0, nrows;
};

__device__ __forceinline__ void gather_chunk(const GatherRun *__restrict__ runs,
                                             const GatherChunk &ch,
                                             double *__restrict__ buf) {
  const GatherRun r = runs[ch.run];
  double *const d = r.dstp ? r.dstp : buf;
  const int npts = ch.nrows * r.nx;
  for (int i = threadIdx.x; i < npts; i += 256) {
    const int rr = i / r.nx;
    const int x = i - rr * r.nx;
    const int row = ch.row0 + rr;
    const int z = row / r.ny;
    const int y = row - z * r.ny;
    const double v = gld(r.src, x + (long long)y * r.src_sy + z * r.src_sz);
    gst(d, r.dst + x + (long long)y * r.dst_sy + z * r.dst_sz, v);
  }
}

// Boundary values of a box (a gathered coarse patch, or a FAB in place):
// every point of the box outside the domain [imin, imax) takes the value
// CarpetX's BoundaryCondition::apply gives it (CarpetX/src/
// boundaries_impl.hxx:322-437), from the box's own points: inside the
// domain in every direction whose face has a condition.  So no point both
// reads and is written.  A run is one (box, component); a chunk is a range
// of the points (x fastest) of one part of the box outside the domain.
struct BcRun {
  double *p;            // the component's box, or null: the patch buffer
  long long off;        // the component's first point from p
  int lo[3], hi[3];     // patch box
  int imin[3], imax[3]; // domain, imax exclusive
  int type[2][3];       // CXP_BC_* per (face, direction)
  int refl[2][3];       // reflection offsets: src = refl - dst
  int parity[3];
  double dirichlet, robin;
  double xmin[3], dx[3];
};
struct BcChunk {
  int run;
  int blo[3];    // the box's first point
  int bnx, bny;  // its x and y extents
  int first, count;
  int direct;    // >= 0: the sources' coarse FAB (a BcDirect); -1: the box
};
// The coarse FAB that holds every source of a chunk's points in its valid
// region: the run's component there, and the periodic shift that takes a
// source point (domain coordinates as the condition gives them) to the
// point that holds its value.
struct BcDirect {
  const double *src;  // component at its FAB's lo
  int lo[3];
  int sy;
  long long sz;
  int shift[3];
};

// The boundary value of point dst of run r, or false if dst is not a
// boundary point of a face with a condition.  fetch(p) gives the value at
// source point p.
template <class Fetch>
__device__ __forceinline__ bool bc_value(const BcRun &r, const int dst[3],
                                         Fetch fetch, double &val) {
  int src[3] = {dst[0], dst[1], dst[2]}, delta[3] = {0, 0, 0};
  bool outside = false, dirichlet = false, lin = false, refl = false;
  int robin = 0;  // bit d: Robin in direction d
  double parity = 1.0;
  for (int d = 0; d < 3; ++d) {
    if (dst[d] >= r.imin[d] && dst[d] < r.imax[d])
      continue;  // not a boundary direction of this point
    const int face = dst[d] >= r.imax[d] ? 1 : 0;
    if (r.type[face][d] == CXP_BC_NONE || r.type[face][d] == CXP_BC_PERIODIC)
      continue;  // nor is a face without a condition, or a periodic one
    outside = true;
    const int edge = face ? r.imax[d] - 1 : r.imin[d];
    switch (r.type[face][d]) {
    case CXP_BC_DIRICHLET:
      dirichlet = true;
      break;
    case CXP_BC_LINEAR_EXTRAPOLATION:
      src[d] = edge;
      delta[d] = face ? -1 : 1;  // -inormal
      lin = true;
      break;
    case CXP_BC_NEUMANN:
      src[d] = edge;
      break;
    case CXP_BC_ROBIN:
      src[d] = edge;
      robin |= 1 << d;
      break;
    case CXP_BC_REFLECTION:
      src[d] = r.refl[face][d] - dst[d];
      parity *= (double)r.parity[d];
      refl = true;
      break;
    default:
      break;
    }
  }
  if (!outside)
    return false;
  if (dirichlet) {
    val = r.dirichlet;
    return true;
  }
  val = fetch(src);
  for (int d = 0; d < 3; ++d)
    if (robin & (1 << d)) {
      double xb[3], xi[3];
      for (int e = 0; e < 3; ++e) {
        xb[e] = r.xmin[e] + (double)dst[e] * r.dx[e];
        xi[e] = r.xmin[e] + (double)src[e] * r.dx[e];
      }
      const double rb = sqrt(xb[0] * xb[0] + xb[1] * xb[1] + xb[2] * xb[2]);
      const double ri = sqrt(xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2]);
      const double qr = ri / rb;
      val = r.robin + qr * (val - r.robin);
    }
  if (lin) {
    const int sp[3] = {src[0] + delta[0], src[1] + delta[1],
                       src[2] + delta[2]};
    const double grad = val - fetch(sp);
    int num = 0, den = 0;
    for (int e = 0; e < 3; ++e) {
      num += (dst[e] - src[e]) * (dst[e] - src[e]);
      den += delta[e] * delta[e];
    }
    // integer quotient, as the reference divides two int sums
    val += sqrt((double)(num / den)) * grad;
  }
  if (refl)
    val *= parity;
  return true;
}

// One chunk of boundary points: sources from the box itself (k_bc, after
// the gather) or from one coarse FAB (inside the gather launch).
__device__ __forceinline__ void bc_chunk(const BcRun &r, const BcChunk &ch,
                                         const BcDirect *dir, double *buf) {
  const int nx = r.hi[0] - r.lo[0] + 1, ny = r.hi[1] - r.lo[1] + 1;
  double *v = (r.p ? r.p : buf) + r.off;
  auto at = [&](const int p[3]) {
    return (long long)(p[0] - r.lo[0]) + (long long)(p[1] - r.lo[1]) * nx +
           (long long)(p[2] - r.lo[2]) * nx * ny;
  };
  // the host sizes the box to hold every source; a source outside it would
  // be a host error, and gives NaN instead of a stray read
  auto in_box = [&](const int p[3]) {
    return p[0] >= r.lo[0] && p[0] <= r.hi[0] && p[1] >= r.lo[1] &&
           p[1] <= r.hi[1] && p[2] >= r.lo[2] && p[2] <= r.hi[2];
  };
  auto from_box = [&](const int p[3]) {
    return in_box(p) ? v[at(p)] : __builtin_nan("");
  };
  auto from_fab = [&](const int p[3]) {
    return gld(dir->src, (long long)(p[0] + dir->shift[0] - dir->lo[0]) +
                             (long long)(p[1] + dir->shift[1] - dir->lo[1]) *
                                 dir->sy +
                             (long long)(p[2] + dir->shift[2] - dir->lo[2]) *
                                 dir->sz);
  };
  for (int q = ch.first + (int)threadIdx.x; q < ch.first + ch.count;
       q += 256) {
    const int dst[3] = {ch.blo[0] + q % ch.bnx, ch.blo[1] + (q / ch.bnx) % ch.bny,
                        ch.blo[2] + q / (ch.bnx * ch.bny)};
    double val;
    const bool b = dir ? bc_value(r, dst, from_fab, val)
                       : bc_value(r, dst, from_box, val);
    if (b)
      v[at(dst)] = val;
  }
}

__global__ void __launch_bounds__(256)
    k_bc(const BcRun *__restrict__ runs, const BcChunk *__restrict__ chunks,
         double *__restrict__ buf) {
  const BcChunk ch = chunks[blockIdx.x];
  const BcRun r = runs[ch.run];
  bc_chunk(r, ch, nullptr, buf);
}

// The gather launch: blocks below ngather copy (coarse patches, old-level
// data), the others compute boundary values whose sources one coarse FAB
// holds; neither reads what the other writes.
__global__ void __launch_bounds__(256)
    k_gather(const GatherRun *__restrict__ runs,
             const GatherChunk *__restrict__ chunks, int ngather,
             const BcRun *__restrict__ bruns, const BcChunk *__restrict__ bchunks,
             const BcDirect *__restrict__ bdir, double *__restrict__ buf) {
  if ((int)blockIdx.x < ngather) {
    gather_chunk(runs, chunks[blockIdx.x], buf);
    return;
  }
  const BcChunk ch = bchunks[blockIdx.x - ngather];
  const BcRun r = bruns[ch.run];
  bc_chunk(r, ch, bdir + ch.direct, buf);
}

// ---------------------------------------------------------------------------
// Point kernels for ENO / fallback operators

struct PointTask {
  DFab crse;
  DFab fine;
  int rlo[3], rhi[3];
  int sid[3];
  int fb_per_var, fb_per_group;
  int ncomp;
  long long ccs, fcs;
  // |undivided differences| of the coarse data along x, y, z (3 arrays of
  // ccs values laid out like crse, already the max over the components the
  // search covers), or null: computed from crse by the search itself
  const double *d2;
};

__device__ __forceinline__ double crse_at(const DFab &c, int i, int j, int k) {
  return c.p[didx(c, i, j, k)];
}

// Tensor product with shifts (call_stencil_3d; x innermost).
__device__ inline double pt_tensor(const DFab &c, const int sid[3],
                                   const int shift[3], const int off[3],
                                   const int ic[3]) {
  int ylo, yhi, zlo, zhi;
  stencil_radius(sid[1], shift[1], off[1], ylo, yhi);
  stencil_radius(sid[2], shift[2], off[2], zlo, zhi);
  double yb[9];
  for (int dk = zlo; dk <= zhi; ++dk) {
    double xb[9];
    for (int dj = ylo; dj <= yhi; ++dj) {
      const double *row = c.p + didx(c, ic[0], ic[1] + dj, ic[2] + dk);
      xb[dj + 4] = apply1d(sid[0], [&](int o) { return row[o]; }, shift[0],
                           off[0]);
    }
    yb[dk + 4] = apply1d(sid[1], [&](int o) { return xb[o + 4]; }, shift[1],
                         off[1]);
  }
  return apply1d(sid[2], [&](int o) { return yb[o + 4]; }, shift[2], off[2]);
}

// Second undivided difference (CC ENO 2; weights +1 -2 +1, :750-773).
__device__ __forceinline__ double undiv2(double m, double c, double p) {
  double dd = -2.0 * c;
  dd += 1.0 * (m + p);
  return dd;
}

// ENO shift search (:1417-1517 per-var; per-group takes the max over comps).
// Score of one shift candidate (si, sj, sk): the penalised largest
// undivided difference over the transverse stencil lines of every ENO
// direction (prolongate_3d_rf2_impl.hxx:1424-1503).
__device__ inline double pt_eno_dd(const PointTask &t, long long comp0,
                                   int ncomp, const int off[3],
                                   const int ic[3], int si, int sj, int sk) {
  int use[3];
  for (int d = 0; d < 3; ++d)
    use[d] = t.sid[d] == SID_ENO2;
  const int sv[3] = {si, sj, sk};
  int r[3][2];
  for (int d = 0; d < 3; ++d) {
    stencil_radius(t.sid[d], sv[d], off[d], r[d][0], r[d][1]);
    if (use[d]) {
      r[d][0] -= sv[d];
      r[d][1] -= sv[d];
    }
  }
  double ddx = 0, ddy = 0, ddz = 0;
  if (t.d2) {
    const DFab &c = t.crse;
    const double *dx2 = t.d2, *dy2 = t.d2 + t.ccs, *dz2 = t.d2 + 2 * t.ccs;
    if (use[0])
      for (int dk = r[2][0]; dk <= r[2][1]; ++dk)
        for (int dj = r[1][0]; dj <= r[1][1]; ++dj)
          ddx = fmax(ddx, dx2[didx(c, ic[0] + si, ic[1] + sj + dj,
                                   ic[2] + sk + dk)]);
    if (use[1])
      for (int dk = r[2][0]; dk <= r[2][1]; ++dk)
        for (int di = r[0][0]; di <= r[0][1]; ++di)
          ddy = fmax(ddy, dy2[didx(c, ic[0] + si + di, ic[1] + sj,
                                   ic[2] + sk + dk)]);
    if (use[2])
      for (int dj = r[1][0]; dj <= r[1][1]; ++dj)
        for (int di = r[0][0]; di <= r[0][1]; ++di)
          ddz = fmax(ddz, dz2[didx(c, ic[0] + si + di, ic[1] + sj + dj,
                                   ic[2] + sk)]);
  }
  for (int cc = 0; cc < (t.d2 ? 0 : ncomp); ++cc) {
    DFab c = t.crse;
    c.p += (comp0 + cc) * t.ccs;
    if (use[0])
      for (int dk = r[2][0]; dk <= r[2][1]; ++dk)
        for (int dj = r[1][0]; dj <= r[1][1]; ++dj) {
          const int i = ic[0] + si, j = ic[1] + sj + dj, k = ic[2] + sk + dk;
          ddx = fmax(ddx, fabs(undiv2(crse_at(c, i - 1, j, k),
                                      crse_at(c, i, j, k),
                                      crse_at(c, i + 1, j, k))));
        }
    if (use[1])
      for (int dk = r[2][0]; dk <= r[2][1]; ++dk)
        for (int di = r[0][0]; di <= r[0][1]; ++di) {
          const int i = ic[0] + si + di, j = ic[1] + sj, k = ic[2] + sk + dk;
          ddy = fmax(ddy, fabs(undiv2(crse_at(c, i, j - 1, k),
                                      crse_at(c, i, j, k),
                                      crse_at(c, i, j + 1, k))));
        }
    if (use[2])
      for (int dj = r[1][0]; dj <= r[1][1]; ++dj)
        for (int di = r[0][0]; di <= r[0][1]; ++di) {
          const int i = ic[0] + si + di, j = ic[1] + sj + dj, k = ic[2] + sk;
          ddz = fmax(ddz, fabs(undiv2(crse_at(c, i, j, k - 1),
                                      crse_at(c, i, j, k),
                                      crse_at(c, i, j, k + 1))));
        }
  }
  // 1 + sqrt(DBL_EPSILON) * (|si| + |sj| + |sk|); sqrt(eps) = 2^-26
  const double penalty = 1.0 + 0x1p-26 * (double)(abs(si) + abs(sj) + abs(sk));
  return penalty * fmax(fmax(ddx, ddy), ddz);
}

// ENO shift search: candidates in the reference's lexicographic order (sk
// outermost, si innermost), the first strictly smallest score wins
// (:1418-1517).
__device__ inline void pt_eno_search(const PointTask &t, long long comp0,
                                     int ncomp, const int off[3],
                                     const int ic[3], int shift[3]) {
  int ms[3];
  for (int d = 0; d < 3; ++d)
    ms[d] = t.sid[d] == SID_ENO2 ? 1 : 0;
  shift[0] = shift[1] = shift[2] = 0;
  double min_dd = __builtin_inf();
  for (int sk = -ms[2]; sk <= ms[2]; ++sk)
    for (int sj = -ms[1]; sj <= ms[1]; ++sj)
      for (int si = -ms[0]; si <= ms[0]; ++si) {
        const double dd = pt_eno_dd(t, comp0, ncomp, off, ic, si, sj, sk);
        if (dd < min_dd) {
          min_dd = dd;
          shift[0] = si;
          shift[1] = sj;
          shift[2] = sk;
        }
      }
}

__device__ inline void pt_minmax(const DFab &c, const int sr[3][2],
                                 const int ic[3], double &mn, double &mx) {
  for (int dk = sr[2][0]; dk <= sr[2][1]; ++dk)
    for (int dj = sr[1][0]; dj <= sr[1][1]; ++dj)
      for (int di = sr[0][0]; di <= sr[0][1]; ++di) {
        const double v = crse_at(c, ic[0] + di, ic[1] + dj, ic[2] + dk);
        mn = fmin(mn, v);
        mx = fmax(mx, v);
      }
}

// Slope sign change along d (:1609-1664): s0 = c(ic+1) - c(ic) compared with
// each difference from srad[d][0]+2 up.
__device__ inline bool pt_slope_change(const DFab &c, int d, const int sr[3][2],
                                       const int ic[3]) {
  const int a = d == 0 ? 1 : 0;
  const int b = d == 2 ? 1 : 2;
  bool need = false;
  for (int tb = sr[b][0]; tb <= sr[b][1]; ++tb)
    for (int ta = sr[a][0]; ta <= sr[a][1]; ++ta) {
      int p[3];
      p[a] = ic[a] + ta;
      p[b] = ic[b] + tb;
      p[d] = ic[d];
      const double c0 = crse_at(c, p[0], p[1], p[2]);
      p[d] = ic[d] + 1;
      const double s0 = crse_at(c, p[0], p[1], p[2]) - c0;
      for (int tt = sr[d][0] + 2; tt <= sr[d][1]; ++tt) {
        p[d] = ic[d] + tt;
        const double v1 = crse_at(c, p[0], p[1], p[2]);
        p[d] = ic[d] + tt - 1;
        const double v0 = crse_at(c, p[0], p[1], p[2]);
        const double s = v1 - v0;
        need |= s * s0 < 0;
      }
    }
  return need;
}

// Branch of a per-var operator at one fine point and component: the ENO
// shift and, for fallback operators, whether the linear value replaces the
// high-order one (prolongate_3d_rf2_impl.hxx:1417-1667).  val is the
// high-order value.
// given: the ENO shift, already searched (or null: search here).
__device__ inline void pt_decide_var(const PointTask &t, long long comp,
                                     const int off[3], const int ic[3],
                                     int shift[3], double &val, bool &need,
                                     const int *given = nullptr) {
  DFab c = t.crse;
  c.p += comp * t.ccs;
  shift[0] = shift[1] = shift[2] = 0;
  const bool any_shift =
      t.sid[0] == SID_ENO2 || t.sid[1] == SID_ENO2 || t.sid[2] == SID_ENO2;
  if (given)
    for (int d = 0; d < 3; ++d)
      shift[d] = given[d];
  else if (any_shift)
    pt_eno_search(t, comp, 1, off, ic, shift);
  val = pt_tensor(c, t.sid, shift, off, ic);
  need = false;
  if (t.fb_per_var) {
    int sr[3][2];
    for (int d = 0; d < 3; ++d)
      stencil_radius(t.sid[d], shift[d], off[d], sr[d][0], sr[d][1]);
    double mn = __builtin_inf(), mx = -__builtin_inf();
    pt_minmax(c, sr, ic, mn, mx);
    need = val < mn || val > mx;
    for (int d = 0; d < 3; ++d) {
      const SidInfo s = sid_info(t.sid[d]);
      if (s.intp == INTP_CONS || s.intp == INTP_ENO)
        need |= pt_slope_change(c, d, sr, ic);
    }
  }
}

// Branch of a per-group operator at one fine point, shared by all ncomp
// components (prolongate_3d_rf2_impl.hxx:1836-2148); vals[c] are the
// high-order values.
__device__ inline void pt_decide_group(const PointTask &t, const int off[3],
                                       const int ic[3], int shift[3],
                                       double (&vals)[16], bool &need,
                                       const int *given = nullptr) {
  const int nc = t.ncomp;
  shift[0] = shift[1] = shift[2] = 0;
  const bool any_shift =
      t.sid[0] == SID_ENO2 || t.sid[1] == SID_ENO2 || t.sid[2] == SID_ENO2;
  if (given)
    for (int d = 0; d < 3; ++d)
      shift[d] = given[d];
  else if (any_shift)
    pt_eno_search(t, 0, nc, off, ic, shift);
  for (int cc = 0; cc < nc; ++cc) {
    DFab c = t.crse;
    c.p += cc * t.ccs;
    vals[cc] = pt_tensor(c, t.sid, shift, off, ic);
  }
  need = false;
  if (t.fb_per_group) {
    int sr[3][2];
    for (int d = 0; d < 3; ++d)
      stencil_radius(t.sid[d], shift[d], off[d], sr[d][0], sr[d][1]);
    double mn = __builtin_inf(), mx = -__builtin_inf();
    for (int cc = 0; cc < nc; ++cc) {
      DFab c = t.crse;
      c.p += cc * t.ccs;
      pt_minmax(c, sr, ic, mn, mx);  // accumulates across comps (:2010-2024)
      need |= vals[cc] < mn || vals[cc] > mx;
    }
    for (int d = 0; d < 3; ++d)
      if (sid_info(t.sid[d]).intp == INTP_CONS)
        for (int cc = 0; cc < nc; ++cc) {
          DFab c = t.crse;
          c.p += cc * t.ccs;
          need |= pt_slope_change(c, d, sr, ic);
        }
  }
}

// Fine point gid of a region in x-fastest order.
__device__ __forceinline__ void pt_point(const PointTask &t, long long gid,
                                         int f[3], int ic[3], int off[3]) {
  const int nx = t.rhi[0] - t.rlo[0] + 1;
  const int ny = t.rhi[1] - t.rlo[1] + 1;
  f[0] = t.rlo[0] + (int)(gid % nx);
  f[1] = t.rlo[1] + (int)((gid / nx) % ny);
  f[2] = t.rlo[2] + (int)(gid / ((long long)nx * ny));
  for (int d = 0; d < 3; ++d) {
    ic[d] = f[d] >> 1;
    off[d] = f[d] & 1;
  }
}

// Branch code of a fine point (for the adjoint): (sx+1) + 3 (sy+1) +
// 9 (sz+1) + 27 need.
__device__ __forceinline__ unsigned char adj_code(const int shift[3],
                                                  bool need) {
  return (unsigned char)((shift[0] + 1) + 3 * (shift[1] + 1) +
                         9 * (shift[2] + 1) + (need ? 27 : 0));
}

__device__ __forceinline__ long long pt_npts(const PointTask &t) {
  return (long long)(t.rhi[0] - t.rlo[0] + 1) * (t.rhi[1] - t.rlo[1] + 1) *
         (t.rhi[2] - t.rlo[2] + 1);
}

// codes (or null): also record each fine point's branch (adj_code), per
// component, region-linear.
__global__ void __launch_bounds__(256)
    k_point_per_var(PointTask t, unsigned char *__restrict__ codes = nullptr) {
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= pt_npts(t))
    return;
  const long long comp = blockIdx.y;
  int f[3], ic[3], off[3], shift[3];
  pt_point(t, gid, f, ic, off);
  double val;
  bool need;
  pt_decide_var(t, comp, off, ic, shift, val, need);
  if (codes)
    codes[comp * pt_npts(t) + gid] = adj_code(shift, need);
  double res = val;
  if (need) {
    DFab c = t.crse;
    c.p += comp * t.ccs;
    int sl[3];
    for (int d = 0; d < 3; ++d)
      sl[d] = sid_linear_per_var(t.sid[d]);
    const int zero[3] = {0, 0, 0};
    res = pt_tensor(c, sl, zero, off, ic);
  }
  DFab fo = t.fine;
  fo.p += comp * t.fcs;
  fo.p[didx(fo, f[0], f[1], f[2])] = res;
}

__global__ void __launch_bounds__(256)
    k_point_per_group(PointTask t, unsigned char *__restrict__ codes = nullptr) {
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= pt_npts(t))
    return;
  int f[3], ic[3], off[3], shift[3];
  pt_point(t, gid, f, ic, off);
  double vals[16];
  bool need;
  pt_decide_group(t, off, ic, shift, vals, need);
  if (codes)
    codes[gid] = adj_code(shift, need);
  int sl[3];
  for (int d = 0; d < 3; ++d)
    sl[d] = sid_linear_per_group(t.sid[d]);
  const int zero[3] = {0, 0, 0};
  for (int cc = 0; cc < t.ncomp; ++cc) {
    double v = vals[cc];
    if (need) {
      DFab c = t.crse;
      c.p += cc * t.ccs;
      v = pt_tensor(c, sl, zero, off, ic);
    }
    DFab fo = t.fine;
    fo.p += cc * t.fcs;
    fo.p[didx(fo, f[0], f[1], f[2])] = v;
  }
}

// LDS-tiled point kernels: a workgroup owns an 8 x 8 x 8 fine tile and first
// copies CoarseBox(tile) of the coarse FAB into LDS (every read of the
// decision and value code for the tile's points lies inside it, as for the
// whole region, prolongate_3d_rf2_impl.hxx:1226-1254), then runs the same
// per-point code reading the LDS copy through a DFab view.  Same operations
// in the same order, so the results equal k_point_per_var/_per_group bit for
// bit; the ENO search's undivided differences come from LDS instead of
// global memory.
#define CXP_PT_TILE 8

// CoarseBox of fine [f0, f1] in direction d for stencil sid.
__host__ __device__ inline void pt_coarse_range(int sid, int f0, int f1,
                                                int &c0, int &c1) {
  const SidInfo s = sid_info(sid);
  c0 = (f0 >> 1) - s.ghosts;
  c1 = (f1 >> 1) + ((s.cent == CENT_VC && (f1 & 1)) ? 1 : 0) + s.ghosts;
}

// Largest coarse block (doubles per component) of a tile for these stencils;
// the tiled kernels hold ncomp such blocks plus three of undivided
// differences.
__host__ __device__ inline int pt_tile_block(const int sid[3]) {
  int v = 1;
  for (int d = 0; d < 3; ++d)
    v *= CXP_PT_TILE / 2 + 1 + 2 * sid_info(sid[d]).ghosts;
  return v;
}

// The ENO shift search of a fine point reads the coarse parent, the shift
// candidates and, per direction, the stencil radius at the point's fine
// offset (prolongate_3d_rf2_impl.hxx:1418-1517).  In a direction whose
// radius does not depend on the offset (every ENO direction: the support
// moves with the shift only), the two fine children of a coarse point run
// the same search.
__host__ __device__ inline bool radius_depends_on_off(int sid) {
  int a0, a1, b0, b1;
  stencil_radius(sid, 0, 0, a0, a1);
  stencil_radius(sid, 0, 1, b0, b1);
  return a0 != b0 || a1 != b1;
}
// Search units of a tile: (coarse parent, fine offset in the directions
// whose radius depends on it), at most (CXP_PT_TILE / 2 + 1)^3 * 8.
#define CXP_PT_UNITS 1000


// CODES: record each fine point's branch in codes (per component in per-var
// mode, region-linear, x fastest) instead of writing its value; without
// CODES, a non-null codes records the branches beside the values.
template <bool PG, bool CODES = false>
__global__ void __launch_bounds__(256)
    k_point_tiled(PointTask t, int ntx, int nty,
                  unsigned char *__restrict__ codes = nullptr) {
  extern __shared__ double ptile[];
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int tv[3] = {b % ntx, (b / ntx) % nty, b / (ntx * nty)};
  int f0[3], f1[3], c0[3], c1[3], n[3];
  for (int d = 0; d < 3; ++d) {
    f0[d] = (t.rlo[d] & ~1) + CXP_PT_TILE * tv[d];
    f1[d] = min(f0[d] + CXP_PT_TILE - 1, t.rhi[d]);
    f0[d] = max(f0[d], t.rlo[d]);
    pt_coarse_range(t.sid[d], f0[d], f1[d], c0[d], c1[d]);
    n[d] = c1[d] - c0[d] + 1;
  }
  const int vol = n[0] * n[1] * n[2];
  const int ncb = PG ? t.ncomp : 1;
  const long long comp0 = PG ? 0 : blockIdx.y;
  for (int idx = threadIdx.x; idx < vol * ncb; idx += blockDim.x) {
    const int cc = idx / vol;
    const int r = idx - cc * vol;
    const int i = r % n[0], j = (r / n[0]) % n[1], k = r / (n[0] * n[1]);
    DFab c = t.crse;
    c.p += (comp0 + cc) * t.ccs;
    ptile[idx] = crse_at(c, c0[0] + i, c0[1] + j, c0[2] + k);
  }
  __syncthreads();
  // |undivided differences| along each ENO direction, max over the block's
  // components (the search's max over lines and components is taken with
  // fmax, which does not depend on the order)
  double *d2 = ptile + vol * ncb;
  const int stride[3] = {1, n[0], n[0] * n[1]};
  for (int d = 0; d < 3; ++d) {
    if (t.sid[d] != SID_ENO2)
      continue;
    for (int r = threadIdx.x; r < vol; r += blockDim.x) {
      const int id[3] = {r % n[0], (r / n[0]) % n[1], r / (n[0] * n[1])};
      double m = 0.0;
      if (id[d] > 0 && id[d] < n[d] - 1)
        for (int cc = 0; cc < ncb; ++cc) {
          const double *q = ptile + cc * vol + r;
          m = fmax(m, fabs(undiv2(q[-stride[d]], q[0], q[stride[d]])));
        }
      d2[d * vol + r] = m;
    }
  }
  __syncthreads();
  PointTask tl = t;
  tl.crse.p = ptile;
  for (int d = 0; d < 3; ++d)
    tl.crse.lo[d] = c0[d];
  tl.crse.sy = n[0];
  tl.crse.sz = (long long)n[0] * n[1];
  tl.ccs = vol;
  tl.d2 = d2;
  // one search per unit (the same code, so the same shift as each of the
  // unit's fine points would find), kept in LDS for the fine points
  __shared__ unsigned char ushift[CXP_PT_UNITS];
  const bool any_shift =
      t.sid[0] == SID_ENO2 || t.sid[1] == SID_ENO2 || t.sid[2] == SID_ENO2;
  int pc0[3], dep[3], nu[3];
  for (int d = 0; d < 3; ++d) {
    pc0[d] = f0[d] >> 1;
    dep[d] = radius_depends_on_off(t.sid[d]);
    nu[d] = ((f1[d] >> 1) - pc0[d] + 1) * (dep[d] ? 2 : 1);
  }
  if (any_shift) {
    // a group of 32 lanes per unit, a lane per candidate: lane c scores
    // candidate c (lexicographic, si fastest), and the group keeps the
    // smallest valid score, the smallest index on ties -- the candidate the
    // reference's scan (strict <, from an infinite minimum) ends with; no
    // valid score leaves the zero shift
    int ms[3];
    for (int d = 0; d < 3; ++d)
      ms[d] = t.sid[d] == SID_ENO2 ? 1 : 0;
    const int nw[3] = {2 * ms[0] + 1, 2 * ms[1] + 1, 2 * ms[2] + 1};
    const int ncand = nw[0] * nw[1] * nw[2];
    const int nunits = nu[0] * nu[1] * nu[2];
    for (int w = threadIdx.x; w < nunits * 32; w += blockDim.x) {
      const int u = w >> 5, c = w & 31;
      const int uv[3] = {u % nu[0], (u / nu[0]) % nu[1], u / (nu[0] * nu[1])};
      int ic[3], off[3];
      for (int d = 0; d < 3; ++d) {
        ic[d] = pc0[d] + (dep[d] ? uv[d] >> 1 : uv[d]);
        off[d] = dep[d] ? uv[d] & 1 : 0;
      }
      double dd = 0.0;
      int valid = 0;
      if (c < ncand) {
        dd = pt_eno_dd(tl, 0, PG ? t.ncomp : 1, off, ic, c % nw[0] - ms[0],
                       (c / nw[0]) % nw[1] - ms[1], c / (nw[0] * nw[1]) - ms[2]);
        valid = dd < __builtin_inf();  // false for inf and NaN
      }
      int idx = c;
      for (int m = 1; m < 32; m <<= 1) {
        const double od = __shfl_xor(dd, m, 32);
        const int oi = __shfl_xor(idx, m, 32);
        const int ov = __shfl_xor(valid, m, 32);
        if (ov && (!valid || od < dd || (od == dd && oi < idx))) {
          dd = od;
          idx = oi;
          valid = ov;
        }
      }
      if (c == 0) {
        int sh[3] = {0, 0, 0};
        if (valid) {
          sh[0] = idx % nw[0] - ms[0];
          sh[1] = (idx / nw[0]) % nw[1] - ms[1];
          sh[2] = idx / (nw[0] * nw[1]) - ms[2];
        }
        ushift[u] = (unsigned char)((sh[0] + 1) + 3 * (sh[1] + 1) + 9 * (sh[2] + 1));
      }
    }
    __syncthreads();
  }
  const int m[3] = {f1[0] - f0[0] + 1, f1[1] - f0[1] + 1, f1[2] - f0[2] + 1};
  for (int p = threadIdx.x; p < m[0] * m[1] * m[2]; p += blockDim.x) {
    const int f[3] = {f0[0] + p % m[0], f0[1] + (p / m[0]) % m[1],
                      f0[2] + p / (m[0] * m[1])};
    int ic[3], off[3], shift[3], given[3] = {0, 0, 0};
    for (int d = 0; d < 3; ++d) {
      ic[d] = f[d] >> 1;
      off[d] = f[d] & 1;
    }
    if (any_shift) {
      int uv[3];
      for (int d = 0; d < 3; ++d)
        uv[d] = dep[d] ? 2 * (ic[d] - pc0[d]) + off[d] : ic[d] - pc0[d];
      const int code = ushift[uv[0] + nu[0] * (uv[1] + nu[1] * uv[2])];
      given[0] = code % 3 - 1;
      given[1] = (code / 3) % 3 - 1;
      given[2] = code / 9 - 1;
    }
    bool need;
    if constexpr (CODES) {
      const long long rnx = t.rhi[0] - t.rlo[0] + 1;
      const long long rny = t.rhi[1] - t.rlo[1] + 1;
      const long long gid = (f[0] - t.rlo[0]) + (f[1] - t.rlo[1]) * rnx +
                            (f[2] - t.rlo[2]) * rnx * rny;
      if (PG) {
        double vals[16];
        pt_decide_group(tl, off, ic, shift, vals, need, given);
        codes[gid] = adj_code(shift, need);
      } else {
        double val;
        pt_decide_var(tl, 0, off, ic, shift, val, need, given);
        codes[(long long)blockIdx.y * pt_npts(t) + gid] = adj_code(shift, need);
      }
    } else if (!PG) {
      double val;
      pt_decide_var(tl, 0, off, ic, shift, val, need, given);
      if (codes) {
        const long long rnx = t.rhi[0] - t.rlo[0] + 1;
        const long long rny = t.rhi[1] - t.rlo[1] + 1;
        codes[(long long)blockIdx.y * pt_npts(t) + (f[0] - t.rlo[0]) +
              (f[1] - t.rlo[1]) * rnx + (f[2] - t.rlo[2]) * rnx * rny] =
            adj_code(shift, need);
      }
      double res = val;
      if (need) {
        int sl[3];
        for (int d = 0; d < 3; ++d)
          sl[d] = sid_linear_per_var(t.sid[d]);
        const int zero[3] = {0, 0, 0};
        res = pt_tensor(tl.crse, sl, zero, off, ic);
      }
      DFab fo = t.fine;
      fo.p += (long long)blockIdx.y * t.fcs;
      fo.p[didx(fo, f[0], f[1], f[2])] = res;
    } else {
      double vals[16];
      pt_decide_group(tl, off, ic, shift, vals, need, given);
      if (codes) {
        const long long rnx = t.rhi[0] - t.rlo[0] + 1;
        const long long rny = t.rhi[1] - t.rlo[1] + 1;
        codes[(f[0] - t.rlo[0]) + (f[1] - t.rlo[1]) * rnx +
              (f[2] - t.rlo[2]) * rnx * rny] = adj_code(shift, need);
      }
      int sl[3];
      for (int d = 0; d < 3; ++d)
        sl[d] = sid_linear_per_group(t.sid[d]);
      const int zero[3] = {0, 0, 0};
      for (int cc = 0; cc < t.ncomp; ++cc) {
        double v = vals[cc];
        if (need) {
          DFab c = tl.crse;
          c.p += cc * tl.ccs;
          v = pt_tensor(c, sl, zero, off, ic);
        }
        DFab fo = t.fine;
        fo.p += cc * t.fcs;
        fo.p[didx(fo, f[0], f[1], f[2])] = v;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Restriction and its adjoint

struct RestrictTask {
  DFab fine;
  DFab crse;
  int rlo[3], rhi[3];  // coarse region
  int cc[3];           // 1 for cell-centred directions
  long long fcs, ccs;
};

__global__ void __launch_bounds__(256) k_restrict(RestrictTask t) {
  const int nx = t.rhi[0] - t.rlo[0] + 1;
  const int ny = t.rhi[1] - t.rlo[1] + 1;
  const int nz = t.rhi[2] - t.rlo[2] + 1;
  const long long n = (long long)nx * ny * nz;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= n)
    return;
  const long long comp = blockIdx.y;
  const int i = t.rlo[0] + (int)(gid % nx);
  const int j = t.rlo[1] + (int)((gid / nx) % ny);
  const int k = t.rlo[2] + (int)(gid / ((long long)nx * ny));
  DFab f = t.fine;
  f.p += comp * t.fcs;
  DFab c = t.crse;
  c.p += comp * t.ccs;
  const int ncc = t.cc[0] + t.cc[1] + t.cc[2];
  double v;
  if (ncc == 0) {
    v = f.p[didx(f, 2 * i, 2 * j, 2 * k)];
  } else {
    double s = 0;
    for (int kk = 0; kk <= t.cc[2]; ++kk)
      for (int jj = 0; jj <= t.cc[1]; ++jj)
        for (int ii = 0; ii <= t.cc[0]; ++ii)
          s += f.p[didx(f, 2 * i + ii, 2 * j + jj, 2 * k + kk)];
    v = s * (1.0 / (double)(1 << ncc));
  }
  c.p[didx(c, i, j, k)] = v;
}

// fine_bar += R^T crse_bar over the children of the coarse region; each fine
// point has at most one parent, so the gather needs no atomics.
__global__ void __launch_bounds__(256) k_restrict_adjoint(RestrictTask t) {
  // iterate over fine points covered by the coarse region
  int flo[3], fn[3];
  for (int d = 0; d < 3; ++d) {
    flo[d] = 2 * t.rlo[d];
    fn[d] = 2 * (t.rhi[d] - t.rlo[d]) + 1 + t.cc[d];
  }
  const long long n = (long long)fn[0] * fn[1] * fn[2];
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= n)
    return;
  const long long comp = blockIdx.y;
  const int fi = flo[0] + (int)(gid % fn[0]);
  const int fj = flo[1] + (int)((gid / fn[0]) % fn[1]);
  const int fk = flo[2] + (int)(gid / ((long long)fn[0] * fn[1]));
  const int f[3] = {fi, fj, fk};
  int ci[3];
  for (int d = 0; d < 3; ++d) {
    if (!t.cc[d] && (f[d] & 1))
      return;  // odd VC points are not read by injection
    ci[d] = f[d] >> 1;
  }
  const int ncc = t.cc[0] + t.cc[1] + t.cc[2];
  DFab fb = t.fine;
  fb.p += comp * t.fcs;
  DFab cb = t.crse;
  cb.p += comp * t.ccs;
  const double w = 1.0 / (double)(1 << ncc);
  fb.p[didx(fb, fi, fj, fk)] += w * cb.p[didx(cb, ci[0], ci[1], ci[2])];
}

} // namespace cxp
