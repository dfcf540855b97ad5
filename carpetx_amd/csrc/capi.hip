// C ABI of libcxp.so (declarations and reference citations in include/cxp.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cctype>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <queue>
#include <vector>

#include "../../include/cxp.h"
#include "batched.hpp"
#include "ops.hpp"
#include "adjoint.hpp"
#include "eno_launch.hpp"

using namespace cxp;

// Scratch of cxp_prolongate_adjoint_noalloc: the branch codes of the
// recomputing adjoint in a file-scope device array, one copy per device,
// held until the process ends.  In host code `scratch` names a host-side
// stand-in, never a device address: the address comes from
// hipGetSymbolAddress (scratch_address below), and only that address is
// passed to kernels.
static __device__ __align__(256) unsigned char scratch[CXP_NOALLOC_SCRATCH_BYTES];

struct cxp_op_s {
  OpDesc d;
};

namespace {

// ---------------------------------------------------------------------------
// Operator handles: one immutable entry per (family, order, indextype).

std::once_flag g_ops_once;
cxp_op_s g_ops[9][8][8];
bool g_ops_valid[9][8][8];
// the same operators with the fallback set by the caller: [fb]
cxp_op_s g_ops_fb[9][8][8][2];

void init_ops() {
  for (int f = 0; f < 9; ++f)
    for (int o = 0; o < 8; ++o)
      for (int c = 0; c < 8; ++c) {
        const int it[3] = {(c >> 2) & 1, (c >> 1) & 1, c & 1};
        OpDesc d;
        g_ops_valid[f][o][c] = op_lookup(kFamilies[f], o, it, &d) == 0;
        if (g_ops_valid[f][o][c]) {
          g_ops[f][o][c].d = d;
          for (int fb = 0; fb < 2; ++fb) {
            g_ops_fb[f][o][c][fb].d = d;
            g_ops_fb[f][o][c][fb].d.fb = fb;
          }
        }
      }
}

// Value of key in a Cactus tags string (key=value pairs separated by
// blanks; a value is a "quoted string", a {braced list} or a bare word;
// *quoted tells the first kind).  Keys compare case-insensitively.  Returns
// 1 if found, 0 if absent, -1 if the string does not parse.
int tags_value(const char *tags, const char *key, std::string &out,
               bool *quoted) {
  const char *p = tags;
  auto blank = [](char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r'; };
  while (true) {
    while (*p && blank(*p))
      ++p;
    if (!*p)
      return 0;
    const char *k0 = p;
    while (*p && (std::isalnum((unsigned char)*p) || *p == '_'))
      ++p;
    if (p == k0)
      return -1;
    const std::string k(k0, p);
    while (*p && blank(*p))
      ++p;
    if (*p != '=')
      return -1;
    ++p;
    while (*p && blank(*p))
      ++p;
    std::string v;
    const bool q = *p == '"';
    if (*p == '"') {
      const char *v0 = ++p;
      while (*p && *p != '"')
        ++p;
      if (!*p)
        return -1;
      v.assign(v0, p);
      ++p;
    } else if (*p == '{') {
      int depth = 0;
      const char *v0 = p;
      do {
        if (*p == '{')
          ++depth;
        else if (*p == '}')
          --depth;
        ++p;
      } while (*p && depth > 0);
      if (depth)
        return -1;
      v.assign(v0, p);
    } else {
      const char *v0 = p;
      while (*p && !blank(*p))
        ++p;
      if (p == v0)
        return -1;
      v.assign(v0, p);
    }
    if (k.size() == std::strlen(key) &&
        std::equal(k.begin(), k.end(), key, [](char a, char b) {
          return std::tolower((unsigned char)a) == std::tolower((unsigned char)b);
        })) {
      out = v;
      *quoted = q;
      return 1;
    }
  }
}

// ---------------------------------------------------------------------------
// Validation (the reference's asserts, as status codes)

bool fab_ok(const cxp_fab *f) {
  if (!f || !f->data || f->ncomp < 1 || f->comp_stride < 0)
    return false;
  for (int d = 0; d < 3; ++d)
    if (f->hi[d] < f->lo[d])
      return false;
  const int64_t vol = int64_t(f->hi[0] - f->lo[0] + 1) *
                      (f->hi[1] - f->lo[1] + 1) * (f->hi[2] - f->lo[2] + 1);
  return f->ncomp == 1 || f->comp_stride >= vol;
}

// Bytes of one component of a FAB.  The separable kernel addresses a
// component with 32-bit buffer offsets, so it must stay below 2 GiB.
int64_t fab_component_bytes(const cxp_fab *f) {
  return 8 * int64_t(f->hi[0] - f->lo[0] + 1) * (f->hi[1] - f->lo[1] + 1) *
         (f->hi[2] - f->lo[2] + 1);
}
constexpr int64_t kMaxComponentBytes = (int64_t(1) << 31) - 4096;

bool box_in(const int lo[3], const int hi[3], const int blo[3],
            const int bhi[3]) {
  for (int d = 0; d < 3; ++d)
    if (lo[d] < blo[d] || hi[d] > bhi[d])
      return false;
  return true;
}

DFab dfab(const cxp_fab *f, int comp) {
  DFab v;
  v.p = f->data + (int64_t)comp * f->comp_stride;
  for (int d = 0; d < 3; ++d)
    v.lo[d] = f->lo[d];
  v.sy = f->hi[0] - f->lo[0] + 1;
  v.sz = (long long)v.sy * (f->hi[1] - f->lo[1] + 1);
  return v;
}

int check_prolong(const cxp_op_s *op, const cxp_fab *crse, int crse_comp,
                  const cxp_fab *fine, int fine_comp, int ncomp,
                  const int rlo[3], const int rhi[3], int per_group) {
  if (!op || !rlo || !rhi || !fab_ok(crse) || !fab_ok(fine) || ncomp < 1)
    return CXP_ERR_ARG;
  if (crse_comp < 0 || fine_comp < 0 || crse_comp + ncomp > crse->ncomp ||
      fine_comp + ncomp > fine->ncomp)
    return CXP_ERR_COMP;
  for (int d = 0; d < 3; ++d)
    if (rhi[d] < rlo[d])
      return CXP_ERR_ARG;
  if (!box_in(rlo, rhi, fine->lo, fine->hi))
    return CXP_ERR_REGION;
  int clo[3], chi[3];
  op_coarse_box(op->d, rlo, rhi, clo, chi);
  if (!box_in(clo, chi, crse->lo, crse->hi))
    return CXP_ERR_COARSE;
  if (per_group && ncomp > 16)
    return CXP_ERR_NCOMP;
  if (fab_component_bytes(crse) > kMaxComponentBytes ||
      fab_component_bytes(fine) > kMaxComponentBytes)
    return CXP_ERR_ARG;
  return CXP_OK;
}

int hip_status(hipError_t e) { return e == hipSuccess ? CXP_OK : CXP_ERR_HIP; }

// ---------------------------------------------------------------------------
// Tile-shape selection

struct TileCfg {
  int tx, ty, lx, rp;
};
// Tile shapes: TX x TY fine points per plane, TX * TY = 256 RP (2 RP fine y
// per thread).  Every one fits the kernel's LDS budget (checked at first
// use).
const TileCfg kCfgs[] = {
    {64, 4, 6, 1},  {32, 8, 5, 1},  {16, 16, 4, 1}, {8, 32, 3, 1},
    {4, 64, 2, 1},  {64, 8, 6, 2},  {32, 16, 5, 2}, {16, 32, 4, 2},
    {8, 64, 3, 2},
};
constexpr int kNumTileConfigs = sizeof(kCfgs) / sizeof(kCfgs[0]);

// A list tile's shape code (SepTile::cfg): log2 TX, plus 16 for RP = 2.
int cfg_code(int cfg) { return kCfgs[cfg].lx + (kCfgs[cfg].rp == 2 ? 16 : 0); }

bool cfgs_checked() {
  for (int c = 0; c < kNumTileConfigs; ++c) {
    const TileCfg &k = kCfgs[c];
    const Geom g = make_geom(k.lx, k.rp);
    const int max_coarse = (k.tx / 2 + 9) * (k.ty / 2 + 9);
    if ((k.rp != 1 && k.rp != 2) || geom_lds_doubles(g) > kLdsDoubles ||
        g.tx * g.ty != 2 * k.rp * CXP_NT ||
        max_coarse > sep_pre(k.rp) * CXP_NT ||
        g.tx != k.tx || g.ty != k.ty || (k.tx / 2 + 9) > g.cxs ||
        (k.ty / 2 + 9) > g.cys || k.tx < 4 || (k.rp == 2 && k.tx < 8))
      return false;
  }
  return true;
}
const bool kCfgsOk = cfgs_checked();

// Size class of a patch: thin direction (extent <= 4) and a coarse x-size
// bucket, 0 .. 15; a bulky patch (every extent above 64) is class 16 of its
// own, so tuning the boxes of a level (38^3 in a regrid, say) does not
// retune the launches of a 192^3 box.  The tuner keys on (sids, class).
int size_class(const int rlo[3], const int rhi[3]) {
  const int n[3] = {rhi[0] - rlo[0] + 1, rhi[1] - rlo[1] + 1,
                    rhi[2] - rlo[2] + 1};
  if (n[0] > 64 && n[1] > 64 && n[2] > 64)
    return 16;
  int thin = 0;
  if (n[0] <= 4)
    thin = 1;
  else if (n[1] <= 4)
    thin = 2;
  else if (n[2] <= 4)
    thin = 3;
  int xb = n[0] <= 8 ? 0 : n[0] <= 16 ? 1 : n[0] <= 32 ? 2 : 3;
  return thin * 4 + xb;
}

bool cfg_fits(int cfg, const SepTask &t) {
  if (cfg < 0 || cfg >= kNumTileConfigs)
    return false;
  for (int d = 0; d < 3; ++d)
    if (t.whi[d] - t.wlo[d] + 1 > CXP_ZW)
      return false;
  return true;
}

// Untuned tile shape: the one covering the patch's (x, y) extent with the
// fewest fine points (idle lanes), the wider one on ties (longer coalesced
// rows).  Tiles start at even x and y.
int default_cfg(const SepTask &t) {
  const int nx = t.rhi[0] - (t.rlo[0] & ~1) + 1;
  const int ny = t.rhi[1] - (t.rlo[1] & ~1) + 1;
  int best = 0;
  long long best_area = -1;
  for (int c = 0; c < kNumTileConfigs; ++c) {
    const long long area = (long long)((nx + kCfgs[c].tx - 1) / kCfgs[c].tx) *
                           kCfgs[c].tx *
                           ((ny + kCfgs[c].ty - 1) / kCfgs[c].ty) * kCfgs[c].ty;
    if (best_area < 0 || area < best_area) {
      best = c;
      best_area = area;
    }
  }
  return best;
}

using TuneKey = std::tuple<int, int, int, int>;
std::mutex g_tune_mu;
std::map<TuneKey, std::pair<int, int>> g_tune;  // -> (cfg, zchunk)

TuneKey tune_key(const SepTask &t) {
  return TuneKey(t.sid[0], t.sid[1], t.sid[2], size_class(t.rlo, t.rhi));
}

void choose_cfg(const SepTask &t, int &cfg, int &zchunk) {
  cfg = default_cfg(t);
  zchunk = 0;  // not tuned: chosen per launch / list
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(tune_key(t));
  if (it != g_tune.end() && cfg_fits(it->second.first, t)) {
    cfg = it->second.first;
    zchunk = it->second.second;
  }
}

SepTask make_sep_task(const OpDesc &op, const cxp_fab *crse, int crse_comp,
                      const cxp_fab *fine, int fine_comp, const int rlo[3],
                      const int rhi[3]) {
  SepTask t;
  t.crse = dfab(crse, crse_comp);
  t.fine = dfab(fine, fine_comp);
  for (int d = 0; d < 3; ++d) {
    t.rlo[d] = rlo[d];
    t.rhi[d] = rhi[d];
    t.sid[d] = op.sid[d];
    t.wlo[d] = sid_wlo(op.sid[d]);
    t.whi[d] = sid_whi(op.sid[d]);
  }
  t.ccs = crse->comp_stride;
  t.fcs = fine->comp_stride;
  t.cbytes = (int)fab_component_bytes(crse);
  t.fbytes = (int)fab_component_bytes(fine);
  return t;
}

PointTask make_point_task(const OpDesc &op, const cxp_fab *crse, int crse_comp,
                          const cxp_fab *fine, int fine_comp, int ncomp,
                          const int rlo[3], const int rhi[3]) {
  PointTask t;
  t.crse = dfab(crse, crse_comp);
  t.fine = dfab(fine, fine_comp);
  for (int d = 0; d < 3; ++d) {
    t.rlo[d] = rlo[d];
    t.rhi[d] = rhi[d];
    t.sid[d] = op.sid[d];
  }
  t.fb_per_var = op_fallback_per_var(op);
  t.fb_per_group = op_fallback_per_group(op);
  t.ncomp = ncomp;
  t.ccs = crse->comp_stride;
  t.fcs = fine->comp_stride;
  t.d2 = nullptr;
  return t;
}

int launch_point(const PointTask &t, int ncomp, int per_group, hipStream_t s,
                 unsigned char *codes = nullptr) {
  const long long n = (long long)(t.rhi[0] - t.rlo[0] + 1) *
                      (t.rhi[1] - t.rlo[1] + 1) * (t.rhi[2] - t.rlo[2] + 1);
  // LDS-tiled kernels when a tile's coarse block fits in 48 KB (always in
  // per-var mode; per-group mode with few components), else one thread per
  // point reading global memory
  const size_t lds =
      8 * (size_t)pt_tile_block(t.sid) * (size_t)((per_group ? ncomp : 1) + 3);
  if (lds <= 48 * 1024) {
    int nt[3];
    for (int d = 0; d < 3; ++d)
      nt[d] = (t.rhi[d] - (t.rlo[d] & ~1) + CXP_PT_TILE) / CXP_PT_TILE;
    const dim3 grid((unsigned)(nt[0] * nt[1] * nt[2]), per_group ? 1 : ncomp);
    if (per_group)
      hipLaunchKernelGGL(k_point_tiled<true>, grid, dim3(256), lds, s, t,
                         nt[0], nt[1], codes);
    else
      hipLaunchKernelGGL(k_point_tiled<false>, grid, dim3(256), lds, s, t,
                         nt[0], nt[1], codes);
    return hip_status(hipGetLastError());
  }
  const unsigned nb = (unsigned)((n + 255) / 256);
  if (per_group)
    hipLaunchKernelGGL(k_point_per_group, dim3(nb), dim3(256), 0, s, t, codes);
  else
    hipLaunchKernelGGL(k_point_per_var, dim3(nb, ncomp), dim3(256), 0, s, t,
                       codes);
  return hip_status(hipGetLastError());
}

int launch_thin_box(const SepTask &t, int ncomp, hipStream_t s) {
  const int nx = t.rhi[0] - (t.rlo[0] & ~1) + 1;
  const int ny = t.rhi[1] - (t.rlo[1] & ~1) + 1;
  const int ntx = (nx + 15) / 16, nty = (ny + 15) / 16;
  const size_t lds = 8 * (size_t)thin_lds_doubles(t.whi[1] - t.wlo[1],
                                                  t.whi[2] - t.wlo[2]);
  hipLaunchKernelGGL(k_thin_box, dim3(ntx * nty, ncomp), dim3(CXP_NT), lds, s,
                     t, ntx);
  return hip_status(hipGetLastError());
}

int launch_sep_box(const SepTask &t, int ncomp, hipStream_t s) {
  if (thin_eligible(t))
    return launch_thin_box(t, ncomp, s);
  int cfg, zc;
  choose_cfg(t, cfg, zc);
  const TileCfg &c = kCfgs[cfg];
  const int nx = t.rhi[0] - (t.rlo[0] & ~1) + 1;
  const int ny = t.rhi[1] - (t.rlo[1] & ~1) + 1;
  const int nz = t.rhi[2] - t.rlo[2] + 1;
  const int ntx = (nx + c.tx - 1) / c.tx;
  const int nty = (ny + c.ty - 1) / c.ty;
  // per-box: shorten z-chunks until the launch has enough workgroups
  int zchunk = zc > 0 ? std::min(nz, zc) : nz;
  while (zc <= 0 && zchunk > 8 &&
         (long long)ntx * nty * ((nz + zchunk - 1) / zchunk) * ncomp < 2048)
    zchunk = (zchunk + 1) / 2;
  const int nzc = (nz + zchunk - 1) / zchunk;
  const dim3 grid(ntx * nty * nzc, ncomp);
  // uniform stencils (VVV / CCC operators): the kernel compiled for them
  if (t.sid[0] == t.sid[1] && t.sid[1] == t.sid[2]) {
    switch (t.sid[0]) {
#define CXP_SB(S)                                                              \
  case S:                                                                      \
    if (c.rp == 2)                                                             \
      hipLaunchKernelGGL((k_sep_box<S, 2>), grid, dim3(CXP_NT), 0, s, t, c.lx, \
                         ntx, nty, zchunk);                                    \
    else                                                                       \
      hipLaunchKernelGGL((k_sep_box<S, 1>), grid, dim3(CXP_NT), 0, s, t, c.lx, \
                         ntx, nty, zchunk);                                    \
    return hip_status(hipGetLastError());
      CXP_LINEAR_SIDS(CXP_SB)
#undef CXP_SB
    }
  }
  if (c.rp == 2)
    hipLaunchKernelGGL((k_sep_box<-1, 2>), grid, dim3(CXP_NT), 0, s, t, c.lx,
                       ntx, nty, zchunk);
  else
    hipLaunchKernelGGL((k_sep_box<-1, 1>), grid, dim3(CXP_NT), 0, s, t, c.lx,
                       ntx, nty, zchunk);
  return hip_status(hipGetLastError());
}

int prolongate_impl(const cxp_op_s *op, const cxp_fab *crse, int crse_comp,
                    const cxp_fab *fine, int fine_comp, int ncomp,
                    const int rlo[3], const int rhi[3], int per_group,
                    hipStream_t s) {
  if (op_is_separable(op->d, per_group != 0)) {
    // per-var and per-group coincide for operators without shifts/fallback
    const SepTask t =
        make_sep_task(op->d, crse, crse_comp, fine, fine_comp, rlo, rhi);
    return launch_sep_box(t, ncomp, s);
  }
  if (eno_ccc_fits(op->d, per_group))
    return eno_ccc_launch(crse, crse_comp, fine, fine_comp, ncomp, rlo, rhi,
                          nullptr, false, s);
  const PointTask t = make_point_task(op->d, crse, crse_comp, fine, fine_comp,
                                      ncomp, rlo, rhi);
  return launch_point(t, ncomp, per_group, s);
}

// The gather of an ENO or linear-fallback adjoint: regions, branch
// stencils and the coarse points any branch reaches.  crse_in only sets the
// point task's coarse geometry here.
AdjPointTask make_adj_point_task(const cxp_op_s *op, const cxp_fab *crse_in,
                                 const cxp_fab *crse_bar, int crse_comp,
                                 const cxp_fab *fine_bar, int fine_comp,
                                 int ncomp, const int rlo[3], const int rhi[3],
                                 int per_group) {
  const OpDesc &od = op->d;
  AdjPointTask t;
  t.pt = make_point_task(od, crse_in, crse_comp, fine_bar, fine_comp, ncomp,
                         rlo, rhi);
  t.cbar = dfab(crse_bar, crse_comp);
  t.fbar = dfab(fine_bar, fine_comp);
  t.ccs = crse_bar->comp_stride;
  t.fcs = fine_bar->comp_stride;
  t.per_group = per_group ? 1 : 0;
  int clo[3], chi[3];
  op_coarse_box(od, rlo, rhi, clo, chi);
  for (int d = 0; d < 3; ++d) {
    const int sid = od.sid[d];
    t.sl[d] = per_group ? sid_linear_per_group(sid) : sid_linear_per_var(sid);
    int lo = 1 << 20, hi = -(1 << 20);
    const int ms = sid == SID_ENO2 ? 1 : 0;
    for (int off = 0; off < 2; ++off) {
      for (int sh = -ms; sh <= ms; ++sh) {
        int a, b;
        stencil_radius(sid, sh, off, a, b);
        lo = std::min(lo, a);
        hi = std::max(hi, b);
      }
      int a, b;
      stencil_radius(t.sl[d], 0, off, a, b);
      lo = std::min(lo, a);
      hi = std::max(hi, b);
    }
    t.rlo_reach[d] = lo;
    t.rhi_reach[d] = hi;
    t.olo[d] = std::max(clo[d], (rlo[d] >> 1) + lo);
    t.ohi[d] = std::min(chi[d], (rhi[d] >> 1) + hi);
  }
  return t;
}

// crse_bar += P^T fine_bar over the recorded branches: one thread per coarse
// output point and component, coarse tiles with their codes in LDS when the
// window fits.
int launch_adj_gather(const AdjPointTask &t, const unsigned char *codes,
                      int ncomp, hipStream_t s) {
  long long nout = 1;
  int ntl[3], wmax = 1;
  const int tsz[3] = {CXP_ADJ_TX, CXP_ADJ_TY, CXP_ADJ_TZ};
  for (int d = 0; d < 3; ++d) {
    nout *= t.ohi[d] - t.olo[d] + 1;
    ntl[d] = (t.ohi[d] - t.olo[d] + tsz[d]) / tsz[d];
    wmax *= 2 * (tsz[d] + t.rhi_reach[d] - t.rlo_reach[d]);
  }
  if (nout <= 0)
    return CXP_OK;
  const size_t glds =
      8 * (size_t)(CXP_ADJ_NOCODE + 1) * CXP_ADJ_TROW + (size_t)wmax;
  if (glds <= 64 * 1024)
    hipLaunchKernelGGL(k_adj_gather_tiled,
                       dim3((unsigned)(ntl[0] * ntl[1] * ntl[2]), ncomp),
                       dim3(256), glds, s, t, codes, ntl[0], ntl[1]);
  else
    hipLaunchKernelGGL(k_adj_gather,
                       dim3((unsigned)((nout + 255) / 256), ncomp), dim3(256),
                       0, s, t, codes);
  return hip_status(hipGetLastError());
}

// Bytes of branch codes the recomputing ENO / fallback adjoint needs: one
// per fine point, per component in per-var mode.
size_t adjoint_code_bytes(int ncomp, const int rlo[3], const int rhi[3],
                          int per_group) {
  size_t npts = 1;
  for (int d = 0; d < 3; ++d)
    npts *= (size_t)(rhi[d] - rlo[d] + 1);
  return npts * (size_t)(per_group ? 1 : ncomp);
}

// Whether s is capturing a graph (an invalidated capture is an error).
int stream_capturing(hipStream_t s, bool &capturing,
                     unsigned long long *id = nullptr) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  unsigned long long cid = 0;
  if (hipStreamGetCaptureInfo(s, &cs, &cid) != hipSuccess)
    return CXP_ERR_HIP;
  if (cs == hipStreamCaptureStatusInvalidated)
    return CXP_ERR_STATE;
  capturing = cs == hipStreamCaptureStatusActive;
  if (id)
    *id = cid;
  return CXP_OK;
}

// Branch-code buffers of the recomputing adjoint: one per (device, stream),
// grown on demand, stream-ordered (allocated and freed on that stream), and
// reused by later calls on the same stream, which stream order keeps apart.
struct CodePool {
  unsigned char *p = nullptr;
  size_t cap = 0;
};
std::mutex g_pool_mu;
std::map<std::pair<int, hipStream_t>, CodePool> g_code_pool;

int pooled_codes(hipStream_t s, size_t nbytes, unsigned char **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess)
    return CXP_ERR_HIP;
  std::lock_guard<std::mutex> lk(g_pool_mu);
  CodePool &cp = g_code_pool[{dev, s}];
  if (nbytes > cp.cap) {
    if (cp.p && hipFreeAsync(cp.p, s) != hipSuccess)
      return CXP_ERR_HIP;
    cp.p = nullptr;
    cp.cap = 0;
    const size_t cap = std::max(nbytes + nbytes / 4, (size_t)1 << 20);
    if (hipMallocAsync((void **)&cp.p, cap, s) != hipSuccess) {
      cp.p = nullptr;
      return CXP_ERR_HIP;
    }
    cp.cap = cap;
  }
  *out = cp.p;
  return CXP_OK;
}

// The address of the scratch array (top of this file) on the current
// device, looked up once per device with that device current.
std::mutex g_scratch_mu;
std::map<int, unsigned char *> g_scratch_ptr;  // device -> address of scratch

int scratch_address(unsigned char **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess)
    return CXP_ERR_HIP;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it == g_scratch_ptr.end()) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
      return CXP_ERR_HIP;
    it = g_scratch_ptr.emplace(dev, (unsigned char *)p).first;
  }
  *out = it->second;
  return CXP_OK;
}

// Adjoint of an ENO or linear-fallback operator, linearised at crse_in:
// record every fine point's branch into codes, then gather per coarse
// point.  The caller has checked the arguments and sized codes.
int adjoint_point_run(const cxp_op_s *op, const cxp_fab *crse_in,
                      const cxp_fab *crse_bar, int crse_comp,
                      const cxp_fab *fine_bar, int fine_comp, int ncomp,
                      const int rlo[3], const int rhi[3], int per_group,
                      unsigned char *codes, hipStream_t s) {
  if (eno_ccc_fits(op->d, per_group)) {
    // one code per coarse parent: an eighth of the bytes of the per-point
    // record, written and read contiguously
    if (const int e = eno_ccc_launch(crse_in, crse_comp, nullptr, 0, ncomp,
                                     rlo, rhi, codes, true, s, true))
      return e;
    return eno_ccc_adj_launch(crse_bar, crse_comp, fine_bar, fine_comp, ncomp,
                              rlo, rhi, codes, s, true);
  }
  const AdjPointTask t =
      make_adj_point_task(op, crse_in, crse_bar, crse_comp, fine_bar,
                          fine_comp, ncomp, rlo, rhi, per_group);
  long long npts = 1;
  for (int d = 0; d < 3; ++d)
    npts *= rhi[d] - rlo[d] + 1;
  // the branch record: LDS-tiled like the forward when a tile's coarse
  // block fits, else one thread per point
  const size_t lds =
      8 * (size_t)pt_tile_block(t.pt.sid) * (size_t)((per_group ? ncomp : 1) + 3);
  if (lds <= 48 * 1024) {
    int nt[3];
    for (int d = 0; d < 3; ++d)
      nt[d] = (rhi[d] - (rlo[d] & ~1) + CXP_PT_TILE) / CXP_PT_TILE;
    const dim3 grid((unsigned)(nt[0] * nt[1] * nt[2]), per_group ? 1 : ncomp);
    if (per_group)
      hipLaunchKernelGGL((k_point_tiled<true, true>), grid, dim3(256), lds, s,
                         t.pt, nt[0], nt[1], codes);
    else
      hipLaunchKernelGGL((k_point_tiled<false, true>), grid, dim3(256), lds, s,
                         t.pt, nt[0], nt[1], codes);
  } else {
    hipLaunchKernelGGL(k_adj_decide,
                       dim3((unsigned)((npts + 255) / 256), per_group ? 1 : ncomp),
                       dim3(256), 0, s, t.pt, codes, t.per_group);
  }
  int st = hip_status(hipGetLastError());
  if (st == CXP_OK)
    st = launch_adj_gather(t, codes, ncomp, s);
  return st;
}

int adjoint_point_check(const cxp_op_s *op, const cxp_fab *crse_in,
                        int crse_comp, const cxp_fab *fine_bar, int fine_comp,
                        int ncomp, const int rlo[3], const int rhi[3],
                        int per_group) {
  if (!crse_in)
    return CXP_ERR_ARG;
  return check_prolong(op, crse_in, crse_comp, fine_bar, fine_comp, ncomp,
                       rlo, rhi, per_group);
}

// The recomputing adjoint with its codes from the stream's pool.  A
// captured graph would keep the pool's buffer, which a later call may free
// or share, so it refuses capture (CXP_ERR_STATE).
int adjoint_point_launch(const cxp_op_s *op, const cxp_fab *crse_in,
                         const cxp_fab *crse_bar, int crse_comp,
                         const cxp_fab *fine_bar, int fine_comp, int ncomp,
                         const int rlo[3], const int rhi[3], int per_group,
                         hipStream_t s) {
  if (const int e = adjoint_point_check(op, crse_in, crse_comp, fine_bar,
                                        fine_comp, ncomp, rlo, rhi, per_group))
    return e;
  bool capturing = false;
  if (const int e = stream_capturing(s, capturing))
    return e;
  if (capturing)
    return CXP_ERR_STATE;
  unsigned char *codes = nullptr;
  if (const int e = pooled_codes(
          s, adjoint_code_bytes(ncomp, rlo, rhi, per_group), &codes))
    return e;
  return adjoint_point_run(op, crse_in, crse_bar, crse_comp, fine_bar,
                           fine_comp, ncomp, rlo, rhi, per_group, codes, s);
}

// The MINMOD adjoint: for CCC operators one fused launch that recomputes
// the branches; else record the branch codes from crse_in, then gather,
// with the codes from the stream's pool (or from `codes`, when given, of at
// least mm_code_bytes).  Refuses capture like adjoint_point_launch when it
// takes the pool.
int minmod_adjoint_launch(const cxp_op_s *op, const cxp_fab *crse_in,
                          const cxp_fab *crse_bar, int crse_comp,
                          const cxp_fab *fine_bar, int fine_comp, int ncomp,
                          const int rlo[3], const int rhi[3], int per_group,
                          hipStream_t s, unsigned char *codes = nullptr) {
  if (const int e = adjoint_point_check(op, crse_in, crse_comp, fine_bar,
                                        fine_comp, ncomp, rlo, rhi, per_group))
    return e;
  if (op_all_minmod(op->d) && mm_tiled_fits(fine_bar))
    return mm_fused_launch(op->d, crse_in, crse_bar, crse_comp, fine_bar,
                           fine_comp, ncomp, rlo, rhi, s);
  const long long nbytes = mm_code_bytes(op->d, rlo, rhi, ncomp);
  if (!codes && nbytes > 0) {
    bool capturing = false;
    if (const int e = stream_capturing(s, capturing))
      return e;
    if (capturing)
      return CXP_ERR_STATE;
    if (const int e = pooled_codes(s, (size_t)nbytes, &codes))
      return e;
  }
  if (const int e = mm_record_launch(op->d, crse_in, crse_comp, ncomp, rlo,
                                     rhi, codes, s))
    return e;
  return mm_gather_launch(op->d, crse_bar, crse_comp, fine_bar, fine_comp,
                          ncomp, rlo, rhi, codes, s);
}

} // namespace

// ===========================================================================

struct cxp_worklist_s {
  struct Patch {
    const cxp_op_s *op;
    cxp_fab crse, fine;
    int crse_comp, fine_comp, ncomp;
    int rlo[3], rhi[3];
    // >= 0: crse is a gathered coarse patch at this offset (doubles) of the
    // patch buffer, set at upload
    int64_t gather_off = -1;
  };
  std::vector<Patch> patches;
  std::vector<SepTask> tasks;
  std::vector<SepTile> tiles;
  std::vector<int> point_patches;  // indices of non-separable patches
  SepTask *d_tasks = nullptr;
  SepTile *d_tiles = nullptr;
  std::vector<int> wg_start;  // per-workgroup tile ranges of k_fill_batched
  int *d_wg_start = nullptr;
  size_t cap_wg = 0;
  std::vector<int> qstart;  // per-XCD-class queues after the ranges, or empty
  int *d_ctr = nullptr;     // their claim counters and the finish counter
  size_t cap_tasks = 0, cap_tiles = 0;
  // coarse patches of level fills, gathered at the start of every run
  std::vector<GatherRun> gather_runs;  // dst: offsets into d_gather
  std::vector<GatherChunk> gather_chunks;
  int64_t gather_doubles = 0;
  double *d_gather = nullptr;
  size_t cap_gather = 0;
  GatherRun *d_gather_runs = nullptr;
  GatherChunk *d_gather_chunks = nullptr;
  size_t cap_gather_runs = 0, cap_gather_chunks = 0;
  // boundary values of gathered coarse patches, after the gather
  std::vector<BcRun> bc_runs;  // off: offsets into d_gather
  std::vector<BcChunk> bc_chunks;   // sources in the patch: after the gather
  std::vector<BcChunk> bcd_chunks;  // sources in one coarse FAB: with it
  std::vector<BcDirect> bc_direct;
  BcRun *d_bc_runs = nullptr;
  BcChunk *d_bc_chunks = nullptr, *d_bcd_chunks = nullptr;
  BcDirect *d_bc_direct = nullptr;
  size_t cap_bc_runs = 0, cap_bc_chunks = 0, cap_bcd_chunks = 0,
         cap_bc_direct = 0;
  // boundary conditions of levels, in place, after the fills
  std::vector<BcRun> post_runs;
  std::vector<BcChunk> post_chunks;
  BcRun *d_post_runs = nullptr;
  BcChunk *d_post_chunks = nullptr;
  size_t cap_post_runs = 0, cap_post_chunks = 0;
  std::vector<PoisonRun> poison_runs;  // one per (patch, component)
  std::vector<PoisonChunk> poison_chunks;
  PoisonRun *d_poison_runs = nullptr;
  PoisonChunk *d_poison_chunks = nullptr;
  size_t cap_poison_runs = 0, cap_poison_chunks = 0;
  bool uploaded = false;
  // the reference's prolongate_per_group parameter, for ENO and fallback
  // patches
  int per_group = 0;
  int grid = 0;
  int lds = 0;  // bytes of dynamic LDS of k_fill_batched
  int uniform_sid = -1;  // the one stencil of every separable task, or -1
  int64_t fine_points = 0, algo_bytes = 0;
  // Identity and version for captured graphs: every change that may
  // reallocate the device buffers or change the tiles (append, clear,
  // per-group setting, upload, autotune) bumps epoch, and a graph that
  // launched the list at another epoch refuses to replay.
  uint64_t id = 0;
  uint64_t epoch = 0;
  // Launch ordering: one list's claim counters and coarse-patch buffer serve
  // one launch at a time, so a launch on another stream than the previous
  // one first waits for what that stream holds so far (an event recorded
  // there at that moment; none is recorded on the same-stream path, where
  // each event between kernels would cost microseconds of GPU time).
  std::mutex mu;
  hipEvent_t done = nullptr;
  hipStream_t last_stream = nullptr;
  bool launched = false;
};

struct cxp_graph_s {
  hipGraph_t graph;
  hipGraphExec_t exec;
  // (list id, epoch) of every list the captured work launched
  std::vector<std::pair<uint64_t, uint64_t>> lists;
};

namespace {

// ---- list identity, capture records and launch ordering

std::mutex g_lists_mu;  // guards the three maps and every list's epoch
std::map<uint64_t, cxp_worklist_s *> g_lists;  // live lists by id
uint64_t g_next_list_id = 1;
// lists launched on a capturing stream, by capture id, until the capture
// ends in cxp_graph_capture_end (captures ended elsewhere are dropped
// oldest first beyond kMaxCaptureRecords)
std::map<unsigned long long, std::vector<std::pair<uint64_t, uint64_t>>>
    g_capture_lists;
constexpr size_t kMaxCaptureRecords = 256;

void touch(cxp_worklist_s *wl) {
  std::lock_guard<std::mutex> lk(g_lists_mu);
  ++wl->epoch;
}

// Waits on the host for the list's last launch (before its device buffers
// are rewritten or freed).
int list_drain(cxp_worklist_s *wl) {
  std::lock_guard<std::mutex> lk(wl->mu);
  if (wl->launched && hipStreamSynchronize(wl->last_stream) != hipSuccess)
    return CXP_ERR_HIP;
  return CXP_OK;
}

// Makes s wait for everything issued so far on the list's previous launch
// stream, when that is another stream.  Called with wl->mu held.
int list_order_after_last(cxp_worklist_s *wl, hipStream_t s) {
  if (!wl->launched || wl->last_stream == s)
    return CXP_OK;
  if (!wl->done &&
      hipEventCreateWithFlags(&wl->done, hipEventDisableTiming) != hipSuccess) {
    wl->done = nullptr;
    return CXP_ERR_HIP;
  }
  if (hipEventRecord(wl->done, wl->last_stream) != hipSuccess ||
      hipStreamWaitEvent(s, wl->done, 0) != hipSuccess)
    return CXP_ERR_HIP;
  return CXP_OK;
}

// Before launching wl's work on s: on a capturing stream, records the list
// in the capture; otherwise orders s after the list's previous launch when
// that went to another stream.  Called with wl->mu held.
int list_enter(cxp_worklist_s *wl, hipStream_t s, bool &capturing) {
  unsigned long long cid = 0;
  if (const int e = stream_capturing(s, capturing, &cid))
    return e;
  if (capturing) {
    std::lock_guard<std::mutex> lk(g_lists_mu);
    g_capture_lists[cid].emplace_back(wl->id, wl->epoch);
    while (g_capture_lists.size() > kMaxCaptureRecords)
      g_capture_lists.erase(g_capture_lists.begin());
    return CXP_OK;
  }
  return list_order_after_last(wl, s);
}

// After wl's work went to s (not capturing).  Called with wl->mu held.
int list_leave(cxp_worklist_s *wl, hipStream_t s) {
  wl->last_stream = s;
  wl->launched = true;
  return CXP_OK;
}

// Runs fn (the launches of one entry point) between list_enter and
// list_leave with the list's lock held.
template <typename F> int list_launch(cxp_worklist_s *wl, hipStream_t s, F fn) {
  std::lock_guard<std::mutex> lk(wl->mu);
  bool capturing = false;
  if (const int e = list_enter(wl, s, capturing))
    return e;
  const int st = fn();
  if (capturing)
    return st;
  const int e = list_leave(wl, s);
  return st ? st : e;
}

// Union of the coarse points read for one patch.
int64_t patch_coarse_points(const cxp_worklist_s::Patch &p) {
  int clo[3], chi[3];
  op_coarse_box(p.op->d, p.rlo, p.rhi, clo, chi);
  return int64_t(chi[0] - clo[0] + 1) * (chi[1] - clo[1] + 1) *
         (chi[2] - clo[2] + 1);
}

int persistent_grid(int ntiles, int lds, int su);

// Compute units of the current device.
int device_cus() {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess)
      cus = prop.multiProcessorCount;
    if (cus <= 0)
      cus = 256;
  }
  return cus;
}

// Percent of each XCD class's estimated cost that k_fill_batched hands out
// dynamically (its smallest tiles).  On the uneven 512-patch level, 20 to
// 45 percent all give the same time and 10 percent is slower.
constexpr int kQueuePct = 25;

// Relative time a workgroup of the persistent grid needs for one unit of
// work, by its dispatch slot on its compute unit.  Workgroups are dispatched
// round-robin over the compute units, so block b is the (b / cus)-th
// resident workgroup of its unit; the SIMDs issue the oldest ready wave
// first, and the later slots run slower.  Measured on MI355X with 8 slots
// per unit (the uneven 512-patch level, equal work per workgroup: mean
// busy time by slot relative to slot 0).  Other slot counts keep equal
// weights: unmeasured, and a guessed ramp slowed the 10-slot lists of the
// 64-box level.
double slot_weight(int b, int cus, int grid) {
  static const double w8[8] = {1.000, 1.082, 1.085, 1.100,
                               1.204, 1.265, 1.251, 1.280};
  if (grid != 8 * cus)
    return 1.0;
  return w8[std::min(b / cus, 7)];
}

// Orders the tile list into one contiguous range per workgroup of the
// persistent grid, followed, for lists of at least four tiles per
// workgroup, by one queue per XCD class (see k_fill_batched).  Cost
// estimates, in units of one plane iteration of the z-marching routine: the
// coarse planes a separable tile marches through plus a fixed per-tile
// overhead; a thin tile costs about two plane iterations (measured on the
// uneven level).
void schedule_tiles(cxp_worklist_s *wl, int grid) {
  struct Group {
    size_t first, count;
    long long cost;
  };
  const size_t nt = wl->tiles.size();
  std::vector<long long> cost(nt);
  // A group is every tile that writes one component of one fine FAB: all
  // patches of a FAB (the six slabs of a ghost shell) go to one XCD, in
  // patch order.  Their coarse reads overlap at the slabs' edges, and the
  // low-x ghosts of one row and the high-x ghosts of the previous row share
  // 128-byte lines, whose partial writes merge in that XCD's L2 when both
  // slabs run at about the same time (equal-cost tiles keep this order in
  // the longest-first sort below).
  std::vector<Group> groups;
  std::vector<size_t> member;  // tiles in group order
  {
    std::map<std::pair<const double *, int>, size_t> gid;
    std::vector<std::vector<size_t>> gt;
    for (size_t i = 0; i < nt; ++i) {
      const SepTile &tl = wl->tiles[i];
      const SepTask &t = wl->tasks[tl.task];
      // marching steps, weighted by the points a step covers
      cost[i] = tl.cfg < 0 ? 2
                           : (long long)((tl.z1 >> 1) + t.whi[2] - (tl.z0 >> 1) -
                                         t.wlo[2] + 1) *
                                     (tl.cfg >= 16 ? 2 : 1) +
                                 4;
      const auto key = std::make_pair((const double *)t.fine.p, tl.comp);
      auto it = gid.find(key);
      if (it == gid.end()) {
        it = gid.emplace(key, gt.size()).first;
        gt.emplace_back();
      }
      gt[it->second].push_back(i);
    }
    // A group heavier than a quarter of one XCD's share would leave XCDs
    // idle on lists of few FABs: it is split into its patches, and a patch
    // still that heavy into runs of its tiles.
    long long total = 0;
    for (size_t i = 0; i < nt; ++i)
      total += cost[i];
    const int nxcd = grid >= 8 ? 8 : 1;
    const long long cap = std::max(1LL, total / (4LL * nxcd));
    member.reserve(nt);
    auto emit = [&](const std::vector<size_t> &v) {
      Group g{member.size(), 0, 0};
      for (size_t i : v) {
        if (g.count > 0 && g.cost + cost[i] > cap) {
          groups.push_back(g);
          g = Group{member.size(), 0, 0};
        }
        member.push_back(i);
        ++g.count;
        g.cost += cost[i];
      }
      if (g.count > 0)
        groups.push_back(g);
    };
    for (const auto &v : gt) {
      long long c = 0;
      for (size_t i : v)
        c += cost[i];
      if (c <= cap) {
        Group g{member.size(), v.size(), c};
        for (size_t i : v)
          member.push_back(i);
        groups.push_back(g);
        continue;
      }
      // by patch (task), in order
      std::map<int, std::vector<size_t>> by_task;
      std::vector<int> task_order;
      for (size_t i : v) {
        const int tk = wl->tiles[i].task;
        if (!by_task.count(tk))
          task_order.push_back(tk);
        by_task[tk].push_back(i);
      }
      for (int tk : task_order)
        emit(by_task[tk]);
    }
  }
  const int nx = grid >= 8 ? 8 : 1;  // XCD classes of workgroups
  const int cus = device_cus();
  // groups -> XCD, the most expensive first to the least loaded
  std::vector<size_t> order(groups.size());
  for (size_t i = 0; i < order.size(); ++i)
    order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    return groups[a].cost > groups[b].cost;
  });
  std::vector<std::vector<size_t>> xcd_tiles(nx);
  std::vector<long long> xload(nx, 0);
  for (size_t gi : order) {
    int best = 0;
    for (int x = 1; x < nx; ++x)
      if (xload[x] < xload[best])
        best = x;
    for (size_t k = 0; k < groups[gi].count; ++k)
      xcd_tiles[best].push_back(member[groups[gi].first + k]);
    xload[best] += groups[gi].cost;
  }
  // within an XCD: the smallest tiles, kQueuePct percent of the XCD's cost,
  // form its queue; the others go longest first to the least loaded of its
  // workgroups
  std::vector<std::vector<size_t>> wg_tiles(grid), queue(nx);
  const bool use_queue = nt >= 4 * (size_t)grid;
  for (int x = 0; x < nx; ++x) {
    std::vector<size_t> &ts = xcd_tiles[x];
    std::stable_sort(ts.begin(), ts.end(),
                     [&](size_t a, size_t b) { return cost[a] > cost[b]; });
    if (use_queue) {
      long long total = 0, tail = 0;
      for (size_t i : ts)
        total += cost[i];
      size_t cut = ts.size();
      while (cut > 0 && (tail + cost[ts[cut - 1]]) * 100 <= total * kQueuePct) {
        --cut;
        tail += cost[ts[cut]];
      }
      queue[x].assign(ts.begin() + cut, ts.end());
      ts.resize(cut);
    }
    // (weighted load, workgroup); the weight is the workgroup's time per
    // unit of work
    using Slot = std::pair<double, int>;
    std::priority_queue<Slot, std::vector<Slot>, std::greater<Slot>> heap;
    for (int b = x; b < grid; b += nx)
      heap.push({0.0, b});
    for (size_t i : ts) {
      Slot sl = heap.top();
      heap.pop();
      wg_tiles[sl.second].push_back(i);
      sl.first += cost[i] * slot_weight(sl.second, cus, grid);
      heap.push(sl);
    }
  }
  // per workgroup: interleave separable and thin tiles by cost progress
  std::vector<SepTile> out;
  out.reserve(nt);
  wl->wg_start.assign(grid + 1, 0);
  for (int b = 0; b < grid; ++b) {
    wl->wg_start[b] = (int)out.size();
    std::vector<size_t> sep, thin;
    long long csep = 0, cthin = 0;
    for (size_t i : wg_tiles[b]) {
      if (wl->tiles[i].cfg < 0) {
        thin.push_back(i);
        cthin += cost[i];
      } else {
        sep.push_back(i);
        csep += cost[i];
      }
    }
    size_t a = 0, c = 0;
    long long dsep = 0, dthin = 0;
    while (a < sep.size() || c < thin.size()) {
      // the kind whose done fraction is smaller goes next
      const bool take_sep =
          c >= thin.size() ||
          (a < sep.size() && (double)dsep * cthin <= (double)dthin * csep);
      const size_t i = take_sep ? sep[a++] : thin[c++];
      (take_sep ? dsep : dthin) += cost[i];
      out.push_back(wl->tiles[i]);
    }
  }
  wl->wg_start[grid] = (int)out.size();
  wl->qstart.clear();
  if (use_queue) {
    // each queue longest first; equal costs keep their patch order
    for (int x = 0; x < nx; ++x) {
      wl->qstart.push_back((int)out.size());
      for (size_t i : queue[x])
        out.push_back(wl->tiles[i]);
    }
    wl->qstart.push_back((int)out.size());
  }
  wl->tiles.swap(out);
}

// Tiles of the separable patches.  Untuned patches share one z-chunk length,
// the longest (fewest halo planes) that still gives the persistent grid
// several tiles per workgroup slot.
void build_tiles(cxp_worklist_s *wl, int only_class_key_cfg = -1,
                 const TuneKey *only_key = nullptr) {
  wl->tasks.clear();
  wl->tiles.clear();
  wl->lds = 8 * kLdsDoubles;
  wl->point_patches.clear();
  struct Plan {
    int task, cfg, zc, ncomp;
  };
  std::vector<Plan> plans;
  for (size_t n = 0; n < wl->patches.size(); ++n) {
    const auto &p = wl->patches[n];
    if (!op_is_separable(p.op->d, wl->per_group != 0)) {
      wl->point_patches.push_back((int)n);
      continue;
    }
    SepTask t = make_sep_task(p.op->d, &p.crse, p.crse_comp, &p.fine,
                              p.fine_comp, p.rlo, p.rhi);
    if (thin_eligible(t)) {
      // whole-z 16 x 16 tiles; no tile shape to tune
      if (only_key)
        continue;
      const int task = (int)wl->tasks.size();
      wl->tasks.push_back(t);
      const int lds = 8 * thin_lds_doubles(t.whi[1] - t.wlo[1],
                                           t.whi[2] - t.wlo[2]);
      wl->lds = std::max(wl->lds, lds);
      for (int comp = 0; comp < p.ncomp; ++comp)
        for (int y0 = t.rlo[1] & ~1; y0 <= t.rhi[1]; y0 += 16)
          for (int x0 = t.rlo[0] & ~1; x0 <= t.rhi[0]; x0 += 16) {
            SepTile tl;
            tl.task = task;
            tl.comp = comp;
            tl.x0 = x0;
            tl.y0 = y0;
            tl.z0 = t.rlo[2];
            tl.z1 = t.rhi[2];
            tl.cfg = -1;
            tl.pad = 0;
            wl->tiles.push_back(tl);
          }
      continue;
    }
    if (only_key && tune_key(t) != *only_key)
      continue;
    int cfg, zc;
    choose_cfg(t, cfg, zc);
    if (only_key) {
      cfg = only_class_key_cfg >> 8;
      zc = only_class_key_cfg & 0xff;
    }
    plans.push_back({(int)wl->tasks.size(), cfg, zc, p.ncomp});
    wl->tasks.push_back(t);
  }
  auto count = [&](int zdef) {
    long long n = 0;
    for (const Plan &pl : plans) {
      const SepTask &t = wl->tasks[pl.task];
      const TileCfg &c = kCfgs[pl.cfg];
      const int zc = pl.zc > 0 ? pl.zc : zdef;
      const long long nx = (t.rhi[0] - (t.rlo[0] & ~1) + c.tx) / c.tx;
      const long long ny = (t.rhi[1] - (t.rlo[1] & ~1) + c.ty) / c.ty;
      const long long nz = (t.rhi[2] - t.rlo[2] + zc) / zc;
      n += nx * ny * nz * pl.ncomp;
    }
    return n;
  };
  // one stencil in every direction of every task: the specialised kernel
  wl->uniform_sid = -1;
  for (size_t i = 0; i < wl->tasks.size(); ++i) {
    const SepTask &t = wl->tasks[i];
    const bool uni = t.sid[0] == t.sid[1] && t.sid[1] == t.sid[2];
    if (!uni || (i > 0 && t.sid[0] != wl->uniform_sid)) {
      wl->uniform_sid = -1;
      break;
    }
    wl->uniform_sid = t.sid[0];
  }
  // the dispatching kernel runs one row pair per thread: a mixed list takes
  // the one-pair shape of the same width
  if (wl->uniform_sid < 0)
    for (Plan &pl : plans)
      if (kCfgs[pl.cfg].rp != 1)
        for (int c = 0; c < kNumTileConfigs; ++c)
          if (kCfgs[c].rp == 1 && kCfgs[c].tx == kCfgs[pl.cfg].tx)
            pl.cfg = c;
  const long long slots =
      4LL * persistent_grid(1 << 30, wl->lds, wl->uniform_sid);
  int zdef = 64;
  while (zdef > 8 && count(zdef) < slots)
    zdef /= 2;
  for (const Plan &pl : plans) {
    const SepTask &t = wl->tasks[pl.task];
    const TileCfg &c = kCfgs[pl.cfg];
    const int zc = pl.zc > 0 ? pl.zc : zdef;
    for (int comp = 0; comp < pl.ncomp; ++comp)
      for (int z0 = t.rlo[2]; z0 <= t.rhi[2]; z0 += zc)
        for (int y0 = t.rlo[1] & ~1; y0 <= t.rhi[1]; y0 += c.ty)
          for (int x0 = t.rlo[0] & ~1; x0 <= t.rhi[0]; x0 += c.tx) {
            SepTile tl;
            tl.task = pl.task;
            tl.comp = comp;
            tl.x0 = x0;
            tl.y0 = y0;
            tl.z0 = z0;
            tl.z1 = std::min(t.rhi[2], z0 + zc - 1);
            tl.cfg = cfg_code(pl.cfg);
            tl.pad = 0;
            wl->tiles.push_back(tl);
          }
  }
  wl->grid = persistent_grid((int)wl->tiles.size(), wl->lds, wl->uniform_sid);
  schedule_tiles(wl, wl->grid);
}

// Grows a device array to hold n elements (twice that, for later growth);
// cap is set only once the allocation succeeded.
template <typename T> int grow_device(T *&d, size_t &cap, size_t n) {
  if (n <= cap)
    return CXP_OK;
  if (d)
    (void)hipFree(d);
  d = nullptr;
  cap = 0;
  if (hipMalloc(&d, 2 * n * sizeof(T)) != hipSuccess) {
    d = nullptr;
    return CXP_ERR_HIP;
  }
  cap = 2 * n;
  return CXP_OK;
}

// Grows a device array to hold v and copies v into it (asynchronously).
template <typename T>
int upload_vec(const std::vector<T> &v, T *&d, size_t &cap, hipStream_t s) {
  if (int e = grow_device(d, cap, v.size()))
    return e;
  if (!v.empty() && hipMemcpyAsync(d, v.data(), v.size() * sizeof(T),
                                   hipMemcpyHostToDevice, s) != hipSuccess)
    return CXP_ERR_HIP;
  return CXP_OK;
}

// Runs and chunks of k_poison for every patch of the list.
void build_poison(cxp_worklist_s *wl) {
  wl->poison_runs.clear();
  wl->poison_chunks.clear();
  for (const auto &p : wl->patches) {
    const int nx = p.rhi[0] - p.rlo[0] + 1, ny = p.rhi[1] - p.rlo[1] + 1;
    const int nrows = ny * (p.rhi[2] - p.rlo[2] + 1);
    const int per = std::max(1, 2048 / nx);
    for (int c = 0; c < p.ncomp; ++c) {
      const DFab f = dfab(&p.fine, p.fine_comp + c);
      PoisonRun r;
      r.p = f.p + didx(f, p.rlo[0], p.rlo[1], p.rlo[2]);
      r.sy = f.sy;
      r.sz = f.sz;
      r.nx = nx;
      r.ny = ny;
      const int run = (int)wl->poison_runs.size();
      wl->poison_runs.push_back(r);
      for (int row = 0; row < nrows; row += per)
        wl->poison_chunks.push_back({run, row, std::min(per, nrows - row)});
    }
  }
}

// Allocates the coarse-patch buffer of the list's level fills and points
// their patches at it (before the tiles are built from the patches).
int place_gathered(cxp_worklist_s *wl) {
  if (wl->gather_doubles == 0)
    return CXP_OK;
  if ((size_t)wl->gather_doubles > wl->cap_gather) {
    if (wl->d_gather)
      (void)hipFree(wl->d_gather);
    wl->d_gather = nullptr;
    wl->cap_gather = (size_t)wl->gather_doubles;
    if (hipMalloc(&wl->d_gather, wl->cap_gather * sizeof(double)) !=
        hipSuccess) {
      wl->cap_gather = 0;
      return CXP_ERR_HIP;
    }
  }
  for (auto &p : wl->patches)
    if (p.gather_off >= 0)
      p.crse.data = wl->d_gather + p.gather_off;
  return CXP_OK;
}

// Fills the gathered coarse patches: copies from the coarse FABs, then the
// boundary values that are computed from those copies.
int launch_gather(cxp_worklist_s *wl, hipStream_t s) {
  const size_t nblocks = wl->gather_chunks.size() + wl->bcd_chunks.size();
  if (nblocks > 0) {
    hipLaunchKernelGGL(k_gather, dim3((unsigned)nblocks), dim3(256), 0, s,
                       wl->d_gather_runs, wl->d_gather_chunks,
                       (int)wl->gather_chunks.size(), wl->d_bc_runs,
                       wl->d_bcd_chunks, wl->d_bc_direct, wl->d_gather);
    if (const int e = hip_status(hipGetLastError()))
      return e;
  }
  if (wl->bc_chunks.empty())
    return CXP_OK;
  hipLaunchKernelGGL(k_bc, dim3((unsigned)wl->bc_chunks.size()), dim3(256), 0,
                     s, wl->d_bc_runs, wl->d_bc_chunks, wl->d_gather);
  return hip_status(hipGetLastError());
}

// The levels' own boundary conditions, after the fills (their sources are
// points the fills and copies wrote).
int launch_post_bc(cxp_worklist_s *wl, hipStream_t s) {
  if (wl->post_chunks.empty())
    return CXP_OK;
  hipLaunchKernelGGL(k_bc, dim3((unsigned)wl->post_chunks.size()), dim3(256),
                     0, s, wl->d_post_runs, wl->d_post_chunks, nullptr);
  return hip_status(hipGetLastError());
}

int upload_lists(cxp_worklist_s *wl, hipStream_t s) {
  build_poison(wl);
  if (int e = upload_vec(wl->gather_runs, wl->d_gather_runs,
                         wl->cap_gather_runs, s))
    return e;
  if (int e = upload_vec(wl->gather_chunks, wl->d_gather_chunks,
                         wl->cap_gather_chunks, s))
    return e;
  if (int e = upload_vec(wl->bc_runs, wl->d_bc_runs, wl->cap_bc_runs, s))
    return e;
  if (int e = upload_vec(wl->bc_chunks, wl->d_bc_chunks, wl->cap_bc_chunks, s))
    return e;
  if (int e = upload_vec(wl->bcd_chunks, wl->d_bcd_chunks, wl->cap_bcd_chunks, s))
    return e;
  if (int e = upload_vec(wl->bc_direct, wl->d_bc_direct, wl->cap_bc_direct, s))
    return e;
  if (int e = upload_vec(wl->post_runs, wl->d_post_runs, wl->cap_post_runs, s))
    return e;
  if (int e = upload_vec(wl->post_chunks, wl->d_post_chunks,
                         wl->cap_post_chunks, s))
    return e;
  if (int e = upload_vec(wl->poison_runs, wl->d_poison_runs,
                         wl->cap_poison_runs, s))
    return e;
  if (int e = upload_vec(wl->poison_chunks, wl->d_poison_chunks,
                         wl->cap_poison_chunks, s))
    return e;
  if (int e = grow_device(wl->d_tasks, wl->cap_tasks, wl->tasks.size()))
    return e;
  if (int e = grow_device(wl->d_tiles, wl->cap_tiles, wl->tiles.size()))
    return e;
  if (int e = grow_device(wl->d_wg_start, wl->cap_wg, wl->wg_start.size()))
    return e;
  if (!wl->d_ctr) {
    if (hipMalloc(&wl->d_ctr, 9 * 32 * sizeof(int)) != hipSuccess)
      return CXP_ERR_HIP;
    if (hipMemsetAsync(wl->d_ctr, 0, 9 * 32 * sizeof(int), s) != hipSuccess)
      return CXP_ERR_HIP;
  }
  if (!wl->wg_start.empty() &&
      hipMemcpyAsync(wl->d_wg_start, wl->wg_start.data(),
                     wl->wg_start.size() * sizeof(int), hipMemcpyHostToDevice,
                     s) != hipSuccess)
    return CXP_ERR_HIP;
  if (!wl->tasks.empty() &&
      hipMemcpyAsync(wl->d_tasks, wl->tasks.data(),
                     wl->tasks.size() * sizeof(SepTask),
                     hipMemcpyHostToDevice, s) != hipSuccess)
    return CXP_ERR_HIP;
  if (!wl->tiles.empty() &&
      hipMemcpyAsync(wl->d_tiles, wl->tiles.data(),
                     wl->tiles.size() * sizeof(SepTile),
                     hipMemcpyHostToDevice, s) != hipSuccess)
    return CXP_ERR_HIP;
  // the host vectors must stay alive until the copy completed
  if (hipStreamSynchronize(s) != hipSuccess)
    return CXP_ERR_HIP;
  return CXP_OK;
}

int persistent_grid(int ntiles, int lds, int su) {
  const int cus = device_cus();
  int per_cu = 0;
  const hipError_t e = with_fill_kernel(su, [&](auto kern) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, CXP_NT,
                                                        lds);
  });
  if (e != hipSuccess || per_cu <= 0)
    per_cu = 2;
  int g = cus * per_cu;
  g -= g % 8;
  return std::max(1, std::min(ntiles, g));
}

// The separable tiles of a list: one persistent launch.
int launch_sep_list(cxp_worklist_s *wl, hipStream_t s) {
  if (wl->tiles.empty())
    return CXP_OK;
  FillQueue q{};
  const int nq = (int)wl->qstart.size() - 1;
  if (nq >= 1 && nq <= 8 && wl->d_ctr) {
    q.ctr = wl->d_ctr;
    q.nq = nq;
    q.chunk = 1;
    for (int i = 0; i <= nq; ++i)
      q.start[i] = wl->qstart[i];
  }
  with_fill_kernel(wl->uniform_sid, [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(wl->grid), dim3(CXP_NT), (size_t)wl->lds, s,
                       wl->d_tasks, wl->d_tiles, wl->d_wg_start, q);
    return 0;
  });
  return hip_status(hipGetLastError());
}

int launch_batched(cxp_worklist_s *wl, hipStream_t s) {
  if (const int e = launch_sep_list(wl, s))
    return e;

  for (int n : wl->point_patches) {
    const auto &p = wl->patches[n];
    if (eno_ccc_fits(p.op->d, wl->per_group)) {
      if (const int e = eno_ccc_launch(&p.crse, p.crse_comp, &p.fine,
                                       p.fine_comp, p.ncomp, p.rlo, p.rhi,
                                       nullptr, false, s))
        return e;
      continue;
    }
    const PointTask t = make_point_task(p.op->d, &p.crse, p.crse_comp,
                                        &p.fine, p.fine_comp, p.ncomp, p.rlo,
                                        p.rhi);
    const int e = launch_point(t, p.ncomp, wl->per_group, s);
    if (e)
      return e;
  }
  return CXP_OK;
}

} // namespace

extern "C" {

int cxp_version(void) { return 1; }

const char *cxp_status_string(int st) {
  switch (st) {
  case CXP_OK: return "ok";
  case CXP_ERR_TYPE: return "unsupported prolongation type";
  case CXP_ERR_ORDER: return "unsupported prolongation order for this type";
  case CXP_ERR_ARG: return "invalid argument";
  case CXP_ERR_REGION: return "fine region is not contained in the fine FAB";
  case CXP_ERR_COARSE: return "coarse FAB does not cover CoarseBox(region)";
  case CXP_ERR_NCOMP: return "per-group prolongation supports at most 16 components";
  case CXP_ERR_RATIO: return "refinement ratio must be 2";
  case CXP_ERR_HIP: return "HIP runtime error";
  case CXP_ERR_COMP: return "component range outside the FAB";
  case CXP_ERR_STATE: return "object used in the wrong state";
  case CXP_ERR_UNSUPPORTED: return "operation not supported for this operator";
  case CXP_ERR_SCRATCH:
    return "call needs more than the no-allocation scratch array holds";
  }
  return "unknown status";
}

int cxp_op_lookup(const char *type, int order, const int indextype[3],
                  cxp_op *out) {
  if (!type || !indextype || !out)
    return CXP_ERR_ARG;
  std::call_once(g_ops_once, init_ops);
  OpDesc d;
  const int e = op_lookup(type, order, indextype, &d);
  if (e == -1)
    return CXP_ERR_TYPE;
  if (e == -2)
    return CXP_ERR_ORDER;
  if (e)
    return CXP_ERR_ARG;
  const int c = (indextype[0] << 2) | (indextype[1] << 1) | indextype[2];
  *out = &g_ops[d.family][order][c];
  return CXP_OK;
}

int cxp_op_lookup_fallback(const char *type, int order,
                           const int indextype[3], int fallback,
                           cxp_op *out) {
  if (fallback != CXP_FB_KEEP && fallback != CXP_FB_NONE &&
      fallback != CXP_FB_LINEAR)
    return CXP_ERR_ARG;
  cxp_op op = nullptr;
  const int e = cxp_op_lookup(type, order, indextype, &op);
  if (e)
    return e;
  if (fallback == CXP_FB_KEEP) {
    *out = op;
    return CXP_OK;
  }
  const int c = (indextype[0] << 2) | (indextype[1] << 1) | indextype[2];
  *out = &g_ops_fb[op->d.family][order][c][fallback];
  return CXP_OK;
}

int cxp_group_op_lookup(const char *tags, const char *default_type,
                        int default_order, const int indextype[3],
                        cxp_op *out) {
  if (!default_type || !indextype || !out)
    return CXP_ERR_ARG;
  std::string type = default_type;
  int order = default_order;
  if (tags) {
    std::string v;
    bool quoted = false;
    const int t = tags_value(tags, "prolongation_type", v, &quoted);
    if (t < 0)
      return CXP_ERR_ARG;
    if (t == 1)
      type = v;
    const int o = tags_value(tags, "prolongation_order", v, &quoted);
    if (o < 0)
      return CXP_ERR_ARG;
    if (o == 1) {
      char *end = nullptr;
      const long n = std::strtol(v.c_str(), &end, 10);
      if (quoted || v.empty() || *end || n < INT32_MIN || n > INT32_MAX)
        return CXP_ERR_ARG;  // the reference asserts an integer value
      order = (int)n;
    }
  }
  return cxp_op_lookup(type.c_str(), order, indextype, out);
}

int cxp_op_describe(cxp_op op, int cent[3], int intp[3], int order[3],
                    int ghosts[3], int *fb) {
  if (!op)
    return CXP_ERR_ARG;
  for (int d = 0; d < 3; ++d) {
    const SidInfo s = sid_info(op->d.sid[d]);
    if (cent)
      cent[d] = s.cent;
    if (intp)
      intp[d] = (op->d.family == 5 && s.cent == CENT_VC) ? INTP_HERMITE : s.intp;
    if (order)
      order[d] = s.order;
    if (ghosts)
      ghosts[d] = s.ghosts;
  }
  if (fb)
    *fb = op->d.fb;
  return CXP_OK;
}

int cxp_coarse_box(cxp_op op, const int fine_lo[3], const int fine_hi[3],
                   int crse_lo[3], int crse_hi[3]) {
  if (!op || !fine_lo || !fine_hi || !crse_lo || !crse_hi)
    return CXP_ERR_ARG;
  op_coarse_box(op->d, fine_lo, fine_hi, crse_lo, crse_hi);
  return CXP_OK;
}

int cxp_prolongate(cxp_op op, const cxp_fab *crse, int crse_comp,
                   const cxp_fab *fine, int fine_comp, int ncomp,
                   const int region_lo[3], const int region_hi[3],
                   int per_group, void *stream) {
  if (!kCfgsOk)
    return CXP_ERR_STATE;
  const int e = check_prolong(op, crse, crse_comp, fine, fine_comp, ncomp,
                              region_lo, region_hi, per_group);
  if (e)
    return e;
  return prolongate_impl(op, crse, crse_comp, fine, fine_comp, ncomp,
                         region_lo, region_hi, per_group,
                         (hipStream_t)stream);
}

int cxp_prolongate_regions(cxp_op op, const cxp_fab *crse, int crse_comp,
                           const cxp_fab *fine, int fine_comp, int ncomp,
                           int nregions, const int *region_lo,
                           const int *region_hi, int per_group, void *stream) {
  if (!kCfgsOk)
    return CXP_ERR_STATE;
  if (nregions < 0 || (nregions > 0 && (!region_lo || !region_hi)))
    return CXP_ERR_ARG;
  for (int i = 0; i < nregions; ++i)
    if (const int e = check_prolong(op, crse, crse_comp, fine, fine_comp, ncomp,
                                    region_lo + 3 * i, region_hi + 3 * i,
                                    per_group))
      return e;
  // no two regions may share a point (each tile writes its own)
  for (int i = 0; i < nregions; ++i)
    for (int j = i + 1; j < nregions; ++j) {
      bool meet = true;
      for (int d = 0; d < 3; ++d)
        meet = meet && region_lo[3 * i + d] <= region_hi[3 * j + d] &&
               region_lo[3 * j + d] <= region_hi[3 * i + d];
      if (meet)
        return CXP_ERR_REGION;
    }
  hipStream_t s = (hipStream_t)stream;
  if (nregions == 0)
    return CXP_OK;
  // one launch for the separable operators; the others region by region
  if (!op_is_separable(op->d, per_group != 0) || nregions > CXP_MAX_REGIONS) {
    for (int i = 0; i < nregions; ++i)
      if (const int e = prolongate_impl(op, crse, crse_comp, fine, fine_comp,
                                        ncomp, region_lo + 3 * i,
                                        region_hi + 3 * i, per_group, s))
        return e;
    return CXP_OK;
  }
  RegionLaunch r{};
  r.n = nregions;
  int nz[CXP_MAX_REGIONS];
  size_t lds = 8 * (size_t)kLdsDoubles;
  for (int i = 0; i < nregions; ++i) {
    const SepTask t = make_sep_task(op->d, crse, crse_comp, fine, fine_comp,
                                    region_lo + 3 * i, region_hi + 3 * i);
    r.t[i] = t;
    nz[i] = t.rhi[2] - t.rlo[2] + 1;
    if (thin_eligible(t)) {
      r.lx[i] = -1;
      r.ntx[i] = (t.rhi[0] - (t.rlo[0] & ~1) + 16) / 16;
      r.nty[i] = (t.rhi[1] - (t.rlo[1] & ~1) + 16) / 16;
      r.zchunk[i] = nz[i];
      lds = std::max(lds, 8 * (size_t)thin_lds_doubles(t.whi[1] - t.wlo[1],
                                                       t.whi[2] - t.wlo[2]));
      continue;
    }
    int cfg, zc;
    choose_cfg(t, cfg, zc);
    // one row pair per thread: the shape of the same width
    if (kCfgs[cfg].rp != 1)
      for (int c = 0; c < kNumTileConfigs; ++c)
        if (kCfgs[c].rp == 1 && kCfgs[c].tx == kCfgs[cfg].tx)
          cfg = c;
    r.lx[i] = kCfgs[cfg].lx;
    r.ntx[i] = (t.rhi[0] - (t.rlo[0] & ~1) + kCfgs[cfg].tx) / kCfgs[cfg].tx;
    r.nty[i] = (t.rhi[1] - (t.rlo[1] & ~1) + kCfgs[cfg].ty) / kCfgs[cfg].ty;
    r.zchunk[i] = zc > 0 ? std::min(nz[i], zc) : nz[i];
  }
  // as for one box: shorten the untuned z-chunks until the launch has
  // enough workgroups
  auto blocks = [&](int i) {
    return r.ntx[i] * r.nty[i] * ((nz[i] + r.zchunk[i] - 1) / r.zchunk[i]);
  };
  for (;;) {
    long long total = 0;
    int widest = -1;
    for (int i = 0; i < nregions; ++i) {
      total += blocks(i);
      if (r.lx[i] >= 0 && r.zchunk[i] > 8 &&
          (widest < 0 || r.zchunk[i] > r.zchunk[widest]))
        widest = i;
    }
    if (total * ncomp >= 2048 || widest < 0)
      break;
    r.zchunk[widest] = (r.zchunk[widest] + 1) / 2;
  }
  r.start[0] = 0;
  for (int i = 0; i < nregions; ++i)
    r.start[i + 1] = r.start[i] + blocks(i);
  hipLaunchKernelGGL(k_sep_regions, dim3((unsigned)r.start[nregions], ncomp),
                     dim3(CXP_NT), lds, s, r);
  return hip_status(hipGetLastError());
}

int cxp_restrict(const int indextype[3], const cxp_fab *fine, int fine_comp,
                 const cxp_fab *crse, int crse_comp, int ncomp,
                 const int rlo[3], const int rhi[3], void *stream) {
  if (!indextype || !rlo || !rhi || !fab_ok(fine) || !fab_ok(crse) ||
      ncomp < 1)
    return CXP_ERR_ARG;
  if (fine_comp < 0 || crse_comp < 0 || fine_comp + ncomp > fine->ncomp ||
      crse_comp + ncomp > crse->ncomp)
    return CXP_ERR_COMP;
  RestrictTask t;
  for (int d = 0; d < 3; ++d) {
    if (indextype[d] != 0 && indextype[d] != 1)
      return CXP_ERR_ARG;
    if (rhi[d] < rlo[d])
      return CXP_ERR_ARG;
    t.cc[d] = indextype[d];
    t.rlo[d] = rlo[d];
    t.rhi[d] = rhi[d];
  }
  if (!box_in(rlo, rhi, crse->lo, crse->hi))
    return CXP_ERR_REGION;
  for (int d = 0; d < 3; ++d)
    if (2 * rlo[d] < fine->lo[d] || 2 * rhi[d] + t.cc[d] > fine->hi[d])
      return CXP_ERR_COARSE;
  t.fine = dfab(fine, fine_comp);
  t.crse = dfab(crse, crse_comp);
  t.fcs = fine->comp_stride;
  t.ccs = crse->comp_stride;
  const long long n = (long long)(rhi[0] - rlo[0] + 1) * (rhi[1] - rlo[1] + 1) *
                      (rhi[2] - rlo[2] + 1);
  hipLaunchKernelGGL(k_restrict, dim3((unsigned)((n + 255) / 256), ncomp),
                     dim3(256), 0, (hipStream_t)stream, t);
  return hip_status(hipGetLastError());
}

int cxp_restrict_adjoint(const int indextype[3], const cxp_fab *crse_bar,
                         int crse_comp, const cxp_fab *fine_bar, int fine_comp,
                         int ncomp, const int rlo[3], const int rhi[3],
                         void *stream) {
  if (!indextype || !rlo || !rhi || !fab_ok(fine_bar) || !fab_ok(crse_bar) ||
      ncomp < 1)
    return CXP_ERR_ARG;
  if (fine_comp < 0 || crse_comp < 0 || fine_comp + ncomp > fine_bar->ncomp ||
      crse_comp + ncomp > crse_bar->ncomp)
    return CXP_ERR_COMP;
  RestrictTask t;
  for (int d = 0; d < 3; ++d) {
    if (indextype[d] != 0 && indextype[d] != 1)
      return CXP_ERR_ARG;
    if (rhi[d] < rlo[d])
      return CXP_ERR_ARG;
    t.cc[d] = indextype[d];
    t.rlo[d] = rlo[d];
    t.rhi[d] = rhi[d];
  }
  if (!box_in(rlo, rhi, crse_bar->lo, crse_bar->hi))
    return CXP_ERR_REGION;
  for (int d = 0; d < 3; ++d)
    if (2 * rlo[d] < fine_bar->lo[d] || 2 * rhi[d] + t.cc[d] > fine_bar->hi[d])
      return CXP_ERR_COARSE;
  t.fine = dfab(fine_bar, fine_comp);
  t.crse = dfab(crse_bar, crse_comp);
  t.fcs = fine_bar->comp_stride;
  t.ccs = crse_bar->comp_stride;
  long long n = 1;
  for (int d = 0; d < 3; ++d)
    n *= 2 * (rhi[d] - rlo[d]) + 1 + t.cc[d];
  hipLaunchKernelGGL(k_restrict_adjoint,
                     dim3((unsigned)((n + 255) / 256), ncomp), dim3(256), 0,
                     (hipStream_t)stream, t);
  return hip_status(hipGetLastError());
}

int cxp_prolongate_adjoint(cxp_op op, const cxp_fab *crse_in,
                           const cxp_fab *crse_bar, int crse_comp,
                           const cxp_fab *fine_bar, int fine_comp, int ncomp,
                           const int rlo[3], const int rhi[3], int per_group,
                           void *stream) {
  const int e = check_prolong(op, crse_bar, crse_comp, fine_bar, fine_comp,
                              ncomp, rlo, rhi, per_group);
  if (e)
    return e;
  const OpDesc &od = op->d;
  if (op_has_minmod(od))
    return minmod_adjoint_launch(op, crse_in, crse_bar, crse_comp, fine_bar,
                                 fine_comp, ncomp, rlo, rhi, per_group,
                                 (hipStream_t)stream);
  if (op_has_eno(od) || op_fallback_per_var(od) || op_fallback_per_group(od))
    return adjoint_point_launch(op, crse_in, crse_bar, crse_comp, fine_bar,
                                fine_comp, ncomp, rlo, rhi, per_group,
                                (hipStream_t)stream);
  return adjoint_launch(od, crse_in, crse_bar, crse_comp, fine_bar,
                        fine_comp, ncomp, rlo, rhi, per_group,
                        (hipStream_t)stream);
}

int cxp_prolongate_adjoint_noalloc(cxp_op op, const cxp_fab *crse_in,
                                   const cxp_fab *crse_bar, int crse_comp,
                                   const cxp_fab *fine_bar, int fine_comp,
                                   int ncomp, const int rlo[3],
                                   const int rhi[3], int per_group,
                                   void *stream) {
  const int e = check_prolong(op, crse_bar, crse_comp, fine_bar, fine_comp,
                              ncomp, rlo, rhi, per_group);
  if (e)
    return e;
  const OpDesc &od = op->d;
  if (op_has_minmod(od)) {
    if (const int e2 = adjoint_point_check(op, crse_in, crse_comp, fine_bar,
                                           fine_comp, ncomp, rlo, rhi,
                                           per_group))
      return e2;
    if (op_all_minmod(od) && mm_tiled_fits(fine_bar))  // keeps no codes
      return minmod_adjoint_launch(op, crse_in, crse_bar, crse_comp, fine_bar,
                                   fine_comp, ncomp, rlo, rhi, per_group,
                                   (hipStream_t)stream);
    if (mm_code_bytes(od, rlo, rhi, ncomp) > CXP_NOALLOC_SCRATCH_BYTES)
      return CXP_ERR_SCRATCH;
    unsigned char *codes = nullptr;
    if (const int e2 = scratch_address(&codes))
      return e2;
    return minmod_adjoint_launch(op, crse_in, crse_bar, crse_comp, fine_bar,
                                 fine_comp, ncomp, rlo, rhi, per_group,
                                 (hipStream_t)stream, codes);
  }
  if (!(op_has_eno(od) || op_fallback_per_var(od) ||
        op_fallback_per_group(od)))
    // the linear adjoints use no scratch
    return adjoint_launch(od, crse_in, crse_bar, crse_comp, fine_bar,
                          fine_comp, ncomp, rlo, rhi, per_group,
                          (hipStream_t)stream);
  if (const int e2 = adjoint_point_check(op, crse_in, crse_comp, fine_bar,
                                         fine_comp, ncomp, rlo, rhi,
                                         per_group))
    return e2;
  if (adjoint_code_bytes(ncomp, rlo, rhi, per_group) >
      (size_t)CXP_NOALLOC_SCRATCH_BYTES)
    return CXP_ERR_SCRATCH;
  unsigned char *codes = nullptr;
  if (const int e2 = scratch_address(&codes))
    return e2;
  return adjoint_point_run(op, crse_in, crse_bar, crse_comp, fine_bar,
                           fine_comp, ncomp, rlo, rhi, per_group, codes,
                           (hipStream_t)stream);
}

// Operators whose forward takes a branch per fine point (ENO shift, linear
// fallback); MINMOD operators record their per-pass codes instead.
bool op_records_branches(const OpDesc &od, int per_group) {
  if (op_has_minmod(od))
    return false;
  return !op_is_separable(od, per_group != 0);
}

int cxp_branch_bytes(cxp_op op, const int rlo[3], const int rhi[3], int ncomp,
                     int per_group, int64_t *bytes) {
  if (!op || !rlo || !rhi || !bytes || ncomp < 1)
    return CXP_ERR_ARG;
  int64_t n = 1;
  for (int d = 0; d < 3; ++d) {
    if (rhi[d] < rlo[d])
      return CXP_ERR_ARG;
    n *= rhi[d] - rlo[d] + 1;
  }
  if (op_has_minmod(op->d))
    *bytes = mm_code_bytes(op->d, rlo, rhi, ncomp);
  else
    *bytes = op_records_branches(op->d, per_group)
                 ? n * (per_group ? 1 : ncomp)
                 : 0;
  return CXP_OK;
}

int cxp_prolongate_record(cxp_op op, const cxp_fab *crse, int crse_comp,
                          const cxp_fab *fine, int fine_comp, int ncomp,
                          const int rlo[3], const int rhi[3], int per_group,
                          unsigned char *branches, void *stream) {
  const int e = check_prolong(op, crse, crse_comp, fine, fine_comp, ncomp, rlo,
                              rhi, per_group);
  if (e)
    return e;
  if (op_has_minmod(op->d) && mm_code_bytes(op->d, rlo, rhi, ncomp) > 0) {
    if (!branches)
      return CXP_ERR_ARG;
    if (const int e2 = prolongate_impl(op, crse, crse_comp, fine, fine_comp,
                                       ncomp, rlo, rhi, per_group,
                                       (hipStream_t)stream))
      return e2;
    return mm_record_launch(op->d, crse, crse_comp, ncomp, rlo, rhi, branches,
                            (hipStream_t)stream);
  }
  if (!op_records_branches(op->d, per_group))
    return prolongate_impl(op, crse, crse_comp, fine, fine_comp, ncomp, rlo,
                           rhi, per_group, (hipStream_t)stream);
  if (!branches)
    return CXP_ERR_ARG;
  if (eno_ccc_fits(op->d, per_group))
    return eno_ccc_launch(crse, crse_comp, fine, fine_comp, ncomp, rlo, rhi,
                          branches, false, (hipStream_t)stream);
  const PointTask t = make_point_task(op->d, crse, crse_comp, fine, fine_comp,
                                      ncomp, rlo, rhi);
  return launch_point(t, ncomp, per_group, (hipStream_t)stream, branches);
}

int cxp_prolongate_adjoint_regions(cxp_op op, const cxp_fab *crse_bar,
                                   int crse_comp, const cxp_fab *fine_bar,
                                   int fine_comp, int ncomp, int nregions,
                                   const int *region_lo, const int *region_hi,
                                   int per_group, void *stream) {
  if (nregions < 0 || (nregions > 0 && (!region_lo || !region_hi)))
    return CXP_ERR_ARG;
  for (int i = 0; i < nregions; ++i)
    if (const int e = check_prolong(op, crse_bar, crse_comp, fine_bar,
                                    fine_comp, ncomp, region_lo + 3 * i,
                                    region_hi + 3 * i, per_group))
      return e;
  const OpDesc &od = op->d;
  // linear operators only: the others are linearised at a forward input
  if (op_has_minmod(od) || op_has_eno(od) || op_fallback_per_var(od) ||
      op_fallback_per_group(od))
    return CXP_ERR_UNSUPPORTED;
  return adjoint_regions_launch(od, crse_bar, crse_comp, fine_bar, fine_comp,
                                ncomp, nregions, region_lo, region_hi,
                                (hipStream_t)stream);
}

int cxp_prolongate_adjoint_recorded(cxp_op op, const unsigned char *branches,
                                    const cxp_fab *crse_bar, int crse_comp,
                                    const cxp_fab *fine_bar, int fine_comp,
                                    int ncomp, const int rlo[3],
                                    const int rhi[3], int per_group,
                                    void *stream) {
  const int e = check_prolong(op, crse_bar, crse_comp, fine_bar, fine_comp,
                              ncomp, rlo, rhi, per_group);
  if (e)
    return e;
  const OpDesc &od = op->d;
  if (op_has_minmod(od) && mm_code_bytes(od, rlo, rhi, ncomp) > 0) {
    if (!branches)
      return CXP_ERR_ARG;
    return mm_gather_launch(od, crse_bar, crse_comp, fine_bar, fine_comp,
                            ncomp, rlo, rhi, branches, (hipStream_t)stream);
  }
  if (!op_records_branches(od, per_group))
    return adjoint_launch(od, nullptr, crse_bar, crse_comp, fine_bar,
                          fine_comp, ncomp, rlo, rhi, per_group,
                          (hipStream_t)stream);
  if (!branches)
    return CXP_ERR_ARG;
  if (eno_ccc_fits(od, per_group))
    return eno_ccc_adj_launch(crse_bar, crse_comp, fine_bar, fine_comp, ncomp,
                              rlo, rhi, branches, (hipStream_t)stream);
  const AdjPointTask t =
      make_adj_point_task(op, crse_bar, crse_bar, crse_comp, fine_bar,
                          fine_comp, ncomp, rlo, rhi, per_group);
  return launch_adj_gather(t, branches, ncomp, (hipStream_t)stream);
}

// ---- worklists

int cxp_worklist_create(cxp_worklist *out) {
  if (!out)
    return CXP_ERR_ARG;
  cxp_worklist_s *wl = new cxp_worklist_s();
  {
    std::lock_guard<std::mutex> lk(g_lists_mu);
    wl->id = g_next_list_id++;
    g_lists[wl->id] = wl;
  }
  *out = wl;
  return CXP_OK;
}

int cxp_worklist_destroy(cxp_worklist wl) {
  if (!wl)
    return CXP_ERR_ARG;
  {
    std::lock_guard<std::mutex> lk(g_lists_mu);
    g_lists.erase(wl->id);
  }
  (void)list_drain(wl);
  if (wl->done)
    (void)hipEventDestroy(wl->done);
  if (wl->d_tasks)
    (void)hipFree(wl->d_tasks);
  if (wl->d_tiles)
    (void)hipFree(wl->d_tiles);
  if (wl->d_wg_start)
    (void)hipFree(wl->d_wg_start);
  if (wl->d_ctr)
    (void)hipFree(wl->d_ctr);
  if (wl->d_gather)
    (void)hipFree(wl->d_gather);
  if (wl->d_gather_runs)
    (void)hipFree(wl->d_gather_runs);
  if (wl->d_gather_chunks)
    (void)hipFree(wl->d_gather_chunks);
  if (wl->d_bc_runs)
    (void)hipFree(wl->d_bc_runs);
  if (wl->d_bc_chunks)
    (void)hipFree(wl->d_bc_chunks);
  if (wl->d_bcd_chunks)
    (void)hipFree(wl->d_bcd_chunks);
  if (wl->d_bc_direct)
    (void)hipFree(wl->d_bc_direct);
  if (wl->d_post_runs)
    (void)hipFree(wl->d_post_runs);
  if (wl->d_post_chunks)
    (void)hipFree(wl->d_post_chunks);
  if (wl->d_poison_runs)
    (void)hipFree(wl->d_poison_runs);
  if (wl->d_poison_chunks)
    (void)hipFree(wl->d_poison_chunks);
  delete wl;
  return CXP_OK;
}

int cxp_worklist_clear(cxp_worklist wl) {
  if (!wl)
    return CXP_ERR_ARG;
  wl->patches.clear();
  wl->gather_runs.clear();
  wl->gather_chunks.clear();
  wl->bc_runs.clear();
  wl->bc_chunks.clear();
  wl->bcd_chunks.clear();
  wl->bc_direct.clear();
  wl->post_runs.clear();
  wl->post_chunks.clear();
  wl->gather_doubles = 0;
  wl->uploaded = false;
  touch(wl);
  wl->fine_points = wl->algo_bytes = 0;
  return CXP_OK;
}

int cxp_worklist_append(cxp_worklist wl, cxp_op op, const cxp_fab *crse,
                        int crse_comp, const cxp_fab *fine, int fine_comp,
                        int ncomp, const int rlo[3], const int rhi[3]) {
  if (!wl)
    return CXP_ERR_ARG;
  const int e = check_prolong(op, crse, crse_comp, fine, fine_comp, ncomp, rlo,
                              rhi, 0);
  if (e)
    return e;
  cxp_worklist_s::Patch p;
  p.op = op;
  p.crse = *crse;
  p.fine = *fine;
  p.crse_comp = crse_comp;
  p.fine_comp = fine_comp;
  p.ncomp = ncomp;
  for (int d = 0; d < 3; ++d) {
    p.rlo[d] = rlo[d];
    p.rhi[d] = rhi[d];
  }
  wl->patches.push_back(p);
  wl->uploaded = false;
  touch(wl);
  const int64_t npts = int64_t(rhi[0] - rlo[0] + 1) * (rhi[1] - rlo[1] + 1) *
                       (rhi[2] - rlo[2] + 1);
  wl->fine_points += npts * ncomp;
  wl->algo_bytes += 8 * ncomp * (npts + patch_coarse_points(p));
  return CXP_OK;
}

int cxp_worklist_append_shell(cxp_worklist wl, cxp_op op, const cxp_fab *crse,
                              int crse_comp, const cxp_fab *fine,
                              int fine_comp, int ncomp, const int ilo[3],
                              const int ihi[3]) {
  if (!wl || !fine || !ilo || !ihi || !op)
    return CXP_ERR_ARG;
  const int *glo = fine->lo, *ghi = fine->hi;
  for (int d = 0; d < 3; ++d)
    if (ilo[d] < glo[d] || ihi[d] > ghi[d] || ilo[d] > ihi[d])
      return CXP_ERR_REGION;
  // six slabs: z full (x,y grown), y (x grown, z interior), x (y,z interior)
  struct Slab {
    int lo[3], hi[3];
  };
  std::vector<Slab> slabs;
  auto add = [&](int x0, int x1, int y0, int y1, int z0, int z1) {
    if (x0 <= x1 && y0 <= y1 && z0 <= z1)
      slabs.push_back({{x0, y0, z0}, {x1, y1, z1}});
  };
  add(glo[0], ghi[0], glo[1], ghi[1], glo[2], ilo[2] - 1);
  add(glo[0], ghi[0], glo[1], ghi[1], ihi[2] + 1, ghi[2]);
  add(glo[0], ghi[0], glo[1], ilo[1] - 1, ilo[2], ihi[2]);
  add(glo[0], ghi[0], ihi[1] + 1, ghi[1], ilo[2], ihi[2]);
  add(glo[0], ilo[0] - 1, ilo[1], ihi[1], ilo[2], ihi[2]);
  add(ihi[0] + 1, ghi[0], ilo[1], ihi[1], ilo[2], ihi[2]);
  // validate all slabs before appending any
  for (const Slab &sl : slabs) {
    const int e = check_prolong(op, crse, crse_comp, fine, fine_comp, ncomp,
                                sl.lo, sl.hi, 0);
    if (e)
      return e;
  }
  const int64_t bytes_before = wl->algo_bytes;
  for (const Slab &sl : slabs) {
    const int e = cxp_worklist_append(wl, op, crse, crse_comp, fine, fine_comp,
                                      ncomp, sl.lo, sl.hi);
    if (e)
      return e;
  }
  // algorithmic bytes count coarse points once per box over the union of the
  // slabs' coarse boxes: recount them with a mask
  int64_t fine_pts = 0;
  for (const Slab &sl : slabs)
    fine_pts += int64_t(sl.hi[0] - sl.lo[0] + 1) * (sl.hi[1] - sl.lo[1] + 1) *
                (sl.hi[2] - sl.lo[2] + 1);
  int ulo[3] = {1 << 30, 1 << 30, 1 << 30}, uhi[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
  std::vector<std::pair<std::vector<int>, std::vector<int>>> cboxes;
  for (const Slab &sl : slabs) {
    int clo[3], chi[3];
    op_coarse_box(op->d, sl.lo, sl.hi, clo, chi);
    cboxes.push_back({{clo[0], clo[1], clo[2]}, {chi[0], chi[1], chi[2]}});
    for (int d = 0; d < 3; ++d) {
      ulo[d] = std::min(ulo[d], clo[d]);
      uhi[d] = std::max(uhi[d], chi[d]);
    }
  }
  const int ux = uhi[0] - ulo[0] + 1, uy = uhi[1] - ulo[1] + 1,
            uz = uhi[2] - ulo[2] + 1;
  std::vector<unsigned char> mask((size_t)ux * uy * uz, 0);
  for (const auto &cb : cboxes)
    for (int k = cb.first[2]; k <= cb.second[2]; ++k)
      for (int j = cb.first[1]; j <= cb.second[1]; ++j)
        for (int i = cb.first[0]; i <= cb.second[0]; ++i)
          mask[((size_t)(k - ulo[2]) * uy + (j - ulo[1])) * ux + (i - ulo[0])] = 1;
  int64_t ncoarse = 0;
  for (unsigned char m : mask)
    ncoarse += m;
  wl->algo_bytes = bytes_before + 8 * (int64_t)ncomp * (fine_pts + ncoarse);
  return CXP_OK;
}

namespace {

// Inclusive index box for the level fill's box algebra.
struct IBox {
  int lo[3], hi[3];
  bool empty() const {
    return hi[0] < lo[0] || hi[1] < lo[1] || hi[2] < lo[2];
  }
  int64_t volume() const {
    return empty() ? 0
                   : int64_t(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) *
                         (hi[2] - lo[2] + 1);
  }
};

IBox ibox(const int lo[3], const int hi[3]) {
  IBox b;
  for (int d = 0; d < 3; ++d) {
    b.lo[d] = lo[d];
    b.hi[d] = hi[d];
  }
  return b;
}

IBox intersect(const IBox &a, const IBox &b) {
  IBox r;
  for (int d = 0; d < 3; ++d) {
    r.lo[d] = std::max(a.lo[d], b.lo[d]);
    r.hi[d] = std::min(a.hi[d], b.hi[d]);
  }
  return r;
}

bool inside(const IBox &a, const IBox &b) {
  return box_in(a.lo, a.hi, b.lo, b.hi);
}

// a minus b as disjoint boxes: z-low and z-high slabs over a's full x-y
// extent, then y-low and y-high over the overlap's z range, then x-low and
// x-high (the ghost-shell slab order of cxp_worklist_append_shell).
void subtract(const IBox &a, const IBox &b, std::vector<IBox> &out) {
  const IBox c = intersect(a, b);
  if (c.empty()) {
    if (!a.empty())
      out.push_back(a);
    return;
  }
  auto add = [&](int x0, int x1, int y0, int y1, int z0, int z1) {
    IBox r{{x0, y0, z0}, {x1, y1, z1}};
    if (!r.empty())
      out.push_back(r);
  };
  add(a.lo[0], a.hi[0], a.lo[1], a.hi[1], a.lo[2], c.lo[2] - 1);
  add(a.lo[0], a.hi[0], a.lo[1], a.hi[1], c.hi[2] + 1, a.hi[2]);
  add(a.lo[0], a.hi[0], a.lo[1], c.lo[1] - 1, c.lo[2], c.hi[2]);
  add(a.lo[0], a.hi[0], c.hi[1] + 1, a.hi[1], c.lo[2], c.hi[2]);
  add(a.lo[0], c.lo[0] - 1, c.lo[1], c.hi[1], c.lo[2], c.hi[2]);
  add(c.hi[0] + 1, a.hi[0], c.lo[1], c.hi[1], c.lo[2], c.hi[2]);
}

void subtract_all(std::vector<IBox> &pieces, const IBox &b) {
  std::vector<IBox> next;
  for (const IBox &p : pieces)
    subtract(p, b, next);
  pieces.swap(next);
}

bool level_ok(const cxp_level *L) {
  if (!L || L->nfabs < 0 || (L->nfabs > 0 && (!L->fabs || !L->valid)))
    return false;
  for (int d = 0; d < 3; ++d)
    if (L->domain_hi[d] < L->domain_lo[d])
      return false;
  for (int i = 0; i < L->nfabs; ++i) {
    const int *v = L->valid + 6 * i;
    if (!fab_ok(&L->fabs[i]) || !box_in(v, v + 3, L->fabs[i].lo, L->fabs[i].hi))
      return false;
    for (int d = 0; d < 3; ++d)
      if (v[3 + d] < v[d])
        return false;
  }
  return true;
}

// Distinct points of a union of boxes.
int64_t union_points(const std::vector<IBox> &boxes) {
  std::vector<IBox> acc;
  int64_t n = 0;
  for (const IBox &b : boxes) {
    std::vector<IBox> fresh{b};
    for (const IBox &a : acc)
      subtract_all(fresh, a);
    for (const IBox &f : fresh)
      n += f.volume();
    acc.push_back(b);
  }
  return n;
}

// Reflection offsets of a domain: src = offset - dst mirrors about the
// first and last points for vertex centring and half a cell outside them
// for cell centring (boundaries_impl.hxx:286-298).
int refl_offset(const IBox &dom, int face, int d, int cell) {
  return face ? 2 * dom.hi[d] + cell : 2 * dom.lo[d] - cell;
}

// The box a coarse patch with boundary points must hold: C and every point
// BoundaryCondition::apply reads for C's points outside the domain
// (boundaries_impl.hxx:249-298 and :322-360); those all lie inside the
// domain.  CXP_ERR_ARG if C reaches a face without a condition or the
// domain is too narrow for the sources.
int bc_closure(const cxp_bc &bc, const OpDesc &od, const IBox &dom,
               const IBox &C, IBox &E) {
  E = C;
  for (int d = 0; d < 3; ++d)
    for (int face = 0; face < 2; ++face) {
      if (face ? C.hi[d] <= dom.hi[d] : C.lo[d] >= dom.lo[d])
        continue;
      const int edge = face ? dom.hi[d] : dom.lo[d];
      auto need = [&](int x) {
        if (x < dom.lo[d] || x > dom.hi[d])
          return false;
        E.lo[d] = std::min(E.lo[d], x);
        E.hi[d] = std::max(E.hi[d], x);
        return true;
      };
      switch (bc.type[face][d]) {
      case CXP_BC_DIRICHLET:
      case CXP_BC_PERIODIC:  // images, not sources in the box
        break;
      case CXP_BC_NEUMANN:
      case CXP_BC_ROBIN:
        need(edge);
        break;
      case CXP_BC_LINEAR_EXTRAPOLATION:
        if (!need(edge) || !need(face ? edge - 1 : edge + 1))
          return CXP_ERR_ARG;
        break;
      case CXP_BC_REFLECTION: {
        // mirrors of C's outside range [a, b] in direction d
        const int a = face ? std::max(C.lo[d], dom.hi[d] + 1) : C.lo[d];
        const int b = face ? C.hi[d] : std::min(C.hi[d], dom.lo[d] - 1);
        const int off = refl_offset(dom, face, d, od.indextype[d]);
        if (!need(off - a) || !need(off - b))
          return CXP_ERR_ARG;
        break;
      }
      default:
        return CXP_ERR_ARG;
      }
    }
  return CXP_OK;
}

bool bc_ok(const cxp_bc &bc, int ncomp) {
  if (ncomp > 16)
    return false;
  bool dir = false, rob = false, ref = false;
  for (int f = 0; f < 2; ++f)
    for (int d = 0; d < 3; ++d) {
      const int t = bc.type[f][d];
      if (t < CXP_BC_NONE || t > CXP_BC_PERIODIC)
        return false;
      // periodic directions are periodic on both faces
      if ((t == CXP_BC_PERIODIC) != (bc.type[1 - f][d] == CXP_BC_PERIODIC))
        return false;
      dir |= t == CXP_BC_DIRICHLET;
      rob |= t == CXP_BC_ROBIN;
      ref |= t == CXP_BC_REFLECTION;
    }
  if ((dir && !bc.dirichlet_values) || (rob && !bc.robin_values) ||
      (ref && !bc.parities))
    return false;
  if (bc.parities)
    for (int i = 0; i < 3 * ncomp; ++i)
      if (bc.parities[i] != 1 && bc.parities[i] != -1)
        return false;
  return true;
}

// Points per chunk of the level fills' copy and boundary launches: one or
// two points per thread keep more independent loads in flight than longer
// per-thread loops (both kernels are latency-bound).
constexpr int kGatherChunkPts = 512;
constexpr int kBcChunkPts = 256;

// Consistency of new chunks with the runs they index (runs[k] is run
// run0 + k): every chunk names one of them, and a boundary chunk's points
// lie in its run's box.  Checked on the host before anything reaches the
// device.
bool gather_chunks_ok(const std::vector<GatherChunk> &chunks, int run0,
                      size_t nruns) {
  for (const GatherChunk &c : chunks)
    if (c.run < run0 || c.run >= run0 + (int)nruns || c.row0 < 0 ||
        c.nrows < 1)
      return false;
  return true;
}
bool bc_chunks_ok(const std::vector<BcChunk> &chunks, int run0,
                  const std::vector<BcRun> &runs) {
  for (const BcChunk &c : chunks) {
    if (c.run < run0 || c.run >= run0 + (int)runs.size() || c.bnx < 1 ||
        c.bny < 1 || c.first < 0 || c.count < 1)
      return false;
    const BcRun &r = runs[c.run - run0];
    const int last = c.first + c.count - 1;
    const int ext[3] = {c.bnx - 1, c.bny - 1, last / (c.bnx * c.bny)};
    const int ext0[3] = {0, 0, c.first / (c.bnx * c.bny)};
    for (int d = 0; d < 3; ++d)
      if (c.blo[d] + ext0[d] < r.lo[d] || c.blo[d] + ext[d] > r.hi[d])
        return false;
  }
  return true;
}

// For a gathered coarse patch: the coarse level, whose valid data holds
// the sources of the patch's boundary points, the periodic directions and
// periods, and where chunks that read it directly go.
struct DirectSrc {
  const cxp_level *crse;
  int crse_comp;
  const bool *periodic;
  const int *period;
  std::vector<BcChunk> *chunks;
  std::vector<BcDirect> *dirs;
  int dir0;
};

// Boundary runs of box E (its ncomp components at p + off + c comp_stride;
// p null: the patch buffer) over a domain: one run per component, chunks
// of kBcChunkPts points over the parts of E outside the domain.  runs[k]
// becomes run run0 + k.  With direct, each part is split where each
// direction is below, inside or above the domain; a part whose sources
// (edge points, mirror images, periodic images, and for linear
// extrapolation the points beside the edge) one coarse FAB's valid region
// holds reads them there, in the gather launch; the others read the patch
// after the gather.  Parts outside only in periodic directions are the
// gather's images and get no chunks.
void add_bc_runs(const cxp_bc &bc, const int it[3], const IBox &E,
                 const IBox &dom, double *p, int64_t off, int64_t comp_stride,
                 int ncomp, int run0, std::vector<BcRun> &runs,
                 std::vector<BcChunk> &chunks,
                 const DirectSrc *direct = nullptr) {
  std::vector<IBox> outs;
  subtract(E, dom, outs);
  for (int c = 0; c < ncomp; ++c) {
    BcRun r;
    r.p = p;
    r.off = off + c * comp_stride;
    for (int d = 0; d < 3; ++d) {
      r.lo[d] = E.lo[d];
      r.hi[d] = E.hi[d];
      r.imin[d] = dom.lo[d];
      r.imax[d] = dom.hi[d] + 1;
      for (int f = 0; f < 2; ++f) {
        r.type[f][d] = bc.type[f][d];
        r.refl[f][d] = refl_offset(dom, f, d, it[d]);
      }
      r.parity[d] = bc.parities ? bc.parities[3 * c + d] : 1;
      r.xmin[d] = bc.xmin[d];
      r.dx[d] = bc.dx[d];
    }
    r.dirichlet = bc.dirichlet_values ? bc.dirichlet_values[c] : 0.0;
    r.robin = bc.robin_values ? bc.robin_values[c] : 0.0;
    const int run = run0 + (int)runs.size();
    runs.push_back(r);
    auto emit = [&](std::vector<BcChunk> &into, const IBox &P, int dir) {
      const int nx = P.hi[0] - P.lo[0] + 1, ny = P.hi[1] - P.lo[1] + 1;
      const int vol = (int)P.volume();
      for (int q = 0; q < vol; q += kBcChunkPts) {
        BcChunk ch;
        ch.run = run;
        for (int d = 0; d < 3; ++d)
          ch.blo[d] = P.lo[d];
        ch.bnx = nx;
        ch.bny = ny;
        ch.first = q;
        ch.count = std::min(kBcChunkPts, vol - q);
        ch.direct = dir;
        into.push_back(ch);
      }
    };
    for (const IBox &P : outs) {
      if (!direct) {
        emit(chunks, P, -1);
        continue;
      }
      for (int cl = 0; cl < 27; ++cl) {
        const int cls[3] = {cl % 3, (cl / 3) % 3, cl / 9};  // below, in, above
        IBox S;
        for (int d = 0; d < 3; ++d) {
          S.lo[d] = cls[d] == 0 ? P.lo[d]
                  : cls[d] == 1 ? std::max(P.lo[d], dom.lo[d])
                                : std::max(P.lo[d], dom.hi[d] + 1);
          S.hi[d] = cls[d] == 0 ? std::min(P.hi[d], dom.lo[d] - 1)
                  : cls[d] == 1 ? std::min(P.hi[d], dom.hi[d])
                                : P.hi[d];
        }
        if (S.empty())
          continue;
        // how each direction maps S's points to their sources: identity
        // (inside, or outside a periodic face: then shifted to the image),
        // the edge (and for linear extrapolation the point beside it), or
        // the mirror o - dst; a Dirichlet direction needs no source at all
        enum { kIdentity, kEdge, kMirror };
        int kind[3] = {kIdentity, kIdentity, kIdentity};
        int shift[3] = {0, 0, 0}, e0[3] = {0, 0, 0}, e1[3] = {0, 0, 0},
            mo[3] = {0, 0, 0};
        bool cond = false, dirichlet = false;
        for (int d = 0; d < 3; ++d) {
          if (cls[d] == 1)
            continue;
          const int face = cls[d] == 2 ? 1 : 0;
          const int t = bc.type[face][d];
          const int edge = face ? dom.hi[d] : dom.lo[d];
          if (t == CXP_BC_PERIODIC) {
            shift[d] = face ? -direct->period[d] : direct->period[d];
            continue;
          }
          if (t == CXP_BC_NONE)
            continue;
          cond = true;
          if (t == CXP_BC_DIRICHLET) {
            dirichlet = true;
          } else if (t == CXP_BC_REFLECTION) {
            kind[d] = kMirror;
            mo[d] = refl_offset(dom, face, d, it[d]);
          } else {
            kind[d] = kEdge;
            e0[d] = edge;
            e1[d] = t == CXP_BC_LINEAR_EXTRAPOLATION
                        ? (face ? edge - 1 : edge + 1)
                        : edge;
          }
        }
        if (!cond)
          continue;  // outside only in periodic directions
        // split S by the coarse FAB whose valid region holds the sources:
        // the preimage of each valid box under the map
        std::vector<IBox> left{S};
        for (int k = 0; k < direct->crse->nfabs && !left.empty(); ++k) {
          const IBox v = ibox(direct->crse->valid + 6 * k,
                              direct->crse->valid + 6 * k + 3);
          IBox Q = S;
          for (int d = 0; d < 3 && !dirichlet; ++d) {
            if (kind[d] == kIdentity) {
              Q.lo[d] = std::max(S.lo[d], v.lo[d] - shift[d]);
              Q.hi[d] = std::min(S.hi[d], v.hi[d] - shift[d]);
            } else if (kind[d] == kEdge) {
              if (e0[d] < v.lo[d] || e0[d] > v.hi[d] || e1[d] < v.lo[d] ||
                  e1[d] > v.hi[d])
                Q.hi[d] = Q.lo[d] - 1;
            } else {
              Q.lo[d] = std::max(S.lo[d], mo[d] - v.hi[d]);
              Q.hi[d] = std::min(S.hi[d], mo[d] - v.lo[d]);
            }
          }
          std::vector<IBox> rest;
          for (const IBox &L : left) {
            const IBox X = intersect(L, Q);
            if (X.empty()) {
              rest.push_back(L);
              continue;
            }
            const DFab f = dfab(&direct->crse->fabs[k], direct->crse_comp + c);
            BcDirect dd;
            dd.src = f.p;
            for (int d = 0; d < 3; ++d) {
              dd.lo[d] = f.lo[d];
              dd.shift[d] = shift[d];
            }
            dd.sy = f.sy;
            dd.sz = f.sz;
            direct->dirs->push_back(dd);
            emit(*direct->chunks, X, direct->dir0 + (int)direct->dirs->size() - 1);
            subtract(L, X, rest);
          }
          left.swap(rest);
        }
        for (const IBox &L : left)
          emit(chunks, L, -1);
      }
    }
  }
}

// What a level fill prolongates in each fine FAB:
//  * ghosts: its box within the fine domain grown by the ghost width,
//    minus every fine valid box (FillPatch_ProlongateGhosts);
//  * whole: its box within the fine domain (FillPatch_NewLevel,
//    fillpatch.cxx:160-170: the domain is grown in periodic directions
//    only, and none are modelled);
//  * remake: its box within the grown fine domain minus every valid box of
//    the old fine level, whose data is copied into the rest
//    (FillPatch_RemakeLevel, fillpatch.cxx:185-227).
enum class FillMode { ghosts, whole, remake };

int append_level_impl(cxp_worklist wl, cxp_op op, const cxp_level *crse,
                      int crse_comp, const cxp_level *fine, int fine_comp,
                      int ncomp, const cxp_bc *bc,
                      FillMode mode = FillMode::ghosts,
                      const cxp_level *old = nullptr) {
  if (!wl || !op || !level_ok(crse) || !level_ok(fine) || ncomp < 1)
    return CXP_ERR_ARG;
  if (mode == FillMode::remake && !level_ok(old))
    return CXP_ERR_ARG;
  if (bc && !bc_ok(*bc, ncomp))
    return CXP_ERR_ARG;
  for (const cxp_level *L : {crse, fine, old})
    for (int i = 0; L && i < L->nfabs; ++i) {
      const int c0 = L == crse ? crse_comp : fine_comp;
      if (c0 < 0 || c0 + ncomp > L->fabs[i].ncomp)
        return CXP_ERR_COMP;
      if (fab_component_bytes(&L->fabs[i]) > kMaxComponentBytes)
        return CXP_ERR_ARG;
    }
  // ghost width of the fine FABs and the grown fine domain (dest_domain,
  // fillpatch.cxx:116-119)
  int g[3] = {0, 0, 0};
  for (int i = 0; i < fine->nfabs; ++i)
    for (int d = 0; d < 3; ++d)
      g[d] = std::max(g[d], std::max(fine->valid[6 * i + d] - fine->fabs[i].lo[d],
                                     fine->fabs[i].hi[d] - fine->valid[6 * i + 3 + d]));
  // periodic directions of the coarse conditions, and their periods
  bool periodic[3] = {false, false, false};
  int period[3] = {0, 0, 0};
  for (int d = 0; d < 3 && bc; ++d) {
    periodic[d] = bc->type[0][d] == CXP_BC_PERIODIC;
    period[d] = crse->domain_hi[d] - crse->domain_lo[d] + op->d.indextype[d];
  }
  IBox fdom = ibox(fine->domain_lo, fine->domain_hi);
  for (int d = 0; d < 3; ++d)
    if (mode != FillMode::whole || periodic[d]) {
      fdom.lo[d] -= g[d];
      fdom.hi[d] += g[d];
    }
  const IBox cdom = ibox(crse->domain_lo, crse->domain_hi);
  auto cvalid = [&](int j) { return ibox(crse->valid + 6 * j, crse->valid + 6 * j + 3); };
  auto cfabbox = [&](int j) { return ibox(crse->fabs[j].lo, crse->fabs[j].hi); };

  std::vector<cxp_worklist_s::Patch> patches;
  std::vector<GatherRun> runs;
  std::vector<GatherChunk> chunks;
  std::vector<BcRun> bruns;
  std::vector<BcChunk> bchunks, bdchunks;
  std::vector<BcDirect> bdirs;
  int64_t gather = wl->gather_doubles;
  int64_t fine_pts = 0, coarse_pts = 0;
  const int run0 = (int)wl->gather_runs.size();
  const int brun0 = (int)wl->bc_runs.size();

  // one copy of region X of coarse FAB j into the coarse patch C at offset
  // off, every component
  // (with shift sh, into X + sh: a periodic image)
  auto add_copy = [&](int j, const IBox &X, const IBox &C, int64_t off,
                      const int *sh = nullptr) {
    const int64_t cvol = C.volume();
    const int d0 = sh ? sh[0] : 0, d1 = sh ? sh[1] : 0, d2 = sh ? sh[2] : 0;
    const int nx = X.hi[0] - X.lo[0] + 1, ny = X.hi[1] - X.lo[1] + 1;
    const int nrows = ny * (X.hi[2] - X.lo[2] + 1);
    const int per = std::max(1, kGatherChunkPts / nx);
    const int cx = C.hi[0] - C.lo[0] + 1, cy = C.hi[1] - C.lo[1] + 1;
    for (int c = 0; c < ncomp; ++c) {
      const DFab f = dfab(&crse->fabs[j], crse_comp + c);
      GatherRun r;
      r.src = f.p + didx(f, X.lo[0], X.lo[1], X.lo[2]);
      r.src_sy = f.sy;
      r.src_sz = f.sz;
      r.dstp = nullptr;
      r.dst = off + c * cvol + (X.lo[0] + d0 - C.lo[0]) +
              (long long)(X.lo[1] + d1 - C.lo[1]) * cx +
              (long long)(X.lo[2] + d2 - C.lo[2]) * cx * cy;
      r.dst_sy = cx;
      r.dst_sz = (long long)cx * cy;
      r.nx = nx;
      r.ny = ny;
      const int run = run0 + (int)runs.size();
      runs.push_back(r);
      for (int row = 0; row < nrows; row += per)
        chunks.push_back({run, row, std::min(per, nrows - row)});
    }
  };
  // boundary values of the coarse patch E at offset off, every component
  auto add_bc = [&](const IBox &E, int64_t off) {
    const DirectSrc direct{crse, crse_comp, periodic, period, &bdchunks, &bdirs,
                           (int)wl->bc_direct.size()};
    add_bc_runs(*bc, op->d.indextype, E, cdom, nullptr, off, E.volume(), ncomp,
                brun0, bruns, bchunks, &direct);
  };
  // the fine level's periodic shifts (fine index units), the unshifted one
  // first: a valid box also covers its periodic images (the reference's
  // copies and syncs use the geometry's periodicity)
  std::vector<std::array<int, 3>> fshifts;
  for (int s0 = 0; s0 < 27; ++s0) {
    const int s = (s0 + 13) % 27;
    const int sh[3] = {s % 3 - 1, (s / 3) % 3 - 1, s / 9 - 1};
    if ((sh[0] && !periodic[0]) || (sh[1] && !periodic[1]) ||
        (sh[2] && !periodic[2]))
      continue;
    std::array<int, 3> v;
    for (int d = 0; d < 3; ++d)
      v[d] = sh[d] * (fine->domain_hi[d] - fine->domain_lo[d] +
                      op->d.indextype[d]);
    fshifts.push_back(v);
  }
  auto shifted = [](IBox b, const std::array<int, 3> &v) {
    for (int d = 0; d < 3; ++d) {
      b.lo[d] += v[d];
      b.hi[d] += v[d];
    }
    return b;
  };
  // copies of the old fine level's valid data (and its periodic images)
  // into fine FAB i (remake); each point from the first image covering it
  int64_t copied = 0;
  auto add_fab_copies = [&](int i) {
    const IBox B = ibox(fine->fabs[i].lo, fine->fabs[i].hi);
    std::vector<IBox> covered;
    for (const auto &v : fshifts)
      for (int j = 0; j < old->nfabs; ++j) {
        const IBox X = intersect(
            B, shifted(ibox(old->valid + 6 * j, old->valid + 6 * j + 3), v));
        if (X.empty())
          continue;
        std::vector<IBox> parts{X};
        for (const IBox &cv : covered)
          subtract_all(parts, cv);
        covered.push_back(X);
        for (const IBox &P : parts) {
          copied += P.volume();
          const int nx = P.hi[0] - P.lo[0] + 1, ny = P.hi[1] - P.lo[1] + 1;
          const int nrows = ny * (P.hi[2] - P.lo[2] + 1);
          const int per = std::max(1, kGatherChunkPts / nx);
          for (int c = 0; c < ncomp; ++c) {
            const DFab f = dfab(&old->fabs[j], fine_comp + c);
            const DFab t = dfab(&fine->fabs[i], fine_comp + c);
            GatherRun r;
            r.src = f.p + didx(f, P.lo[0] - v[0], P.lo[1] - v[1], P.lo[2] - v[2]);
            r.src_sy = f.sy;
            r.src_sz = f.sz;
            r.dstp = t.p;
            r.dst = didx(t, P.lo[0], P.lo[1], P.lo[2]);
            r.dst_sy = t.sy;
            r.dst_sz = t.sz;
            r.nx = nx;
            r.ny = ny;
            const int run = run0 + (int)runs.size();
            runs.push_back(r);
            for (int row = 0; row < nrows; row += per)
              chunks.push_back({run, row, std::min(per, nrows - row)});
          }
        }
      }
  };

  // Gathers coarse box C at offset `gather` of the patch buffer, into box
  // E: inside the domain from the valid regions (and in periodic
  // directions from the images one period away); outside it, with
  // boundary conditions, the boundary values (E is then C and their
  // sources), without, the first FAB holding each point (E = C).  On
  // failure nothing is added.
  auto gather_box = [&](const IBox &C, IBox &E) -> int {
    const size_t n_runs = runs.size(), n_chunks = chunks.size();
    const size_t n_bruns = bruns.size(), n_bchunks = bchunks.size();
    const size_t n_bdchunks = bdchunks.size(), n_bdirs = bdirs.size();
    auto fail = [&](int e) {
      runs.resize(n_runs);
      chunks.resize(n_chunks);
      bruns.resize(n_bruns);
      bchunks.resize(n_bchunks);
      bdchunks.resize(n_bdchunks);
      bdirs.resize(n_bdirs);
      return e;
    };
    const bool use_bc = bc && !inside(C, cdom);
    E = C;
    if (use_bc)
      if (const int e = bc_closure(*bc, op->d, cdom, C, E))
        return fail(e);
    if (E.volume() > (int64_t)INT32_MAX)
      return fail(CXP_ERR_ARG);
    const int64_t off = gather;
    std::vector<IBox> left;
    // The unshifted part goes first and each image takes only points no
    // earlier one covered (with vertex centring the domain's first and last
    // points are images of each other; the point inside is kept).
    std::vector<IBox> covered;
    for (int s0 = 0; s0 < 27; ++s0) {
      const int s = (s0 + 13) % 27;  // (0, 0, 0) first
      const int sh[3] = {s % 3 - 1, (s / 3) % 3 - 1, s / 9 - 1};
      if ((sh[0] && !periodic[0]) || (sh[1] && !periodic[1]) ||
          (sh[2] && !periodic[2]))
        continue;
      IBox img = cdom;  // the domain moved by the shift
      int dsh[3];
      for (int d = 0; d < 3; ++d) {
        dsh[d] = sh[d] * period[d];
        img.lo[d] += dsh[d];
        img.hi[d] += dsh[d];
      }
      const IBox part = intersect(E, img);
      if (part.empty())
        continue;
      std::vector<IBox> pieces{part};
      for (const IBox &cv : covered)
        subtract_all(pieces, cv);
      covered.push_back(part);
      for (IBox piece : pieces) {
        for (int d = 0; d < 3; ++d) {  // its source, inside the domain
          piece.lo[d] -= dsh[d];
          piece.hi[d] -= dsh[d];
        }
        left.assign(1, piece);
        for (int j = 0; j < crse->nfabs; ++j) {
          const IBox X = intersect(piece, cvalid(j));
          if (!X.empty()) {
            add_copy(j, X, E, off, dsh);
            subtract_all(left, cvalid(j));
          }
        }
        if (!left.empty())
          return fail(CXP_ERR_COARSE);
      }
    }
    left.clear();
    if (use_bc) {
      add_bc(E, off);
    } else {
      subtract(E, cdom, left);
      for (int j = 0; j < crse->nfabs && !left.empty(); ++j) {
        std::vector<IBox> rest;
        for (const IBox &r : left) {
          const IBox X = intersect(r, cfabbox(j));
          if (!X.empty())
            add_copy(j, X, E, off);
          subtract(r, cfabbox(j), rest);
        }
        left.swap(rest);
      }
      if (!left.empty())
        return fail(CXP_ERR_COARSE);
    }
    return CXP_OK;
  };

  for (int i = 0; i < fine->nfabs; ++i) {
    const cxp_fab &ff = fine->fabs[i];
    const IBox fv = ibox(fine->valid + 6 * i, fine->valid + 6 * i + 3);
    std::vector<IBox> pieces;
    if (mode == FillMode::ghosts) {
      subtract(intersect(ibox(ff.lo, ff.hi), fdom), fv, pieces);
      for (const auto &v : fshifts)
        for (int j = 0; j < fine->nfabs; ++j)
          if (j != i || v[0] || v[1] || v[2])
            subtract_all(pieces, shifted(ibox(fine->valid + 6 * j,
                                              fine->valid + 6 * j + 3), v));
    } else {
      const IBox B = intersect(ibox(ff.lo, ff.hi), fdom);
      if (!B.empty())
        pieces.push_back(B);
      if (mode == FillMode::remake) {
        for (const auto &v : fshifts)
          for (int j = 0; j < old->nfabs; ++j)
            subtract_all(pieces, shifted(ibox(old->valid + 6 * j,
                                              old->valid + 6 * j + 3), v));
        add_fab_copies(i);
      }
    }
    // patches of this FAB: in place where one coarse FAB holds the coarse
    // box (inside the coarse domain within its valid region, and inside
    // the domain when boundary conditions apply); the others share one
    // gathered box, the hull of their coarse boxes (their boxes overlap at
    // edges and corners), or, if that hull cannot be filled, one each
    std::vector<IBox> cboxes;
    std::vector<cxp_worklist_s::Patch> fp;
    std::vector<size_t> gathered;
    for (const IBox &P : pieces) {
      IBox C;
      op_coarse_box(op->d, P.lo, P.hi, C.lo, C.hi);
      cboxes.push_back(C);
      fine_pts += P.volume();
      cxp_worklist_s::Patch p;
      p.op = op;
      p.fine = ff;
      p.fine_comp = fine_comp;
      p.ncomp = ncomp;
      for (int d = 0; d < 3; ++d) {
        p.rlo[d] = P.lo[d];
        p.rhi[d] = P.hi[d];
      }
      const bool bcp = bc && !inside(C, cdom);
      int src = -1;
      for (int j = 0; j < crse->nfabs && src < 0 && !bcp; ++j)
        if (inside(C, cfabbox(j)) &&
            (intersect(C, cdom).empty() || inside(intersect(C, cdom), cvalid(j))))
          src = j;
      if (src >= 0) {
        p.crse = crse->fabs[src];
        p.crse_comp = crse_comp;
      } else {
        gathered.push_back(fp.size());
      }
      fp.push_back(p);
    }
    auto point_at = [&](cxp_worklist_s::Patch &p, const IBox &E, int64_t off) {
      p.crse.data = nullptr;  // placed at upload
      for (int d = 0; d < 3; ++d) {
        p.crse.lo[d] = E.lo[d];
        p.crse.hi[d] = E.hi[d];
      }
      p.crse.comp_stride = E.volume();
      p.crse.ncomp = ncomp;
      p.crse_comp = 0;
      p.gather_off = off;
    };
    if (!gathered.empty()) {
      IBox H = cboxes[gathered[0]];
      int64_t separate = 0;
      for (size_t k : gathered) {
        separate += cboxes[k].volume();
        for (int d = 0; d < 3; ++d) {
          H.lo[d] = std::min(H.lo[d], cboxes[k].lo[d]);
          H.hi[d] = std::max(H.hi[d], cboxes[k].hi[d]);
        }
      }
      // shared only where it gathers no more than the separate boxes
      IBox E;
      if (gathered.size() > 1 && H.volume() <= separate &&
          gather_box(H, E) == CXP_OK) {
        for (size_t k : gathered)
          point_at(fp[k], E, gather);
        gather += E.volume() * ncomp;
      } else {
        for (size_t k : gathered) {
          if (const int e = gather_box(cboxes[k], E))
            return e;
          point_at(fp[k], E, gather);
          gather += E.volume() * ncomp;
        }
      }
    }
    patches.insert(patches.end(), fp.begin(), fp.end());
    coarse_pts += union_points(cboxes);
  }
  if (!gather_chunks_ok(chunks, run0, runs.size()) ||
      !bc_chunks_ok(bchunks, brun0, bruns) ||
      !bc_chunks_ok(bdchunks, brun0, bruns))
    return CXP_ERR_STATE;
  for (const BcChunk &c : bdchunks)
    if (c.direct < (int)wl->bc_direct.size() ||
        c.direct >= (int)(wl->bc_direct.size() + bdirs.size()))
      return CXP_ERR_STATE;
  for (const BcChunk &c : bchunks)
    if (c.direct != -1)
      return CXP_ERR_STATE;
  wl->patches.insert(wl->patches.end(), patches.begin(), patches.end());
  wl->gather_runs.insert(wl->gather_runs.end(), runs.begin(), runs.end());
  wl->gather_chunks.insert(wl->gather_chunks.end(), chunks.begin(), chunks.end());
  wl->bc_runs.insert(wl->bc_runs.end(), bruns.begin(), bruns.end());
  wl->bc_chunks.insert(wl->bc_chunks.end(), bchunks.begin(), bchunks.end());
  wl->bcd_chunks.insert(wl->bcd_chunks.end(), bdchunks.begin(), bdchunks.end());
  wl->bc_direct.insert(wl->bc_direct.end(), bdirs.begin(), bdirs.end());
  wl->gather_doubles = gather;
  wl->uploaded = false;
  touch(wl);
  wl->fine_points += fine_pts * ncomp;
  wl->algo_bytes += 8 * (int64_t)ncomp * (fine_pts + coarse_pts + 2 * copied);
  return CXP_OK;
}

} // namespace

int cxp_worklist_append_level(cxp_worklist wl, cxp_op op,
                              const cxp_level *crse, int crse_comp,
                              const cxp_level *fine, int fine_comp,
                              int ncomp) {
  return append_level_impl(wl, op, crse, crse_comp, fine, fine_comp, ncomp,
                           nullptr);
}

int cxp_worklist_append_level_bc(cxp_worklist wl, cxp_op op,
                                 const cxp_level *crse, int crse_comp,
                                 const cxp_level *fine, int fine_comp,
                                 int ncomp, const cxp_bc *bc) {
  return append_level_impl(wl, op, crse, crse_comp, fine, fine_comp, ncomp,
                           bc);
}

int cxp_worklist_append_new_level(cxp_worklist wl, cxp_op op,
                                  const cxp_level *crse, int crse_comp,
                                  const cxp_level *fine, int fine_comp,
                                  int ncomp, const cxp_bc *bc) {
  return append_level_impl(wl, op, crse, crse_comp, fine, fine_comp, ncomp,
                           bc, FillMode::whole);
}

int cxp_worklist_append_remake_level(cxp_worklist wl, cxp_op op,
                                     const cxp_level *crse, int crse_comp,
                                     const cxp_level *old_fine,
                                     const cxp_level *fine, int fine_comp,
                                     int ncomp, const cxp_bc *bc) {
  if (!old_fine)
    return CXP_ERR_ARG;
  return append_level_impl(wl, op, crse, crse_comp, fine, fine_comp, ncomp,
                           bc, FillMode::remake, old_fine);
}

int cxp_worklist_append_boundaries(cxp_worklist wl, const cxp_level *level,
                                   const int indextype[3], int comp,
                                   int ncomp, const cxp_bc *bc) {
  if (!wl || !level_ok(level) || !indextype || !bc || ncomp < 1 ||
      !bc_ok(*bc, ncomp))
    return CXP_ERR_ARG;
  for (int d = 0; d < 3; ++d)
    if (indextype[d] != 0 && indextype[d] != 1)
      return CXP_ERR_ARG;
  for (int i = 0; i < level->nfabs; ++i) {
    if (comp < 0 || comp + ncomp > level->fabs[i].ncomp)
      return CXP_ERR_COMP;
    if (fab_component_bytes(&level->fabs[i]) > kMaxComponentBytes)
      return CXP_ERR_ARG;
  }
  const IBox dom = ibox(level->domain_lo, level->domain_hi);
  std::vector<BcRun> runs;
  std::vector<BcChunk> chunks;
  for (int i = 0; i < level->nfabs; ++i) {
    const cxp_fab &f = level->fabs[i];
    const IBox B = ibox(f.lo, f.hi);
    if (inside(B, dom))
      continue;
    // every source of the FAB's boundary points must lie in the FAB, as
    // the reference asserts (boundaries_impl.hxx:249-298); faces without a
    // condition are left alone
    cxp_bc only = *bc;
    for (int side = 0; side < 2; ++side)
      for (int d = 0; d < 3; ++d)
        if (only.type[side][d] == CXP_BC_NONE ||
            only.type[side][d] == CXP_BC_PERIODIC)
          only.type[side][d] = CXP_BC_DIRICHLET;  // reads nothing
    OpDesc od;
    for (int d = 0; d < 3; ++d)
      od.indextype[d] = indextype[d];
    IBox E;
    if (bc_closure(only, od, dom, B, E) || !inside(E, B))
      return CXP_ERR_ARG;
    const DFab d0 = dfab(&f, comp);
    add_bc_runs(*bc, indextype, B, dom, d0.p, 0, (int64_t)f.comp_stride, ncomp,
                (int)wl->post_runs.size(), runs, chunks);
  }
  if (!bc_chunks_ok(chunks, (int)wl->post_runs.size(), runs))
    return CXP_ERR_STATE;
  wl->post_runs.insert(wl->post_runs.end(), runs.begin(), runs.end());
  wl->post_chunks.insert(wl->post_chunks.end(), chunks.begin(), chunks.end());
  wl->uploaded = false;
  touch(wl);
  return CXP_OK;
}

int cxp_worklist_set_per_group(cxp_worklist wl, int per_group) {
  if (!wl || (per_group != 0 && per_group != 1))
    return CXP_ERR_ARG;
  wl->per_group = per_group;
  wl->uploaded = false;
  touch(wl);
  return CXP_OK;
}

int cxp_worklist_upload(cxp_worklist wl, void *stream) {
  if (!wl)
    return CXP_ERR_ARG;
  // per-group decisions cover at most 16 components (impl.hxx:1707-1708)
  if (wl->per_group)
    for (const auto &p : wl->patches)
      if (p.ncomp > 16 && !op_is_separable(p.op->d, true))
        return CXP_ERR_NCOMP;
  touch(wl);
  // a pending launch may still read the buffers about to be rewritten
  if (const int e = list_drain(wl))
    return e;
  if (const int e = place_gathered(wl))
    return e;
  build_tiles(wl);
  const int e = upload_lists(wl, (hipStream_t)stream);
  if (e)
    return e;
  wl->uploaded = true;
  return CXP_OK;
}

int cxp_worklist_stats(cxp_worklist wl, int64_t *npatches, int64_t *ntiles,
                       int64_t *fine_points, int64_t *algo_bytes) {
  if (!wl)
    return CXP_ERR_ARG;
  if (npatches)
    *npatches = (int64_t)wl->patches.size();
  if (ntiles)
    *ntiles = (int64_t)wl->tiles.size();
  if (fine_points)
    *fine_points = wl->fine_points;
  if (algo_bytes)
    *algo_bytes = wl->algo_bytes;
  return CXP_OK;
}

int cxp_worklist_poison(cxp_worklist wl, double value, void *stream) {
  if (!wl)
    return CXP_ERR_ARG;
  if (!wl->uploaded)
    return CXP_ERR_STATE;
  if (wl->poison_chunks.empty())
    return CXP_OK;
  hipStream_t s = (hipStream_t)stream;
  return list_launch(wl, s, [&] {
    hipLaunchKernelGGL(k_poison, dim3((unsigned)wl->poison_chunks.size()),
                       dim3(256), 0, s, wl->d_poison_runs,
                       wl->d_poison_chunks, value);
    return hip_status(hipGetLastError());
  });
}

int cxp_prolongate_batched(cxp_worklist wl, void *stream) {
  if (!wl)
    return CXP_ERR_ARG;
  if (!wl->uploaded)
    return CXP_ERR_STATE;
  hipStream_t s = (hipStream_t)stream;
  return list_launch(wl, s, [&] {
    if (const int e = launch_gather(wl, s))
      return e;
    if (const int e = launch_batched(wl, s))
      return e;
    return launch_post_bc(wl, s);
  });
}

int cxp_prolongate_per_patch(cxp_worklist wl, void *stream) {
  if (!wl)
    return CXP_ERR_ARG;
  const bool staged = !wl->gather_runs.empty() || !wl->post_runs.empty() ||
                      !wl->bc_runs.empty();
  // gathers, copies and boundary runs live on the device only after an
  // upload
  if (staged && !wl->uploaded)
    return CXP_ERR_STATE;
  hipStream_t s = (hipStream_t)stream;
  return list_launch(wl, s, [&] {
    if (staged)
      if (const int e = launch_gather(wl, s))
        return e;
    for (const auto &p : wl->patches) {
      const int e = prolongate_impl(p.op, &p.crse, p.crse_comp, &p.fine,
                                    p.fine_comp, p.ncomp, p.rlo, p.rhi,
                                    wl->per_group, s);
      if (e)
        return e;
    }
    return launch_post_bc(wl, s);
  });
}

// ---- graphs

int cxp_graph_capture_begin(void *stream) {
  if (!stream)
    return CXP_ERR_ARG;  // the legacy default stream cannot be captured
  return hip_status(hipStreamBeginCapture((hipStream_t)stream,
                                          hipStreamCaptureModeThreadLocal));
}

int cxp_graph_capture_end(void *stream, cxp_graph *out) {
  if (!stream || !out)
    return CXP_ERR_ARG;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  unsigned long long cid = 0;
  if (hipStreamGetCaptureInfo((hipStream_t)stream, &cs, &cid) != hipSuccess)
    return CXP_ERR_HIP;
  std::vector<std::pair<uint64_t, uint64_t>> lists;
  {
    std::lock_guard<std::mutex> lk(g_lists_mu);
    auto it = g_capture_lists.find(cid);
    if (it != g_capture_lists.end()) {
      lists = std::move(it->second);
      g_capture_lists.erase(it);
    }
  }
  std::sort(lists.begin(), lists.end());
  lists.erase(std::unique(lists.begin(), lists.end()), lists.end());
  hipGraph_t g;
  if (hipStreamEndCapture((hipStream_t)stream, &g) != hipSuccess)
    return CXP_ERR_HIP;
  hipGraphExec_t ex;
  if (hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGraphDestroy(g);
    return CXP_ERR_HIP;
  }
  *out = new cxp_graph_s{g, ex, std::move(lists)};
  return CXP_OK;
}

int cxp_graph_launch(cxp_graph g, void *stream) {
  if (!g)
    return CXP_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  // every captured list must be alive and unchanged since the capture
  std::vector<cxp_worklist_s *> wls;
  {
    std::lock_guard<std::mutex> lk(g_lists_mu);
    for (const auto &r : g->lists) {
      auto it = g_lists.find(r.first);
      if (it == g_lists.end() || it->second->epoch != r.second)
        return CXP_ERR_STATE;
      wls.push_back(it->second);
    }
  }
  std::vector<std::unique_lock<std::mutex>> locks;
  for (cxp_worklist_s *wl : wls) {
    locks.emplace_back(wl->mu);
    if (const int e = list_order_after_last(wl, s))
      return e;
  }
  if (hipGraphLaunch(g->exec, s) != hipSuccess)
    return CXP_ERR_HIP;
  for (cxp_worklist_s *wl : wls)
    if (const int e = list_leave(wl, s))
      return e;
  return CXP_OK;
}

int cxp_graph_destroy(cxp_graph g) {
  if (!g)
    return CXP_ERR_ARG;
  (void)hipGraphExecDestroy(g->exec);
  (void)hipGraphDestroy(g->graph);
  delete g;
  return CXP_OK;
}

// ---- autotuner

int cxp_tile_config_count(void) { return kNumTileConfigs; }

int cxp_tile_config(int cfg, int *tx, int *ty, int *zwindow) {
  if (cfg < 0 || cfg >= kNumTileConfigs)
    return CXP_ERR_ARG;
  if (tx)
    *tx = kCfgs[cfg].tx;
  if (ty)
    *ty = kCfgs[cfg].ty;
  if (zwindow)
    *zwindow = CXP_ZW;
  return CXP_OK;
}

// In-context tuning.  A tile shape and z-chunk that are fastest for one
// class run alone need not be fastest inside the whole list, where the
// persistent kernel overlaps every class's tiles; so every candidate is
// timed on the whole list.  Size classes are tuned one after another
// (coordinate descent): each class takes, for all its stencil triples, the
// (shape, z-chunk) that gives the shortest whole-list time, if that beats
// the best so far by more than kTuneGain; otherwise it keeps its setting.
// Passes over the classes repeat until one changes nothing.
constexpr double kTuneGain = 0.005;
constexpr int kTunePasses = 3;

int cxp_autotune(cxp_worklist wl, int reps, void *stream) {
  if (!wl || reps < 1)
    return CXP_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  touch(wl);
  if (const int e = list_drain(wl))
    return e;
  if (int e = place_gathered(wl))
    return e;
  // separable, non-thin keys of the list, grouped by size class
  std::map<int, std::vector<TuneKey>> classes;
  for (const auto &p : wl->patches) {
    if (!op_is_separable(p.op->d, wl->per_group != 0))
      continue;
    const SepTask t = make_sep_task(p.op->d, &p.crse, p.crse_comp, &p.fine,
                                    p.fine_comp, p.rlo, p.rhi);
    if (thin_eligible(t))
      continue;
    const TuneKey k = tune_key(t);
    auto &ks = classes[std::get<3>(k)];
    if (std::find(ks.begin(), ks.end(), k) == ks.end())
      ks.push_back(k);
  }
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return CXP_ERR_HIP;
  int status = CXP_OK;
  // median of reps whole-list launches (after one warm-up) with the table
  // as it stands
  auto time_list = [&](float &ms) {
    build_tiles(wl);
    if ((status = upload_lists(wl, s)))
      return false;
    if (wl->tiles.empty()) {
      ms = 0;
      return true;
    }
    std::vector<float> t;
    for (int r = 0; r <= reps; ++r) {
      (void)hipEventRecord(e0, s);
      if ((status = launch_sep_list(wl, s)))
        return false;
      (void)hipEventRecord(e1, s);
      if (hipEventSynchronize(e1) != hipSuccess) {
        status = CXP_ERR_HIP;
        return false;
      }
      float x = 0;
      (void)hipEventElapsedTime(&x, e0, e1);
      if (r > 0)
        t.push_back(x);
    }
    std::sort(t.begin(), t.end());
    ms = t[t.size() / 2];
    return true;
  };
  auto set_class = [&](const std::vector<TuneKey> &ks, int cfg, int zc) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (const TuneKey &k : ks) {
      if (cfg < 0)
        g_tune.erase(k);
      else
        g_tune[k] = {cfg, zc};
    }
  };
  // start from the defaults for the list's classes
  for (const auto &cl : classes)
    set_class(cl.second, -1, 0);
  float best = 0;
  if (!time_list(best))
    classes.clear();
  const int zchunks[4] = {0, 16, 32, 64};  // 0: the list-wide default
  std::map<int, std::pair<int, int>> choice;  // class -> (cfg or -1, zchunk)
  for (const auto &cl : classes)
    choice[cl.first] = {-1, 0};
  // passes until one changes nothing (at most kTunePasses)
  for (int pass = 0; pass < kTunePasses && !status; ++pass) {
    bool changed = false;
    for (const auto &cl : classes) {
      std::pair<int, int> &ch = choice[cl.first];
      for (int cfg = 0; cfg < kNumTileConfigs && !status; ++cfg)
        for (int zi = 0; zi < 4 && !status; ++zi) {
          if (cfg == ch.first && zchunks[zi] == ch.second)
            continue;
          if (kCfgs[cfg].rp != 1 && wl->uniform_sid < 0)
            continue;  // a mixed list runs one-pair shapes only
          set_class(cl.second, cfg, zchunks[zi]);
          float ms = 0;
          if (!time_list(ms))
            break;
          if (ms < best * (1.0 - kTuneGain)) {
            best = ms;
            ch = {cfg, zchunks[zi]};
            changed = true;
          }
        }
      set_class(cl.second, ch.first, ch.second);
      if (status)
        break;
    }
    if (!changed)
      break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  wl->uploaded = false;
  touch(wl);
  if (status)
    return status;
  return cxp_worklist_upload(wl, stream);
}

int cxp_tune_save(const char *path) {
  if (!path)
    return CXP_ERR_ARG;
  FILE *f = std::fopen(path, "w");
  if (!f)
    return CXP_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  std::fprintf(f, "# sid_x sid_y sid_z size_class cfg zchunk\n");
  for (const auto &kv : g_tune)
    std::fprintf(f, "%d %d %d %d %d %d\n", std::get<0>(kv.first),
                 std::get<1>(kv.first), std::get<2>(kv.first),
                 std::get<3>(kv.first), kv.second.first, kv.second.second);
  std::fclose(f);
  return CXP_OK;
}

int cxp_tune_set(cxp_op op, const int region_lo[3], const int region_hi[3],
                 int cfg, int zchunk) {
  if (!op || !region_lo || !region_hi || cfg < 0 || cfg >= kNumTileConfigs ||
      zchunk < 0 || zchunk == 1)
    return CXP_ERR_ARG;
  const TuneKey k(op->d.sid[0], op->d.sid[1], op->d.sid[2],
                  size_class(region_lo, region_hi));
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune[k] = {cfg, zchunk};
  return CXP_OK;
}

int cxp_tune_clear(void) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune.clear();
  return CXP_OK;
}

int cxp_tune_load(const char *path) {
  if (!path)
    return CXP_ERR_ARG;
  FILE *f = std::fopen(path, "r");
  if (!f)
    return CXP_ERR_ARG;
  char line[256];
  std::lock_guard<std::mutex> lk(g_tune_mu);
  while (std::fgets(line, sizeof line, f)) {
    if (line[0] == '#')
      continue;
    int a, b, c, d, cfg, zc;
    if (std::sscanf(line, "%d %d %d %d %d %d", &a, &b, &c, &d, &cfg, &zc) == 6 &&
        cfg >= 0 && cfg < kNumTileConfigs && zc >= 2)
      g_tune[TuneKey(a, b, c, d)] = {cfg, zc};
  }
  std::fclose(f);
  return CXP_OK;
}

} // extern "C"
