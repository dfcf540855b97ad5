// Reflux (CarpetX/src/schedule.cxx:2632-2716) on the device: the flux
// register correction of the coarse level beside restriction.
//
// The reference calls, per direction d, AMReX's FluxRegister::CrseInit
// (coarse flux, mult -1) and FineAdd (fine flux, mult +1), then Reflux
// (coarse state, scale 1, coarse geometry) (:2694-2704), on a register
// made for the fine level's cell-centred boxes, ratio 2 (driver.cxx:918-924).
// Its arithmetic, restated here for each register face F (a coarse face on
// the low or high side of a coarsened fine box):
//   reg = crse_mult * Fc(F);  reg += fine_mult * Ff(sub) for its 4 fine
//   sub-faces (the higher transverse offset outer);
//   the coarse cell beyond F (index - 1 for a low face, the index for a
//   high face, also across a periodic boundary) gets -scale * reg / V or
//   +scale * reg / V, V = dx[0] dx[1] dx[2], orientation by orientation in
//   the order x-lo, y-lo, z-lo, x-hi, y-hi, z-hi.
//
// The host splits each (fine box, orientation) into rectangles of faces
// that share one coarse-flux FAB and one target FAB (pieces); a launch per
// orientation runs a thread per face over all its pieces.  Within one
// orientation no two faces reach the same coarse cell, so a launch has no
// write conflicts, and the orientation order of the sums is kept.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "../../include/cxp.h"

namespace {

struct RfFab {
  double *p;  // component 0 at the FAB's lo
  int lo[3];
  long long sy, sz, cs;  // row, plane and component strides
};

struct RfPiece {
  RfFab cflux, fflux, state;
  int lo[3], hi[3];  // register faces: coarse face index in dir, cells across
  int dir, side, shift;
};

struct RfChunk {
  int piece, start;  // 256 faces of one piece from `start`
};

constexpr int kRfThreads = 256;

__host__ __device__ inline long long rf_idx(const RfFab &f, int i, int j,
                                            int k) {
  return (long long)(i - f.lo[0]) + (j - f.lo[1]) * f.sy +
         (k - f.lo[2]) * f.sz;
}

__global__ void __launch_bounds__(kRfThreads)
    k_reflux(const RfPiece *__restrict__ pieces,
             const RfChunk *__restrict__ chunks, int comp, int ncomp,
             double cm, double fm, double scale, double vol) {
  const RfChunk ch = chunks[blockIdx.x];
  const RfPiece &p = pieces[ch.piece];
  const int d = p.dir;
  const int t0 = d == 0 ? 1 : 0, t1 = d == 2 ? 1 : 2;
  const int n0 = p.hi[t0] - p.lo[t0] + 1, n1 = p.hi[t1] - p.lo[t1] + 1;
  const int f = ch.start + (int)threadIdx.x;
  if (f >= n0 * n1)
    return;
  int face[3];
  face[d] = p.lo[d];
  face[t0] = p.lo[t0] + f % n0;
  face[t1] = p.lo[t1] + f / n0;
  int cell[3] = {face[0], face[1], face[2]};
  cell[d] += (p.side ? 0 : -1) + p.shift;
  const long long ci = rf_idx(p.cflux, face[0], face[1], face[2]);
  const long long si = rf_idx(p.state, cell[0], cell[1], cell[2]);
  int sub[3] = {2 * face[0], 2 * face[1], 2 * face[2]};
  for (int n = comp; n < comp + ncomp; ++n) {
    double reg = cm * p.cflux.p[n * p.cflux.cs + ci];
    for (int o1 = 0; o1 < 2; ++o1)
      for (int o0 = 0; o0 < 2; ++o0) {
        int q[3] = {sub[0], sub[1], sub[2]};
        q[t0] += o0;
        q[t1] += o1;
        reg = reg + fm * p.fflux.p[n * p.fflux.cs + rf_idx(p.fflux, q[0], q[1], q[2])];
      }
    const double v = p.side ? (scale * reg) / vol : (-scale * reg) / vol;
    p.state.p[n * p.state.cs + si] += v;
  }
}

struct Rect {
  int lo[3], hi[3];
};

bool rect_empty(const Rect &r) {
  for (int d = 0; d < 3; ++d)
    if (r.hi[d] < r.lo[d])
      return true;
  return false;
}

Rect rect_and(const Rect &a, const Rect &b) {
  Rect r;
  for (int d = 0; d < 3; ++d) {
    r.lo[d] = std::max(a.lo[d], b.lo[d]);
    r.hi[d] = std::min(a.hi[d], b.hi[d]);
  }
  return r;
}

// a minus b, as disjoint rectangles appended to out.
void rect_minus(const Rect &a, const Rect &b, std::vector<Rect> &out) {
  const Rect x = rect_and(a, b);
  if (rect_empty(x)) {
    out.push_back(a);
    return;
  }
  Rect rest = a;
  for (int d = 0; d < 3; ++d) {
    if (rest.lo[d] < x.lo[d]) {
      Rect r = rest;
      r.hi[d] = x.lo[d] - 1;
      out.push_back(r);
      rest.lo[d] = x.lo[d];
    }
    if (rest.hi[d] > x.hi[d]) {
      Rect r = rest;
      r.lo[d] = x.hi[d] + 1;
      out.push_back(r);
      rest.hi[d] = x.hi[d];
    }
  }
}

bool fab_holds(const cxp_fab &f, const Rect &r, int comp, int ncomp) {
  if (!f.data || f.ncomp < comp + ncomp || comp < 0 || ncomp < 1)
    return false;
  for (int d = 0; d < 3; ++d)
    if (r.lo[d] < f.lo[d] || r.hi[d] > f.hi[d])
      return false;
  const long long vol = (long long)(f.hi[0] - f.lo[0] + 1) *
                        (f.hi[1] - f.lo[1] + 1) * (f.hi[2] - f.lo[2] + 1);
  return f.ncomp == 1 || f.comp_stride >= vol;
}

RfFab rf_fab(const cxp_fab &f) {
  RfFab v;
  v.p = f.data;
  for (int d = 0; d < 3; ++d)
    v.lo[d] = f.lo[d];
  v.sy = f.hi[0] - f.lo[0] + 1;
  v.sz = v.sy * (f.hi[1] - f.lo[1] + 1);
  v.cs = f.comp_stride;
  return v;
}

Rect valid_rect(const cxp_level *L, int i) {
  Rect r;
  for (int d = 0; d < 3; ++d) {
    r.lo[d] = L->valid[6 * i + d];
    r.hi[d] = L->valid[6 * i + 3 + d];
  }
  return r;
}

bool level_shape_ok(const cxp_level *L) {
  if (!L || L->nfabs < 1 || !L->fabs || !L->valid)
    return false;
  for (int d = 0; d < 3; ++d)
    if (L->domain_hi[d] < L->domain_lo[d])
      return false;
  for (int i = 0; i < L->nfabs; ++i)
    if (rect_empty(valid_rect(L, i)))
      return false;
  return true;
}

} // namespace

struct cxp_reflux_s {
  std::vector<RfPiece> pieces;
  std::vector<RfChunk> chunks;
  int chunk_off[7];  // chunks of orientation o: [chunk_off[o], chunk_off[o+1])
  RfPiece *d_pieces = nullptr;
  RfChunk *d_chunks = nullptr;
  int comp = 0, ncomp = 1;
  double cm = -1, fm = 1, scale = 1, vol = 1;
};

extern "C" {

int cxp_reflux_plan_create(const cxp_level *crse,
                           const cxp_fab *const crse_flux[3],
                           const cxp_level *fine,
                           const cxp_fab *const fine_flux[3], int comp,
                           int ncomp, const double dx[3],
                           const int periodic[3], double crse_mult,
                           double fine_mult, double scale,
                           cxp_reflux_plan *out) {
  if (!out || !crse_flux || !fine_flux || !dx || !periodic ||
      !level_shape_ok(crse) || !level_shape_ok(fine) || comp < 0 ||
      ncomp < 1)
    return CXP_ERR_ARG;
  for (int d = 0; d < 3; ++d)
    if (!crse_flux[d] || !fine_flux[d] || !(dx[d] > 0))
      return CXP_ERR_ARG;
  // the state FABs hold their valid cells and the components
  for (int i = 0; i < crse->nfabs; ++i)
    if (!fab_holds(crse->fabs[i], valid_rect(crse, i), comp, ncomp))
      return CXP_ERR_REGION;
  // coarse fluxes: each FAB holds the faces of its coarse valid box
  for (int d = 0; d < 3; ++d)
    for (int i = 0; i < crse->nfabs; ++i) {
      Rect r = valid_rect(crse, i);
      ++r.hi[d];
      if (!fab_holds(crse_flux[d][i], r, comp, ncomp))
        return CXP_ERR_REGION;
    }
  // fine boxes: coarsenable by 2 (AMReX's register needs it), fluxes on
  // each box's faces
  for (int b = 0; b < fine->nfabs; ++b) {
    const Rect B = valid_rect(fine, b);
    for (int d = 0; d < 3; ++d) {
      if ((B.lo[d] & 1) != 0 || (B.hi[d] & 1) != 1)
        return CXP_ERR_ARG;
      Rect r = B;
      ++r.hi[d];
      if (!fab_holds(fine_flux[d][b], r, comp, ncomp))
        return CXP_ERR_REGION;
    }
  }
  cxp_reflux_s *pl = new cxp_reflux_s();
  pl->comp = comp;
  pl->ncomp = ncomp;
  pl->cm = crse_mult;
  pl->fm = fine_mult;
  pl->scale = scale;
  pl->vol = dx[0] * dx[1] * dx[2];
  static const int kOrient[6][2] = {{0, 0}, {1, 0}, {2, 0}, {0, 1}, {1, 1}, {2, 1}};
  for (int o = 0; o < 6; ++o) {
    pl->chunk_off[o] = (int)pl->chunks.size();
    const int d = kOrient[o][0], side = kOrient[o][1];
    const int N = crse->domain_hi[d] - crse->domain_lo[d] + 1;
    for (int b = 0; b < fine->nfabs; ++b) {
      const Rect B = valid_rect(fine, b);
      Rect faces;
      for (int a = 0; a < 3; ++a) {
        faces.lo[a] = B.lo[a] >> 1;
        faces.hi[a] = B.hi[a] >> 1;
      }
      faces.lo[d] = faces.hi[d] = side ? (B.hi[d] + 1) >> 1 : B.lo[d] >> 1;
      // faces by coarse-flux FAB, the first whose valid faces hold them
      std::vector<Rect> left{faces};
      for (int ci = 0; ci < crse->nfabs && !left.empty(); ++ci) {
        Rect cf = valid_rect(crse, ci);
        ++cf.hi[d];
        std::vector<Rect> still;
        for (const Rect &r : left) {
          const Rect got = rect_and(r, cf);
          if (rect_empty(got)) {
            still.push_back(r);
            continue;
          }
          rect_minus(r, cf, still);
          // targets: the cell beyond each face (and its periodic images)
          const int shifts[3] = {0, -N, N};
          for (int s = 0; s < (periodic[d] ? 3 : 1); ++s) {
            const int off = (side ? 0 : -1) + shifts[s];
            for (int ti = 0; ti < crse->nfabs; ++ti) {
              Rect tv = valid_rect(crse, ti);
              tv.lo[d] -= off;  // target cells as face indices
              tv.hi[d] -= off;
              const Rect piece = rect_and(got, tv);
              if (rect_empty(piece))
                continue;
              RfPiece p;
              p.cflux = rf_fab(crse_flux[d][ci]);
              p.fflux = rf_fab(fine_flux[d][b]);
              p.state = rf_fab(crse->fabs[ti]);
              for (int a = 0; a < 3; ++a) {
                p.lo[a] = piece.lo[a];
                p.hi[a] = piece.hi[a];
              }
              p.dir = d;
              p.side = side;
              p.shift = shifts[s];
              const int t0 = d == 0 ? 1 : 0, t1 = d == 2 ? 1 : 2;
              const int nf = (piece.hi[t0] - piece.lo[t0] + 1) *
                             (piece.hi[t1] - piece.lo[t1] + 1);
              const int pi = (int)pl->pieces.size();
              pl->pieces.push_back(p);
              for (int st = 0; st < nf; st += kRfThreads)
                pl->chunks.push_back({pi, st});
            }
          }
        }
        left.swap(still);
      }
      if (!left.empty()) {
        delete pl;
        return CXP_ERR_COARSE;  // a register face no coarse flux holds
      }
    }
  }
  pl->chunk_off[6] = (int)pl->chunks.size();
  if (!pl->pieces.empty()) {
    if (hipMalloc(&pl->d_pieces, pl->pieces.size() * sizeof(RfPiece)) !=
            hipSuccess ||
        hipMalloc(&pl->d_chunks, pl->chunks.size() * sizeof(RfChunk)) !=
            hipSuccess ||
        hipMemcpy(pl->d_pieces, pl->pieces.data(),
                  pl->pieces.size() * sizeof(RfPiece),
                  hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(pl->d_chunks, pl->chunks.data(),
                  pl->chunks.size() * sizeof(RfChunk),
                  hipMemcpyHostToDevice) != hipSuccess) {
      if (pl->d_pieces)
        (void)hipFree(pl->d_pieces);
      if (pl->d_chunks)
        (void)hipFree(pl->d_chunks);
      delete pl;
      return CXP_ERR_HIP;
    }
  }
  *out = pl;
  return CXP_OK;
}

int cxp_reflux_run(cxp_reflux_plan pl, void *stream) {
  if (!pl)
    return CXP_ERR_ARG;
  for (int o = 0; o < 6; ++o) {
    const int n = pl->chunk_off[o + 1] - pl->chunk_off[o];
    if (n == 0)
      continue;
    hipLaunchKernelGGL(k_reflux, dim3((unsigned)n), dim3(kRfThreads), 0,
                       (hipStream_t)stream, pl->d_pieces,
                       pl->d_chunks + pl->chunk_off[o], pl->comp, pl->ncomp,
                       pl->cm, pl->fm, pl->scale, pl->vol);
    if (hipGetLastError() != hipSuccess)
      return CXP_ERR_HIP;
  }
  return CXP_OK;
}

int cxp_reflux_plan_stats(cxp_reflux_plan pl, int64_t *faces,
                          int64_t *bytes) {
  if (!pl)
    return CXP_ERR_ARG;
  int64_t nf = 0;
  for (const RfPiece &p : pl->pieces) {
    const int d = p.dir, t0 = d == 0 ? 1 : 0, t1 = d == 2 ? 1 : 2;
    nf += (int64_t)(p.hi[t0] - p.lo[t0] + 1) * (p.hi[t1] - p.lo[t1] + 1);
  }
  if (faces)
    *faces = nf;
  // per face and component: the coarse flux and 4 fine fluxes read, the
  // target cell read and written
  if (bytes)
    *bytes = nf * pl->ncomp * 8 * 7;
  return CXP_OK;
}

int cxp_reflux_plan_destroy(cxp_reflux_plan pl) {
  if (!pl)
    return CXP_ERR_ARG;
  if (pl->d_pieces)
    (void)hipFree(pl->d_pieces);
  if (pl->d_chunks)
    (void)hipFree(pl->d_chunks);
  delete pl;
  return CXP_OK;
}

} // extern "C"
