"""Reflux (CarpetX/src/schedule.cxx:2632-2716) restated in numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of
cxp_reflux_*; never by the product.

The reference calls, for each direction d, AMReX's
FluxRegister::CrseInit(coarse flux_d, d, 0, 0, n, -1) and
FluxRegister::FineAdd(fine flux_d, d, 0, 0, n, +1) (schedule.cxx:2694-2701),
then FluxRegister::Reflux(coarse state, 1.0, 0, 0, n, coarse geometry)
(:2702-2704).  The register was made for the fine level's cell-centred
BoxArray with ratio 2 (driver.cxx:918-924).  AMReX (25.11, pinned by the
reference's docker/*.dockerfile, not vendored) defines these as:

* register faces: for every fine valid box B and direction d, the coarse
  faces on the low and high sides of coarsen(B): face index B.lo[d] / 2 and
  (B.hi[d] + 1) / 2 in d, the coarsened cell range of B across;
* CrseInit(mult): register = mult * coarse flux on those faces (a copy);
* FineAdd(mult): register += mult * fine flux, for each of the 4 fine faces
  covering the coarse face (the higher transverse offset outer);
* Reflux(scale, geom): with V = dx[0] * dx[1] * dx[2] of the coarse level,
  for each orientation in the order x-lo, y-lo, z-lo, x-hi, y-hi, z-hi, the
  coarse cell on the far side of each register face (index - 1 for a low
  face, index for a high face; also across a periodic boundary) gets
  += -scale * reg / V (low) or += scale * reg / V (high), if it is a
  valid coarse cell.

The summation order inside AMReX is not vendored: parity unpinned at the
rounding level.  The formula is pinned by conservation and brute force
(tests/test_oracle_reflux.py).
"""
import numpy as np

ORIENTATIONS = [(0, 0), (1, 0), (2, 0), (0, 1), (1, 1), (2, 1)]  # (dir, side)


def _face_value(arr, lo, idx):
    """arr: numpy (nz, ny, nx) over a box from lo; idx (x, y, z)."""
    return arr[idx[2] - lo[2], idx[1] - lo[1], idx[0] - lo[0]]


def reflux(crse, crse_flux, fine_boxes, fine_flux, dx, domain, periodic=(0, 0, 0),
           crse_mult=-1.0, fine_mult=1.0, scale=1.0):
    """crse: list of (lo, hi, data, valid_lo, valid_hi), data (ncomp, nz, ny, nx)
    over [lo, hi], updated in place over its valid cells.
    crse_flux[d]: list of (lo, hi, data, vlo, vhi) face-centred in d, valid
    face boxes (vhi[d] = cell hi + 1).
    fine_boxes: list of (lo, hi) fine valid cell boxes (lo even, hi odd).
    fine_flux[d][b]: (lo, hi, data) of fine box b's flux FAB in direction d.
    domain: (lo, hi) coarse cell domain."""
    vol = dx[0] * dx[1] * dx[2]
    ncomp = crse[0][2].shape[0]
    dlo, dhi = domain
    for d, side in ORIENTATIONS:
        t = [a for a in range(3) if a != d]  # transverse, lower first
        N = dhi[d] - dlo[d] + 1
        for b, (blo, bhi) in enumerate(fine_boxes):
            I = blo[d] // 2 if side == 0 else (bhi[d] + 1) // 2
            flo, ffhi, fdata = fine_flux[d][b]
            r0 = range(blo[t[0]] // 2, bhi[t[0]] // 2 + 1)
            r1 = range(blo[t[1]] // 2, bhi[t[1]] // 2 + 1)
            for c1 in r1:
                for c0 in r0:
                    face = [0, 0, 0]
                    face[d], face[t[0]], face[t[1]] = I, c0, c1
                    src = None
                    for (clo, chi, cdata, vlo, vhi) in crse_flux[d]:
                        if all(vlo[a] <= face[a] <= vhi[a] for a in range(3)):
                            src = (clo, cdata)
                            break
                    if src is None:
                        raise ValueError("register face %s not covered by a coarse flux" % face)
                    for n in range(ncomp):
                        reg = crse_mult * _face_value(src[1][n], src[0], face)
                        for o1 in range(2):
                            for o0 in range(2):
                                ff = [0, 0, 0]
                                ff[d] = 2 * I
                                ff[t[0]] = 2 * c0 + o0
                                ff[t[1]] = 2 * c1 + o1
                                reg = reg + fine_mult * _face_value(fdata[n], flo, ff)
                        shifts = (0, -N, N) if periodic[d] else (0,)
                        for sh in shifts:
                            cell = list(face)
                            cell[d] = I + sh - 1 if side == 0 else I + sh
                            for (lo, hi, data, vlo, vhi) in crse:
                                if all(vlo[a] <= cell[a] <= vhi[a] for a in range(3)):
                                    v = (-scale * reg) / vol if side == 0 else (scale * reg) / vol
                                    data[n, cell[2] - lo[2], cell[1] - lo[1], cell[0] - lo[0]] += v
                                    break
