"""Boundary values of CarpetX's BoundaryCondition::apply on one FAB, restated
point by point (CarpetX/src/boundaries_impl.hxx:150-437; boundaries.cxx:
34-56 for imin, imax, and the coordinates).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
boundary kernel, never by the product (carpetx_amd).

Every point dst of the FAB outside the domain [dom_lo, dom_hi] is a point of
the face region whose normal is nonzero exactly in the directions where dst
is outside (apply_on_face, boundaries_impl.hxx:150-196).  In each such
direction d, with the face's condition (:322-371):
  * dirichlet:             no source;
  * linear extrapolation:  src[d] = edge, delta[d] = -normal[d];
  * neumann, robin:        src[d] = edge;
  * reflection:            src[d] = offset - dst[d], offset = 2 imin - cell
                           (lower) or 2 (imax - 1) + cell (upper), :286-298;
  * none or periodic:      the direction is not a boundary (src[d] = dst[d]);
                           periodic images are the coarse-patch copy's job.
If every such direction is "none" the point is left alone.  The value
(:380-417): any dirichlet direction -> the component's Dirichlet value;
otherwise var(src), then for each Robin direction
val = robin + (|x(src)| / |x(dst)|) (val - robin), then with linear
extrapolation val += sqrt(S / T) (val - var(src + delta)), S = sum (dst -
src)^2 and T = sum delta^2 divided as integers, as the reference divides two
int sums; then the product of the reflection directions' parities.

The restatement is pinned by tests/test_oracle_boundary.py against closed
forms: constants, linear data, data of given parity about the mirror point,
and f = finf + C / r.
"""
import math

import numpy as np

NONE, DIRICHLET, LINEAR_EXTRAPOLATION, NEUMANN, ROBIN, REFLECTION, PERIODIC = range(7)


def periodic_fill(arr, lo, dom_lo, dom_hi, dirs, cell):
    """The coarse-patch copy's periodic images (the reference copies with
    the geometry's periodicity, CarpetX/src/fillpatch.cxx:91-93): every
    point of arr outside the domain only in periodic directions ``dirs``
    takes the value of its image inside the domain, period dom_hi - dom_lo
    + cell[d].  Points inside the domain keep their values."""
    hi = [lo[d] + arr.shape[3 - d] - 1 for d in range(3)]
    src = arr.copy()
    for k in range(lo[2], hi[2] + 1):
        for j in range(lo[1], hi[1] + 1):
            for i in range(lo[0], hi[0] + 1):
                p = [i, j, k]
                q = list(p)
                ok, moved = True, False
                for d in range(3):
                    if dom_lo[d] <= p[d] <= dom_hi[d]:
                        continue
                    if d not in dirs:
                        ok = False
                        break
                    per = dom_hi[d] - dom_lo[d] + cell[d]
                    while q[d] < dom_lo[d]:
                        q[d] += per
                    while q[d] > dom_hi[d]:
                        q[d] -= per
                    moved = True
                if ok and moved:
                    arr[:, k - lo[2], j - lo[1], i - lo[0]] = \
                        src[:, q[2] - lo[2], q[1] - lo[1], q[0] - lo[0]]
    return arr


def apply_bc(arr, lo, dom_lo, dom_hi, faces, cell, dirichlet=None, robin=None,
             parities=None, xmin=(0.0, 0.0, 0.0), dx=(1.0, 1.0, 1.0)):
    """Fill, in place, the points of arr (ncomp, nz, ny, nx; box from lo)
    outside the domain [dom_lo, dom_hi] (inclusive).  faces[f][d]: condition
    of face f (0 lower, 1 upper) in direction d; cell[d]: 1 for cell
    centring; parities[c][d]: +1 / -1."""
    ncomp = arr.shape[0]
    hi = [lo[d] + arr.shape[3 - d] - 1 for d in range(3)]
    imin = list(dom_lo)
    imax = [dom_hi[d] + 1 for d in range(3)]

    def at(p):
        return (slice(None), p[2] - lo[2], p[1] - lo[1], p[0] - lo[0])

    for k in range(lo[2], hi[2] + 1):
        for j in range(lo[1], hi[1] + 1):
            for i in range(lo[0], hi[0] + 1):
                dst = (i, j, k)
                src = list(dst)
                delta = [0, 0, 0]
                active = dirich = lin = False
                robin_dirs, refl_dirs = [], []
                for d in range(3):
                    if imin[d] <= dst[d] < imax[d]:
                        continue
                    face = 1 if dst[d] >= imax[d] else 0
                    edge = imax[d] - 1 if face else imin[d]
                    t = faces[face][d]
                    if t in (NONE, PERIODIC):
                        continue  # periodic faces are translated to none (:160-161)
                    active = True
                    if t == DIRICHLET:
                        dirich = True
                    elif t == LINEAR_EXTRAPOLATION:
                        src[d] = edge
                        delta[d] = -1 if face else 1
                        lin = True
                    elif t in (NEUMANN, ROBIN):
                        src[d] = edge
                        if t == ROBIN:
                            robin_dirs.append(d)
                    elif t == REFLECTION:
                        off = (2 * (imax[d] - 1) + cell[d] if face
                               else 2 * imin[d] - cell[d])
                        src[d] = off - dst[d]
                        refl_dirs.append(d)
                    else:
                        raise ValueError("unknown condition %r" % (t,))
                if not active:
                    continue
                if dirich:
                    val = np.array(dirichlet[:ncomp], dtype=np.float64)
                else:
                    val = arr[at(src)].copy()
                    for d in robin_dirs:
                        xb = [xmin[e] + dst[e] * dx[e] for e in range(3)]
                        xi = [xmin[e] + src[e] * dx[e] for e in range(3)]
                        rb = math.sqrt(0.0 + xb[0] * xb[0] + xb[1] * xb[1] + xb[2] * xb[2])
                        ri = math.sqrt(0.0 + xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2])
                        q = ri / rb
                        rv = np.array(robin[:ncomp], dtype=np.float64)
                        val = rv + q * (val - rv)
                    if lin:
                        sp = [src[e] + delta[e] for e in range(3)]
                        grad = val - arr[at(sp)]
                        num = sum((dst[e] - src[e]) ** 2 for e in range(3))
                        den = sum(delta[e] ** 2 for e in range(3))
                        val = val + math.sqrt(num // den) * grad
                    if refl_dirs:
                        par = np.ones(ncomp)
                        for c in range(ncomp):
                            for d in refl_dirs:
                                par[c] *= parities[c][d]
                        val = val * par
                arr[at(dst)] = val
    return arr
