"""Pins of the Reflux oracle (oracle/reflux.py), CPU only.

* Conservation: a flux-divergence update of both levels (the same fluxes
  the register sees, weighted by dt and face area as AMReX expects),
  then Reflux and restriction, leaves the total of u over the union of the
  uncovered coarse cells and the fine cells unchanged; without Reflux it
  changes.  A wrong sign, face, neighbour cell, sub-face or volume breaks it.
* Brute force per coarse cell with the reference's own multipliers
  (CrseInit -1, FineAdd +1, Reflux 1.0; schedule.cxx:2694-2704): each coarse
  cell looks at its own six faces and takes the correction of any register
  face among them, a traversal independent of the oracle's per-face loop;
  periodic images included.
AMReX's summation order is not vendored, so the oracle's arithmetic order is
parity unpinned; these pins fix the formula.
"""
import itertools

import numpy as np
import pytest

from oracle import reflux as rf


def coarse_fabs(domain, xsplit, ncomp, rng):
    """Two coarse FABs split at x = xsplit (valid boxes), 1 ghost cell."""
    (dlo, dhi) = domain
    boxes = [((dlo[0], dlo[1], dlo[2]), (xsplit - 1, dhi[1], dhi[2])),
             ((xsplit, dlo[1], dlo[2]), (dhi[0], dhi[1], dhi[2]))]
    out = []
    for vlo, vhi in boxes:
        lo = tuple(v - 1 for v in vlo)
        hi = tuple(v + 1 for v in vhi)
        data = rng.uniform(-1, 1, (ncomp,) + tuple(hi[d] - lo[d] + 1 for d in (2, 1, 0)))
        out.append((lo, hi, data, vlo, vhi))
    return out


def face_box(lo, hi, d):
    h = list(hi)
    h[d] += 1
    return tuple(lo), tuple(h)


def random_face_array(lo, hi, ncomp, rng):
    return rng.uniform(-1, 1, (ncomp,) + tuple(hi[d] - lo[d] + 1 for d in (2, 1, 0)))


def make_fluxes(domain, crse, fine_boxes, ncomp, rng, zero_boundary):
    """Coarse face fluxes over the whole domain (each coarse FAB's valid face
    box a view of one global array, so shared faces agree) and fine fluxes
    per fine box from one global fine array (shared faces agree)."""
    dlo, dhi = domain
    cflux, fflux, gfine = [], [], []
    for d in range(3):
        glo, ghi = face_box(dlo, dhi, d)
        g = random_face_array(glo, ghi, ncomp, rng)
        if zero_boundary:
            sl = [slice(None)] * 4
            sl[3 - d] = 0
            g[tuple(sl)] = 0
            sl[3 - d] = -1
            g[tuple(sl)] = 0
        fabs = []
        for (_, _, _, vlo, vhi) in crse:
            flo, fhi = face_box(vlo, vhi, d)
            sl = (slice(None),) + tuple(slice(flo[a] - glo[a], fhi[a] - glo[a] + 1) for a in (2, 1, 0))
            fabs.append((flo, fhi, g[sl].copy(), flo, fhi))
        cflux.append(fabs)
        # fine: a global fine face array over the refined domain
        fdlo = tuple(2 * v for v in dlo)
        fdhi = tuple(2 * v + 1 for v in dhi)
        flo, fhi = face_box(fdlo, fdhi, d)
        G = random_face_array(flo, fhi, ncomp, rng)
        gfine.append((flo, G))
        per_box = []
        for blo, bhi in fine_boxes:
            blo_f, bhi_f = face_box(blo, bhi, d)
            sl = (slice(None),) + tuple(slice(blo_f[a] - flo[a], bhi_f[a] - flo[a] + 1) for a in (2, 1, 0))
            per_box.append((blo_f, bhi_f, G[sl].copy()))
        fflux.append(per_box)
    return cflux, fflux, gfine


def divergence_update(u, lo, cells_lo, cells_hi, flux_at, dt, dxv):
    """u -= dt / dx sum_d (F(i+1) - F(i)) on cells [cells_lo, cells_hi]."""
    for k, j, i in itertools.product(*(range(cells_lo[a], cells_hi[a] + 1) for a in (2, 1, 0))):
        c = [i, j, k]
        for d in range(3):
            up = list(c)
            up[d] += 1
            u[:, k - lo[2], j - lo[1], i - lo[0]] -= dt / dxv * (flux_at(d, up) - flux_at(d, c))


def test_reflux_restores_conservation():
    rng = np.random.default_rng(1)
    ncomp = 2
    domain = ((0, 0, 0), (7, 7, 7))
    crse = coarse_fabs(domain, 4, ncomp, rng)
    # two fine boxes sharing a face at fine x = 8 (coarse face 4), and a third
    # touching the second only at an edge
    fine_boxes = [((4, 4, 6), (7, 11, 9)), ((8, 4, 6), (11, 11, 9)), ((12, 12, 2), (13, 13, 3))]
    cflux, fflux, gfine = make_fluxes(domain, crse, fine_boxes, ncomp, rng, zero_boundary=True)
    dt, dxc = 0.01, 1.0 / 8
    dxf = dxc / 2
    Ac, Af, Vc, Vf = dxc ** 2, dxf ** 2, dxc ** 3, dxf ** 3
    # fine state on each fine box
    fine_u = [rng.uniform(-1, 1, (ncomp,) + tuple(hi[d] - lo[d] + 1 for d in (2, 1, 0)))
              for lo, hi in fine_boxes]

    def covered(c):
        return any(all(blo[a] // 2 <= c[a] <= bhi[a] // 2 for a in range(3)) for blo, bhi in fine_boxes)

    def total():
        t = np.zeros(ncomp)
        for (lo, hi, data, vlo, vhi) in crse:
            for k, j, i in itertools.product(*(range(vlo[a], vhi[a] + 1) for a in (2, 1, 0))):
                if not covered((i, j, k)):
                    t += Vc * data[:, k - lo[2], j - lo[1], i - lo[0]]
        for u in fine_u:
            t += Vf * u.sum(axis=(1, 2, 3))
        return t

    def cflux_at(d, c):
        for (flo, fhi, fd, vlo, vhi) in cflux[d]:
            if all(vlo[a] <= c[a] <= vhi[a] for a in range(3)):
                return fd[:, c[2] - flo[2], c[1] - flo[1], c[0] - flo[0]]
        raise AssertionError

    def restrict():
        for (blo, bhi), u in zip(fine_boxes, fine_u):
            for k, j, i in itertools.product(*(range(blo[a] // 2, bhi[a] // 2 + 1) for a in (2, 1, 0))):
                kids = u[:, 2 * k - blo[2]:2 * k - blo[2] + 2, 2 * j - blo[1]:2 * j - blo[1] + 2,
                         2 * i - blo[0]:2 * i - blo[0] + 2]
                for (lo, hi, data, vlo, vhi) in crse:
                    if all(vlo[a] <= (i, j, k)[a] <= vhi[a] for a in range(3)):
                        data[:, k - lo[2], j - lo[1], i - lo[0]] = kids.mean(axis=(1, 2, 3))
    restrict()
    t0 = total()
    for (lo, hi, data, vlo, vhi) in crse:
        divergence_update(data, lo, vlo, vhi, cflux_at, dt, dxc)
    for (blo, bhi), u, in zip(fine_boxes, fine_u):
        def ff(d, c, blo=blo):
            flo, G = gfine[d]
            return G[:, c[2] - flo[2], c[1] - flo[1], c[0] - flo[0]]
        divergence_update(u, blo, blo, bhi, ff, dt, dxf)
    before = [c[2].copy() for c in crse]
    restrict()
    t_noreflux = total()
    for c, b in zip(crse, before):
        c[2][...] = b
    rf.reflux(crse, cflux, fine_boxes, fflux, (dxc,) * 3, domain,
              crse_mult=-dt * Ac, fine_mult=dt * Af, scale=1.0)
    restrict()
    t1 = total()
    assert np.max(np.abs(t1 - t0)) <= 1e-14 * max(1.0, np.max(np.abs(t0)))
    # without the correction the totals drift
    assert np.max(np.abs(t_noreflux - t0)) > 1e-6


def brute_force(crse, cflux, fine_boxes, fflux, dx, domain, periodic, cm, fm, sc):
    """Per coarse cell: each of its faces that is a register face adds its
    correction; a register face of box B on B's low (high) side lies on the
    high (low) face of the cell outside B."""
    vol = dx[0] * dx[1] * dx[2]
    dlo, dhi = domain
    ncomp = crse[0][2].shape[0]
    out = [c[2].copy() for c in crse]
    for fab, (lo, hi, data, vlo, vhi) in zip(out, crse):
        for k, j, i in itertools.product(*(range(vlo[a], vhi[a] + 1) for a in (2, 1, 0))):
            cell = (i, j, k)
            for d in range(3):
                N = dhi[d] - dlo[d] + 1
                for side, fidx in ((0, cell[d] + 1), (1, cell[d])):
                    cands = [fidx] + ([fidx - N, fidx + N] if periodic[d] else [])
                    for b, (blo, bhi) in enumerate(fine_boxes):
                        I = blo[d] // 2 if side == 0 else (bhi[d] + 1) // 2
                        if I not in cands:
                            continue
                        if not all(blo[a] // 2 <= cell[a] <= bhi[a] // 2 for a in range(3) if a != d):
                            continue
                        face = list(cell)
                        face[d] = I
                        cf = next(fd for (flo, fhi, fd, vl, vh) in cflux[d]
                                  if all(vl[a] <= face[a] <= vh[a] for a in range(3)))
                        flo_c = next(flo for (flo, fhi, fd, vl, vh) in cflux[d]
                                     if all(vl[a] <= face[a] <= vh[a] for a in range(3)))
                        flo, fhi, fdat = fflux[d][b]
                        for n in range(ncomp):
                            subs = 0.0
                            for off in itertools.product((0, 1), repeat=2):
                                q = [2 * face[a] for a in range(3)]
                                t = [a for a in range(3) if a != d]
                                q[t[0]] += off[0]
                                q[t[1]] += off[1]
                                subs += fdat[n, q[2] - flo[2], q[1] - flo[1], q[0] - flo[0]]
                            reg = cm * cf[n, face[2] - flo_c[2], face[1] - flo_c[1], face[0] - flo_c[0]] \
                                + fm * subs
                            fab[n, k - lo[2], j - lo[1], i - lo[0]] += (-1 if side == 0 else 1) * sc * reg / vol
    return out


@pytest.mark.parametrize("periodic", [(0, 0, 0), (1, 0, 1)])
def test_reflux_brute_force_per_cell(periodic):
    rng = np.random.default_rng(2 + sum(periodic))
    ncomp = 2
    domain = ((0, 0, 0), (3, 3, 3))
    crse = coarse_fabs(domain, 2, ncomp, rng)
    # a fine box in the middle, one on the low x and high z faces (periodic
    # images when those directions are periodic)
    fine_boxes = [((2, 2, 2), (5, 5, 3)), ((0, 4, 4), (1, 7, 7))]
    cflux, fflux, _ = make_fluxes(domain, crse, fine_boxes, ncomp, rng, zero_boundary=False)
    dx = (0.25, 0.5, 0.125)
    orig = [c[2].copy() for c in crse]
    want = brute_force(crse, cflux, fine_boxes, fflux, dx, domain, periodic, -1.0, 1.0, 1.0)
    rf.reflux(crse, cflux, fine_boxes, fflux, dx, domain, periodic=periodic)
    for c, w in zip(crse, want):
        assert np.max(np.abs(c[2] - w)) <= 1e-13 * max(1.0, np.max(np.abs(w)))
    # valid cells next to register faces changed; ghost cells did not
    assert sum(int(np.sum(c[2] != o)) for c, o in zip(crse, orig)) > 0
    for (lo, hi, data, vlo, vhi), o in zip(crse, orig):
        ghost = np.ones(data.shape[1:], dtype=bool)
        ghost[tuple(slice(vlo[a] - lo[a], vhi[a] - lo[a] + 1) for a in (2, 1, 0))] = False
        assert np.array_equal(data[:, ghost], o[:, ghost])
