"""Synthetic AMR workloads (seeded) for the benchmark and the tests.

* ``single_box``   one box, VC 3rd-order: coarse nodal [-3, 35]^3 (32^3 cells
                   plus 3 ghosts) to the fine interior [0, 63]^3, then VC
                   injection back; uniform[-1, 1] data and a random tensor
                   cubic.
* ``level_64``     a coarse 128^3 level with 64 fine boxes of 32^3 points at
                   seeded non-overlapping coarse origins and 4 grid functions;
                   CCC conservative 2nd order and VVV 5th-order polynomial
                   prolongation of whole boxes; Gaussian pulse (amplitude 1,
                   width 0.1 L) plus 1e-3 uniform noise.
* ``level_fill_64`` the same 128^3 level and 64 fine boxes as level_64, as
                   two levels of FABs: the coarse level split into 2^3 FABs
                   of 64^3 cells plus 6 ghost points (which hold the pulse,
                   standing in for boundary values), the fine boxes with 3
                   ghost points; the ghost fill of the whole fine level.
* ``uneven_512``   512 fine boxes with per-dimension extents drawn from
                   {8, 16, ..., 64}, per-direction VC/CC centring and order
                   1/3/5 of the ``natural`` family; 10 grid functions of smooth
                   sin products plus 1e-6 noise; a 3-point ghost-shell fill.

All data lives on the GPU; coarse levels are read in place.
"""
import math

import numpy as np
import torch

from .interp import Box, FArrayBox, get_interpolator


def _uniform(shape, gen, lo=-1.0, hi=1.0):
    return torch.rand(shape, generator=gen, dtype=torch.float64, device="cuda") * (hi - lo) + lo


def single_box(seed=1, order=3):
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    op = get_interpolator("interpolate", order, (0, 0, 0))
    cbox = Box((-3, -3, -3), (35, 35, 35))
    crse = FArrayBox(cbox, 2)
    crse.t[0] = _uniform(crse.t[0].shape, gen)
    # random tensor cubic in coarse coordinates x = i (exactly reproduced)
    rng = np.random.default_rng(seed + 1)
    a = rng.uniform(-1, 1, size=(3, order + 1))
    axes = []
    for d in range(3):
        x = torch.arange(cbox.lo[d], cbox.hi[d] + 1, dtype=torch.float64, device="cuda") / 32.0
        axes.append(sum(a[d, q] * x ** q for q in range(order + 1)))
    crse.t[1] = axes[2][:, None, None] * axes[1][None, :, None] * axes[0][None, None, :]
    region = Box((0, 0, 0), (63, 63, 63))
    fine = FArrayBox(region, 2)
    return dict(op=op, crse=crse, fine=fine, region=region, poly_coeffs=a,
                restrict_region=Box((0, 0, 0), (31, 31, 31)))


def _gaussian_level(cbox, indextype, ncomp, gen, ncells):
    """Gaussian pulse exp(-r^2 / (2 (0.1 L)^2)) + 1e-3 u on a coarse level of
    ``ncells`` cells of width 1/ncells (L = 1, centred at 0.5)."""
    fab = FArrayBox(cbox, ncomp)
    h = 1.0 / ncells
    coords = []
    for d in range(3):
        i = torch.arange(cbox.lo[d], cbox.hi[d] + 1, dtype=torch.float64, device="cuda")
        coords.append((i + 0.5 * indextype[d]) * h - 0.5)
    r2 = (coords[2][:, None, None] ** 2 + coords[1][None, :, None] ** 2 +
          coords[0][None, None, :] ** 2)
    g = torch.exp(-r2 / (2 * 0.1 ** 2))
    for c in range(ncomp):
        fab.t[c] = g + 1e-3 * _uniform(g.shape, gen)
    return fab


def level_64(seed=3, ncomp=4, nboxes=64, box_pts=32):
    """Returns {name: dict(op, crse, fines, regions)} for the two operators."""
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    rng = np.random.default_rng(seed)
    nc = 128
    slot = box_pts // 2  # coarse extent of a fine box
    nslot = nc // slot
    slots = rng.choice(nslot ** 3, size=nboxes, replace=False)
    out = {}
    for name, typ, order, it in [("cons2_ccc", "conservative", 1, (1, 1, 1)),
                                 ("poly5_vvv", "interpolate", 5, (0, 0, 0))]:
        op = get_interpolator(typ, order, it)
        # coarse level: cells 0..127 (CC) or nodes 0..128 (VC), plus ghosts
        g = 4
        cbox = Box([-g] * 3, [nc - it[d] + g for d in range(3)])
        crse = _gaussian_level(cbox, it, ncomp, gen, nc)
        fines, regions = [], []
        for s in slots:
            sk, sj, si = s // (nslot * nslot), (s // nslot) % nslot, s % nslot
            lo = [2 * slot * si, 2 * slot * sj, 2 * slot * sk]
            region = Box(lo, [lo[d] + box_pts - 1 for d in range(3)])
            fines.append(FArrayBox(region, ncomp))
            regions.append(region)
        out[name] = dict(op=op, crse=crse, fines=fines, regions=regions, ncomp=ncomp)
    return out


def level_fill_64(seed=3, ncomp=4, nboxes=64, box_pts=32, ghosts=3, cghosts=6,
                  regrid=False):
    """Returns {name: dict(op, crse, fine, ncomp)}, crse and fine Levels for
    WorkList.append_level, for the two operators of level_64.  With regrid,
    each entry also holds "new": a fine Level of nboxes boxes on the same
    slot lattice, half of them at old boxes' slots and half at slots the
    old level leaves free (no data yet), for the regrid-time fills."""
    from .batched import Level
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    rng = np.random.default_rng(seed)
    nc = 128
    slot = box_pts // 2
    nslot = nc // slot
    slots = rng.choice(nslot ** 3, size=nboxes, replace=False)
    free = np.setdiff1d(np.arange(nslot ** 3), slots)
    new_slots = np.concatenate([slots[:nboxes // 2],
                                rng.choice(free, size=nboxes - nboxes // 2, replace=False)])

    def fine_level(slot_ids, fdom):
        ffabs, fvalid = [], []
        for s in slot_ids:
            sk, sj, si = s // (nslot * nslot), (s // nslot) % nslot, s % nslot
            lo = [2 * slot * si, 2 * slot * sj, 2 * slot * sk]
            v = Box(lo, [lo[d] + box_pts - 1 for d in range(3)])
            ffabs.append(FArrayBox(v.grow(ghosts), ncomp))
            fvalid.append(v)
        return Level(ffabs, fvalid, fdom)
    out = {}
    for name, typ, order, it in [("cons2_ccc", "conservative", 1, (1, 1, 1)),
                                 ("poly5_vvv", "interpolate", 5, (0, 0, 0))]:
        op = get_interpolator(typ, order, it)
        cdom = Box([0] * 3, [nc - it[d] for d in range(3)])
        whole = _gaussian_level(cdom.grow(cghosts), it, ncomp, gen, nc)
        cfabs, cvalid = [], []
        for c in range(8):
            lo = [0 if not (c >> d) & 1 else nc // 2 for d in range(3)]
            hi = [nc // 2 - 1 if not (c >> d) & 1 else cdom.hi[d] for d in range(3)]
            v = Box(lo, hi)
            fab = FArrayBox(v.grow(cghosts), ncomp)
            fab.t.copy_(whole.view(fab.box))
            cfabs.append(fab)
            cvalid.append(v)
        del whole
        fdom = Box([0] * 3, [2 * nc - it[d] for d in range(3)])
        out[name] = dict(op=op, crse=Level(cfabs, cvalid, cdom),
                         fine=fine_level(slots, fdom), ncomp=ncomp, indextype=it)
        if regrid:
            out[name]["new"] = fine_level(new_slots, fdom)
    return out


def uneven_512(seed=0, nboxes=512, ncomp=10, ghosts=3, family="natural"):
    """512 uneven boxes, mixed centring and order, ghost-shell fill."""
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed + 100)
    exts = rng.choice(np.arange(8, 65, 8), size=(nboxes, 3))
    cents = rng.integers(0, 2, size=(nboxes, 3))
    orders = rng.choice([1, 3, 5], size=nboxes)
    pitch = 80  # fine points per lattice slot (box + ghosts + stencil reach)
    by_cent = {}
    for b in range(nboxes):
        by_cent.setdefault(tuple(int(x) for x in cents[b]), []).append(b)
    boxes = [None] * nboxes
    levels = {}
    for cent, ids in sorted(by_cent.items()):
        ns = max(1, math.ceil(len(ids) ** (1.0 / 3.0) - 1e-9))
        while ns ** 3 < len(ids):
            ns += 1
        ncoarse = ns * pitch // 2
        g = 8
        cbox = Box([-g] * 3, [ncoarse - cent[d] + g for d in range(3)])
        # smooth sin products with per-GF wave numbers, + 1e-6 noise
        crse = FArrayBox(cbox, ncomp)
        h = 1.0 / ncoarse
        coords = [((torch.arange(cbox.lo[d], cbox.hi[d] + 1, dtype=torch.float64,
                                 device="cuda") + 0.5 * cent[d]) * h) for d in range(3)]
        for c in range(ncomp):
            k = rng.uniform(1.0, 4.0, size=3) * math.pi
            f = (torch.sin(k[2] * coords[2])[:, None, None] *
                 torch.sin(k[1] * coords[1])[None, :, None] *
                 torch.sin(k[0] * coords[0])[None, None, :])
            crse.t[c] = f + 1e-6 * _uniform(f.shape, gen)
        levels[cent] = crse
        slots = rng.permutation(ns ** 3)[:len(ids)]
        for b, s in zip(ids, slots):
            sk, sj, si = s // (ns * ns), (s // ns) % ns, s % ns
            lo = []
            for d, sd in enumerate((si, sj, sk)):
                room = pitch - 2 * 8 - int(exts[b, d])
                lo.append(sd * pitch + 8 + 2 * int(rng.integers(0, room // 2 + 1)))
            interior = Box(lo, [lo[d] + int(exts[b, d]) - 1 for d in range(3)])
            fab = FArrayBox(interior.grow(ghosts), ncomp)
            op = get_interpolator(family, int(orders[b]), cent)
            boxes[b] = dict(op=op, crse=crse, fine=fab, interior=interior,
                            cent=cent, order=int(orders[b]))
    return dict(boxes=boxes, levels=levels, ncomp=ncomp, ghosts=ghosts)


def shell_slabs(fab_box, interior):
    """The six ghost-shell slabs (same decomposition as the C ABI)."""
    glo, ghi, ilo, ihi = fab_box.lo, fab_box.hi, interior.lo, interior.hi
    out = []

    def add(x0, x1, y0, y1, z0, z1):
        if x0 <= x1 and y0 <= y1 and z0 <= z1:
            out.append(Box((x0, y0, z0), (x1, y1, z1)))
    add(glo[0], ghi[0], glo[1], ghi[1], glo[2], ilo[2] - 1)
    add(glo[0], ghi[0], glo[1], ghi[1], ihi[2] + 1, ghi[2])
    add(glo[0], ghi[0], glo[1], ilo[1] - 1, ilo[2], ihi[2])
    add(glo[0], ghi[0], ihi[1] + 1, ghi[1], ilo[2], ihi[2])
    add(glo[0], ilo[0] - 1, ilo[1], ihi[1], ilo[2], ihi[2])
    add(ihi[0] + 1, ghi[0], ilo[1], ihi[1], ilo[2], ihi[2])
    return out
