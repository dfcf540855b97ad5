"""GPU level ghost fill (cxp_worklist_append_level) against the oracle.

The fill follows FillPatch_ProlongateGhosts (CarpetX/src/fillpatch.cxx:
42-138): every fine FAB point that lies in no fine valid region, inside the
fine domain grown by the ghost width, is prolongated from the coarse level;
coarse values inside the coarse domain come only from the coarse valid
regions (the reference copies its coarse patch without coarse ghosts,
:91-93), and outside it from the coarse FABs' ghost points, which hold the
boundary values (the reference applies the boundary conditions to its
coarse patch, :107).

Expected values: the oracle prolongates each whole fine FAB box from one
global coarse array (valid data inside the domain, boundary data outside),
and the test keeps those values exactly where a fill must write and the
original values elsewhere.  Coarse FAB ghost points inside the domain hold
1e300, so reading them instead of a neighbour's valid data would show.
"""
import numpy as np
import pytest
import torch

import oracle
from carpetx_amd import _lib
from carpetx_amd.batched import BoundaryConditions, Graph, Level, WorkList
from oracle.boundary import apply_bc, periodic_fill
from carpetx_amd.interp import Box, FArrayBox, get_interpolator

pytestmark = pytest.mark.gpu

N = 24        # coarse cells per direction
CG = 6        # coarse FAB ghosts
FG = 3        # fine FAB ghosts
NC_CRSE = 3   # coarse components; the fill reads components 1..2
NCOMP = 2
FINE_VALID = [((0, 0, 0), (9, 9, 9)),          # reads one coarse FAB in place
              ((10, 0, 0), (21, 9, 9)),        # beside the first: shared ghosts
              ((16, 14, 12), (31, 29, 27)),    # straddles coarse FABs: gathered
              ((36, 36, 36), (47, 47, 47))]    # upper corner: boundary ghosts


def domains(it):
    cdom = Box((0, 0, 0), tuple(N - it[d] for d in range(3)))
    fdom = Box((0, 0, 0), tuple(2 * N - it[d] for d in range(3)))
    return cdom, fdom


def coarse_level(it, g, cg=CG, poison_outside=False, split=(12, 12, 12)):
    """8 coarse FABs (valid split at `split`) from the global array g, whose
    box is the coarse domain grown by CG; with poison_outside, the points
    outside the domain hold 1e300 as well."""
    cdom, _ = domains(it)
    glo = tuple(cdom.lo[d] - CG for d in range(3))
    fabs, valid = [], []
    for c in range(8):
        lo = tuple(0 if not (c >> d) & 1 else split[d] for d in range(3))
        hi = tuple(split[d] - 1 if not (c >> d) & 1 else cdom.hi[d] for d in range(3))
        v = Box(lo, hi)
        b = v.grow(cg)
        sl = tuple(slice(b.lo[d] - glo[d], b.hi[d] - glo[d] + 1) for d in (2, 1, 0))
        arr = g[(slice(None),) + sl].copy()
        # ghost points inside the domain must not be read
        idx = np.meshgrid(*[np.arange(b.lo[d], b.hi[d] + 1) for d in (2, 1, 0)], indexing="ij")
        ind = np.ones(arr.shape[1:], dtype=bool)
        inv = np.ones(arr.shape[1:], dtype=bool)
        for a, d in zip(idx, (2, 1, 0)):
            ind &= (a >= cdom.lo[d]) & (a <= cdom.hi[d])
            inv &= (a >= v.lo[d]) & (a <= v.hi[d])
        arr[:, ind & ~inv] = 1e300
        if poison_outside:
            arr[:, ~ind] = 1e300
        fabs.append(FArrayBox(b, NC_CRSE, tensor=torch.from_numpy(arr).cuda()))
        valid.append(v)
    return Level(fabs, valid, cdom)


def fine_level(it, rng, boxes=FINE_VALID, fill_valid=True):
    """Fine FABs of the given valid boxes grown by FG: uniform[-1, 1] in the
    valid region (7.0 with fill_valid False), 7.0 in the ghosts."""
    _, fdom = domains(it)
    fabs, valid = [], []
    for lo, hi in boxes:
        hi = tuple(min(hi[d], fdom.hi[d]) for d in range(3))
        v = Box(lo, hi)
        b = v.grow(FG)
        t = torch.full((NCOMP,) + tuple(reversed(b.size)), 7.0, dtype=torch.float64)
        inner = tuple(slice(FG, FG + v.size[d]) for d in (2, 1, 0))
        if fill_valid:
            t[(slice(None),) + inner] = torch.from_numpy(
                rng.uniform(-1, 1, (NCOMP,) + tuple(reversed(v.size))))
        fabs.append(FArrayBox(b, NCOMP, tensor=t.cuda()))
        valid.append(v)
    return Level(fabs, valid, fdom)


def expected(op_args, it, g, fine, per_group=False):
    cdom, fdom = domains(it)
    oop = oracle.lookup(*op_args)
    glo = tuple(cdom.lo[d] - CG for d in range(3))
    ghi = tuple(cdom.hi[d] + CG for d in range(3))
    ch = oracle.Fab(glo, ghi, NCOMP, np.ascontiguousarray(g[1:1 + NCOMP]))
    out = []
    for fab in fine.fabs:
        b = fab.box
        orig = fab.t.cpu().numpy()
        fh = oracle.Fab(b.lo, b.hi, NCOMP, np.zeros_like(orig))
        oracle.prolongate(oop, ch, fh, b.lo, b.hi, per_group=per_group)
        idx = np.meshgrid(*[np.arange(b.lo[d], b.hi[d] + 1) for d in (2, 1, 0)], indexing="ij")
        fill = np.ones(orig.shape[1:], dtype=bool)
        for a, d in zip(idx, (2, 1, 0)):
            fill &= (a >= fdom.lo[d] - FG) & (a <= fdom.hi[d] + FG)
        for v in fine.valid:
            inv = np.ones(orig.shape[1:], dtype=bool)
            for a, d in zip(idx, (2, 1, 0)):
                inv &= (a >= v.lo[d]) & (a <= v.hi[d])
            fill &= ~inv
        out.append(np.where(fill[None], fh.arr, orig))
    return out


OPS = [("natural", 5, (0, 0, 0)), ("natural", 3, (1, 1, 1)), ("ddf", 3, (0, 1, 0)),
       ("minmod", 3, (1, 0, 1)), ("eno", 3, (1, 1, 1))]


@pytest.mark.parametrize("op_args", OPS)
def test_level_fill_matches_oracle(op_args):
    """Batched run, per-patch run and two graph replays each equal the
    oracle bit for bit; after the coarse data change, the next run reads the
    new values (the gather runs at every launch)."""
    it = op_args[2]
    rng = np.random.default_rng(sum(it) + op_args[1])
    cdom, _ = domains(it)
    gshape = (NC_CRSE,) + tuple(cdom.size[d] + 2 * CG for d in (2, 1, 0))
    g = rng.uniform(-1, 1, gshape)
    crse = coarse_level(it, g)
    fine = fine_level(it, rng)
    originals = [f.t.clone() for f in fine.fabs]
    want = expected(op_args, it, g, fine)
    op = get_interpolator(*op_args)
    wl = WorkList()
    wl.append_level(op, crse, 1, fine, 0, NCOMP)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        wl.upload(s)
        wl.run(s)
    torch.cuda.synchronize()
    for f, w in zip(fine.fabs, want):
        assert np.array_equal(f.t.cpu().numpy(), w), op_args

    def reset():
        for f, o in zip(fine.fabs, originals):
            f.t.copy_(o)
        torch.cuda.synchronize()

    reset()
    with torch.cuda.stream(s):
        wl.run_per_patch(s)
    torch.cuda.synchronize()
    for f, w in zip(fine.fabs, want):
        assert np.array_equal(f.t.cpu().numpy(), w), ("per patch", op_args)
    with torch.cuda.stream(s):
        graph = Graph(lambda: wl.run(s), s)
    for rep in range(2):
        reset()
        with torch.cuda.stream(s):
            graph.launch(s)
        torch.cuda.synchronize()
        for f, w in zip(fine.fabs, want):
            assert np.array_equal(f.t.cpu().numpy(), w), ("graph", rep, op_args)
    # new coarse data: the next replay gathers it
    g2 = rng.uniform(-1, 1, gshape)
    crse2 = coarse_level(it, g2)
    for a, b in zip(crse.fabs, crse2.fabs):
        a.t.copy_(b.t)
    reset()
    want2 = expected(op_args, it, g2, fine)
    with torch.cuda.stream(s):
        graph.launch(s)
    torch.cuda.synchronize()
    for f, w in zip(fine.fabs, want2):
        assert np.array_equal(f.t.cpu().numpy(), w), ("new coarse data", op_args)
    graph.close()
    wl.close()


def test_level_fill_needs_coarse_boundary_points():
    """Coarse FABs with one ghost point do not hold the boundary values that
    the corner fine box needs: CXP_ERR_COARSE, and nothing is appended."""
    it = (1, 1, 1)
    rng = np.random.default_rng(3)
    cdom, _ = domains(it)
    g = rng.uniform(-1, 1, (NC_CRSE,) + tuple(cdom.size[d] + 2 * CG for d in (2, 1, 0)))
    crse = coarse_level(it, g, cg=1)
    fine = fine_level(it, rng)
    wl = WorkList()
    with pytest.raises(_lib.CxpError) as e:
        wl.append_level(get_interpolator("natural", 3, it), crse, 1, fine, 0, NCOMP)
    assert e.value.status == -5
    assert wl.stats()["patches"] == 0
    wl.close()


BC_KINDS = {"none": 0, "dirichlet": 1, "linear_extrapolation": 2,
            "neumann": 3, "robin": 4, "reflection": 5, "periodic": 6}


def _periodic_dirs(faces):
    return tuple(d for d in range(3) if faces[0][d] == "periodic")
BC_CASES = [
    (("natural", 5, (0, 0, 0)),
     [["reflection", "linear_extrapolation", "robin"],
      ["dirichlet", "neumann", "reflection"]]),
    (("natural", 3, (1, 1, 1)),
     [["neumann", "reflection", "dirichlet"],
      ["linear_extrapolation", "robin", "reflection"]]),
    (("interpolate", 7, (1, 0, 1)),
     [["reflection"] * 3, ["linear_extrapolation"] * 3]),
    (("eno", 3, (1, 1, 1)),
     [["reflection", "robin", "linear_extrapolation"], ["neumann"] * 3]),
    (("minmod", 3, (0, 1, 0)),
     [["dirichlet", "reflection", "neumann"], ["reflection"] * 3]),
    (("natural", 5, (0, 0, 0)),
     [["periodic", "reflection", "neumann"], ["periodic", "linear_extrapolation", "dirichlet"]]),
    (("natural", 3, (1, 1, 1)),
     [["reflection", "periodic", "periodic"], ["neumann", "periodic", "periodic"]]),
]


@pytest.mark.parametrize("op_args,faces", BC_CASES)
def test_level_fill_with_boundary_conditions_matches_oracle(op_args, faces):
    """cxp_worklist_append_level_bc: coarse points outside the coarse domain
    are computed from the gathered coarse patch, never read from the coarse
    FABs (which hold 1e300 there).  Expected: the oracle's boundary
    conditions applied to the global coarse array, then the oracle's
    prolongation of every fine FAB box, kept where the fill writes.  The
    sources of every boundary point lie inside the domain, so a patch's
    boundary values equal the global array's.  Tolerance 1e-13 (the Robin
    and linear-extrapolation values take a square root); batched, per-patch
    and graph runs, then new coarse data through the same graph."""
    it = op_args[2]
    rng = np.random.default_rng(7 + sum(it) + op_args[1])
    cdom, _ = domains(it)
    gshape = (NC_CRSE,) + tuple(cdom.size[d] + 2 * CG for d in (2, 1, 0))
    glo = tuple(cdom.lo[d] - CG for d in range(3))
    kw = dict(dirichlet=[0.5, -1.5], robin=[0.25, 2.0],
              parities=[(-1, 1, -1), (1, -1, 1)],
              xmin=(0.37, -1.21, 0.53), dx=(0.05, 0.07, 0.06))
    bc = BoundaryConditions(faces, dirichlet_values=kw["dirichlet"],
                            robin_values=kw["robin"], parities=kw["parities"],
                            xmin=kw["xmin"], dx=kw["dx"])
    codes = [[BC_KINDS[faces[f][d]] for d in range(3)] for f in range(2)]

    def with_bc(g):
        gb = g.copy()
        sub = np.ascontiguousarray(gb[1:1 + NCOMP])
        periodic_fill(sub, glo, cdom.lo, cdom.hi, _periodic_dirs(faces), it)
        apply_bc(sub, glo, cdom.lo, cdom.hi, codes, it, **kw)
        gb[1:1 + NCOMP] = sub
        return gb

    g = rng.uniform(-1, 1, gshape)
    crse = coarse_level(it, g, poison_outside=True)
    fine = fine_level(it, rng)
    originals = [f.t.clone() for f in fine.fabs]
    want = expected(op_args, it, with_bc(g), fine)
    op = get_interpolator(*op_args)
    wl = WorkList()
    wl.append_level(op, crse, 1, fine, 0, NCOMP, bc=bc)
    s = torch.cuda.Stream()

    def check(tag, want):
        for f, w in zip(fine.fabs, want):
            got = f.t.cpu().numpy()
            np.testing.assert_allclose(got, w, rtol=1e-13, atol=1e-13,
                                       err_msg=str((tag, op_args, faces)))

    def reset():
        for f, o in zip(fine.fabs, originals):
            f.t.copy_(o)
        torch.cuda.synchronize()

    with torch.cuda.stream(s):
        wl.upload(s)
        wl.run(s)
    torch.cuda.synchronize()
    check("batched", want)
    reset()
    with torch.cuda.stream(s):
        wl.run_per_patch(s)
    torch.cuda.synchronize()
    check("per patch", want)
    with torch.cuda.stream(s):
        graph = Graph(lambda: wl.run(s), s)
    for rep in range(2):
        reset()
        with torch.cuda.stream(s):
            graph.launch(s)
        torch.cuda.synchronize()
        check(("graph", rep), want)
    g2 = rng.uniform(-1, 1, gshape)
    crse2 = coarse_level(it, g2, poison_outside=True)
    for a, b in zip(crse.fabs, crse2.fabs):
        a.t.copy_(b.t)
    reset()
    want2 = expected(op_args, it, with_bc(g2), fine)
    with torch.cuda.stream(s):
        graph.launch(s)
    torch.cuda.synchronize()
    check("new coarse data", want2)
    graph.close()
    wl.close()


NEW_VALID = [((0, 0, 0), (15, 11, 9)), ((16, 0, 0), (31, 13, 9)),
             ((30, 30, 30), (47, 47, 47))]
REGRID_CASES = [
    (("natural", 3, (1, 1, 1)), None, None),
    (("natural", 5, (0, 0, 0)),
     [["periodic", "reflection", "neumann"], ["periodic", "linear_extrapolation", "dirichlet"]],
     [["periodic", "neumann", "reflection"], ["periodic", "dirichlet", "neumann"]]),
    (("natural", 5, (0, 0, 0)),
     [["reflection", "linear_extrapolation", "robin"],
      ["dirichlet", "neumann", "reflection"]],
     [["neumann", "reflection", "linear_extrapolation"],
      ["reflection", "dirichlet", "neumann"]]),
    (("eno", 3, (1, 0, 1)),
     [["reflection"] * 3, ["linear_extrapolation"] * 3],
     [["linear_extrapolation"] * 3, ["reflection"] * 3]),
]
BC_KW = dict(dirichlet=[0.5, -1.5], robin=[0.25, 2.0],
             parities=[(-1, 1, -1), (1, -1, 1)],
             xmin=(0.37, -1.21, 0.53), dx=(0.05, 0.07, 0.06))


def _bc(faces):
    return None if faces is None else BoundaryConditions(
        faces, dirichlet_values=BC_KW["dirichlet"], robin_values=BC_KW["robin"],
        parities=BC_KW["parities"], xmin=BC_KW["xmin"], dx=BC_KW["dx"])


def _codes(faces):
    return [[BC_KINDS[faces[f][d]] for d in range(3)] for f in range(2)]


def _coarse_global(it, g, faces):
    """The global coarse array as the fill must see it: boundary values from
    the conditions, or the array itself without them."""
    if faces is None:
        return g
    cdom, _ = domains(it)
    glo = tuple(cdom.lo[d] - CG for d in range(3))
    gb = g.copy()
    sub = np.ascontiguousarray(gb[1:1 + NCOMP])
    periodic_fill(sub, glo, cdom.lo, cdom.hi, _periodic_dirs(faces), it)
    apply_bc(sub, glo, cdom.lo, cdom.hi, _codes(faces), it, **BC_KW)
    gb[1:1 + NCOMP] = sub
    return gb


def _prolongated(op_args, it, gc, fab):
    """The oracle's prolongation of the whole fine FAB box from the global
    coarse array gc."""
    cdom, _ = domains(it)
    glo = tuple(cdom.lo[d] - CG for d in range(3))
    ghi = tuple(cdom.hi[d] + CG for d in range(3))
    ch = oracle.Fab(glo, ghi, NCOMP, np.ascontiguousarray(gc[1:1 + NCOMP]))
    b = fab.box
    fh = oracle.Fab(b.lo, b.hi, NCOMP, np.zeros((NCOMP,) + tuple(reversed(b.size))))
    oracle.prolongate(oracle.lookup(*op_args), ch, fh, b.lo, b.hi)
    return fh.arr


def _mask(b, lo, hi):
    idx = np.meshgrid(*[np.arange(b.lo[d], b.hi[d] + 1) for d in (2, 1, 0)], indexing="ij")
    m = np.ones(idx[0].shape, dtype=bool)
    for a, d in zip(idx, (2, 1, 0)):
        m &= (a >= lo[d]) & (a <= hi[d])
    return m


def _run_all_ways(wl, fabs, originals, want, tag):
    """Batched, per-patch and two graph replays, each from the original
    fine data, compared with want within 1e-13."""
    s = torch.cuda.Stream()

    def reset():
        for f, o in zip(fabs, originals):
            f.t.copy_(o)
        torch.cuda.synchronize()

    def check(how):
        for f, w in zip(fabs, want):
            np.testing.assert_allclose(f.t.cpu().numpy(), w, rtol=1e-13, atol=1e-13,
                                       err_msg=str((how, tag)))

    with torch.cuda.stream(s):
        wl.upload(s)
        wl.run(s)
    torch.cuda.synchronize()
    check("batched")
    reset()
    with torch.cuda.stream(s):
        wl.run_per_patch(s)
    torch.cuda.synchronize()
    check("per patch")
    with torch.cuda.stream(s):
        graph = Graph(lambda: wl.run(s), s)
    for rep in range(2):
        reset()
        with torch.cuda.stream(s):
            graph.launch(s)
        torch.cuda.synchronize()
        check(("graph", rep))
    graph.close()


@pytest.mark.parametrize("op_args,cfaces,ffaces", REGRID_CASES)
def test_new_level_matches_oracle(op_args, cfaces, ffaces):
    """cxp_worklist_append_new_level (FillPatch_NewLevel): every fine FAB's
    box within the fine domain is prolongated, valid points included;
    points outside the fine domain are left alone or, with the fine level's
    conditions appended (cxp_worklist_append_boundaries), take their
    boundary values after the fill, as the oracle's conditions give them on
    each prolongated FAB."""
    it = op_args[2]
    rng = np.random.default_rng(11 + sum(it) + op_args[1])
    cdom, fdom = domains(it)
    gshape = (NC_CRSE,) + tuple(cdom.size[d] + 2 * CG for d in (2, 1, 0))
    g = rng.uniform(-1, 1, gshape)
    crse = coarse_level(it, g, poison_outside=cfaces is not None)
    fine = fine_level(it, rng, boxes=NEW_VALID, fill_valid=False)
    originals = [f.t.clone() for f in fine.fabs]
    gc = _coarse_global(it, g, cfaces)
    # the new level's domain, grown by the ghost width in periodic directions
    per = () if cfaces is None else _periodic_dirs(cfaces)
    nlo = tuple(fdom.lo[d] - (FG if d in per else 0) for d in range(3))
    nhi = tuple(fdom.hi[d] + (FG if d in per else 0) for d in range(3))
    want = []
    for fab, orig in zip(fine.fabs, originals):
        p = _prolongated(op_args, it, gc, fab)
        w = np.where(_mask(fab.box, nlo, nhi)[None], p, orig.cpu().numpy())
        if ffaces is not None:
            w = np.ascontiguousarray(w)
            apply_bc(w, fab.box.lo, fdom.lo, fdom.hi, _codes(ffaces), it, **BC_KW)
        want.append(w)
    wl = WorkList()
    wl.append_new_level(get_interpolator(*op_args), crse, 1, fine, 0, NCOMP, bc=_bc(cfaces))
    if ffaces is not None:
        wl.append_boundaries(fine, it, 0, NCOMP, _bc(ffaces))
    _run_all_ways(wl, fine.fabs, originals, want, ("new level", op_args))
    wl.close()


@pytest.mark.parametrize("op_args,cfaces,ffaces", REGRID_CASES)
def test_remake_level_matches_oracle(op_args, cfaces, ffaces):
    """cxp_worklist_append_remake_level (FillPatch_RemakeLevel): in each new
    fine FAB the old fine level's valid data where an old valid box covers
    the point, the oracle's prolongation elsewhere within the fine domain
    grown by the ghost width, the original value outside it; then the fine
    level's conditions when appended."""
    it = op_args[2]
    rng = np.random.default_rng(23 + sum(it) + op_args[1])
    cdom, fdom = domains(it)
    gshape = (NC_CRSE,) + tuple(cdom.size[d] + 2 * CG for d in (2, 1, 0))
    g = rng.uniform(-1, 1, gshape)
    crse = coarse_level(it, g, poison_outside=cfaces is not None)
    old = fine_level(it, rng)
    fine = fine_level(it, rng, boxes=NEW_VALID, fill_valid=False)
    originals = [f.t.clone() for f in fine.fabs]
    gc = _coarse_global(it, g, cfaces)
    want = []
    for fab, orig in zip(fine.fabs, originals):
        b = fab.box
        w = orig.cpu().numpy().copy()
        grown = (tuple(fdom.lo[d] - FG for d in range(3)), tuple(fdom.hi[d] + FG for d in range(3)))
        p = _prolongated(op_args, it, gc, fab)
        w = np.where(_mask(b, *grown)[None], p, w)
        for ofab, ov in zip(old.fabs, old.valid):
            m = _mask(b, ov.lo, ov.hi)
            if not m.any():
                continue
            # the old FAB's values at the same global points
            x = ofab.t.cpu().numpy()
            lo = [max(b.lo[d], ov.lo[d]) for d in range(3)]
            hi = [min(b.hi[d], ov.hi[d]) for d in range(3)]
            src = x[(slice(None),) + tuple(slice(lo[d] - ofab.box.lo[d], hi[d] - ofab.box.lo[d] + 1)
                                           for d in (2, 1, 0))]
            w[(slice(None),) + tuple(slice(lo[d] - b.lo[d], hi[d] - b.lo[d] + 1)
                                     for d in (2, 1, 0))] = src
        if ffaces is not None:
            w = np.ascontiguousarray(w)
            apply_bc(w, b.lo, fdom.lo, fdom.hi, _codes(ffaces), it, **BC_KW)
        want.append(w)
    wl = WorkList()
    wl.append_remake_level(get_interpolator(*op_args), crse, 1, old, fine, 0, NCOMP,
                           bc=_bc(cfaces))
    if ffaces is not None:
        wl.append_boundaries(fine, it, 0, NCOMP, _bc(ffaces))
    _run_all_ways(wl, fine.fabs, originals, want, ("remake", op_args))
    wl.close()


PER_FACES = [["periodic", "reflection", "neumann"], ["periodic", "linear_extrapolation", "dirichlet"]]


def _covering(b, boxes, shifts):
    """For each point of box b: the index into (shift, box) pairs of the
    first image covering it (shifts in the given order, boxes within a
    shift), or -1."""
    idx = np.meshgrid(*[np.arange(b.lo[d], b.hi[d] + 1) for d in (2, 1, 0)], indexing="ij")
    out = np.full(idx[0].shape, -1)
    k = 0
    pairs = []
    for sh in shifts:
        for bi, v in enumerate(boxes):
            m = np.ones(idx[0].shape, dtype=bool)
            for a, d in zip(idx, (2, 1, 0)):
                m &= (a >= v.lo[d] + sh[d]) & (a <= v.hi[d] + sh[d])
            out[(out < 0) & m] = k
            pairs.append((sh, bi))
            k += 1
    return out, pairs


@pytest.mark.parametrize("op_args", [("natural", 3, (1, 1, 1)), ("natural", 5, (0, 0, 0))])
def test_periodic_fine_images(op_args):
    """x periodic on both levels, fine boxes at both x faces, so that each
    box's x ghosts lie in the other's periodic image.  Ghost fill: those
    ghosts are left alone (the same-level sync fills them in the
    reference).  Remake: new-level points covered by an old valid box or
    by its periodic image take the old data (the unshifted box first),
    the rest of the grown domain is prolongated."""
    it = op_args[2]
    rng = np.random.default_rng(31 + sum(it))
    cdom, fdom = domains(it)
    per = fdom.hi[0] - fdom.lo[0] + it[0]
    shifts = [(0, 0, 0), (-per, 0, 0), (per, 0, 0)]
    gshape = (NC_CRSE,) + tuple(cdom.size[d] + 2 * CG for d in (2, 1, 0))
    g = rng.uniform(-1, 1, gshape)
    crse = coarse_level(it, g, poison_outside=True)
    gc = _coarse_global(it, g, PER_FACES)
    grown = (tuple(fdom.lo[d] - FG for d in range(3)), tuple(fdom.hi[d] + FG for d in range(3)))
    boxes = [((0, 0, 0), (9, 9, 9)), ((38, 0, 0), (fdom.hi[0], 9, 9))]
    op = get_interpolator(*op_args)
    # ghost fill
    fine = fine_level(it, rng, boxes=boxes)
    originals = [f.t.clone() for f in fine.fabs]
    want = []
    for f, o in zip(fine.fabs, originals):
        cov, _ = _covering(f.box, fine.valid, shifts)
        p = _prolongated(op_args, it, gc, f)
        m = _mask(f.box, *grown) & (cov < 0)
        want.append(np.where(m[None], p, o.cpu().numpy()))
    assert any((_covering(f.box, fine.valid, shifts)[0] >= 3).any() for f in fine.fabs)
    wl = WorkList()
    wl.append_level(op, crse, 1, fine, 0, NCOMP, bc=_bc(PER_FACES))
    _run_all_ways(wl, fine.fabs, originals, want, ("periodic ghosts", op_args))
    wl.close()
    # remake from that level onto shifted boxes
    old = fine
    new_boxes = [((0, 2, 0), (11, 11, 9)), ((36, 0, 0), (fdom.hi[0], 9, 11))]
    new = fine_level(it, rng, boxes=new_boxes, fill_valid=False)
    originals = [f.t.clone() for f in new.fabs]
    old_np = [f.t.cpu().numpy() for f in old.fabs]
    want = []
    for f, o in zip(new.fabs, originals):
        b = f.box
        cov, pairs = _covering(b, old.valid, shifts)
        p = _prolongated(op_args, it, gc, f)
        w = np.where(_mask(b, *grown)[None], p, o.cpu().numpy())
        idx = np.meshgrid(*[np.arange(b.lo[d], b.hi[d] + 1) for d in (2, 1, 0)], indexing="ij")
        for k, (sh, bi) in enumerate(pairs):
            m = cov == k
            if not m.any():
                continue
            ob = old.fabs[bi].box
            zz, yy, xx = (idx[0][m] - sh[2] - ob.lo[2], idx[1][m] - sh[1] - ob.lo[1],
                          idx[2][m] - sh[0] - ob.lo[0])
            w[:, m] = old_np[bi][:, zz, yy, xx]
        want.append(w)
    assert any((_covering(f.box, old.valid, shifts)[0] >= 2).any() for f in new.fabs)
    wl = WorkList()
    wl.append_remake_level(op, crse, 1, old, new, 0, NCOMP, bc=_bc(PER_FACES))
    _run_all_ways(wl, new.fabs, originals, want, ("periodic remake", op_args))
    wl.close()


@pytest.mark.parametrize("op_args", [("eno", 3, (1, 1, 1)), ("poly-cons3lfb", 3, (1, 0, 1)),
                                     ("poly-eno3lfb", 3, (1, 1, 1))])
def test_level_fill_per_group_matches_oracle(op_args):
    """cxp_worklist_set_per_group(1): the list's ENO and fallback patches
    take the per-group decisions (one shift and one fallback choice for all
    components), as the reference's prolongate_per_group parameter selects;
    batched, per-patch and graph runs equal the oracle's per-group
    prolongation bit for bit, and switching back to per-variable gives the
    per-variable values."""
    it = op_args[2]
    rng = np.random.default_rng(41 + sum(it))
    cdom, _ = domains(it)
    gshape = (NC_CRSE,) + tuple(cdom.size[d] + 2 * CG for d in (2, 1, 0))
    g = rng.uniform(-1, 1, gshape)
    crse = coarse_level(it, g)
    fine = fine_level(it, rng)
    originals = [f.t.clone() for f in fine.fabs]
    want_pg = expected(op_args, it, g, fine, per_group=True)
    want_pv = expected(op_args, it, g, fine, per_group=False)
    assert any(not np.array_equal(a, b) for a, b in zip(want_pg, want_pv))
    wl = WorkList()
    wl.append_level(get_interpolator(*op_args), crse, 1, fine, 0, NCOMP)
    wl.set_per_group(True)
    _run_all_ways(wl, fine.fabs, originals, want_pg, ("per group", op_args))
    wl.set_per_group(False)
    _run_all_ways(wl, fine.fabs, originals, want_pv, ("per var", op_args))
    wl.close()


def test_boundary_sources_in_two_coarse_fabs():
    """Coarse FABs split at x = 1: linear extrapolation at the lower x face
    reads the edge (x = 0) and the point beside it (x = 1) from two FABs,
    so those boundary values are computed from the gathered patch after the
    gather; the others read the coarse FABs directly in the gather launch.
    Both equal the oracle."""
    op_args = ("natural", 3, (1, 1, 1))
    faces = [["linear_extrapolation", "reflection", "neumann"],
             ["dirichlet", "linear_extrapolation", "reflection"]]
    it = op_args[2]
    rng = np.random.default_rng(57)
    cdom, _ = domains(it)
    gshape = (NC_CRSE,) + tuple(cdom.size[d] + 2 * CG for d in (2, 1, 0))
    g = rng.uniform(-1, 1, gshape)
    crse = coarse_level(it, g, poison_outside=True, split=(1, 12, 12))
    fine = fine_level(it, rng)
    originals = [f.t.clone() for f in fine.fabs]
    want = expected(op_args, it, _coarse_global(it, g, faces), fine)
    wl = WorkList()
    wl.append_level(get_interpolator(*op_args), crse, 1, fine, 0, NCOMP, bc=_bc(faces))
    _run_all_ways(wl, fine.fabs, originals, want, ("two-FAB sources", op_args))
    wl.close()


def test_new_level_placed_by_boxinbox_tags():
    """The regrid sequence end to end: box-in-box tags on each coarse FAB's
    valid box (BoxInBox_Setup), ErrorEst's threshold test per box, the
    union of the tagged cells' bounding boxes grown by one coarse cell and
    refined (a box of tags clusters to its bounding box), then
    FillPatch_NewLevel on that box, against the oracle.  Coarse cells of
    [0, 1]^3 (N = 24, h = 1/24), a cube of radius 1/4 about the centre:
    |i + 1/2 - 12| <= 6 - 1 tags cells 7..16, half a cell from either
    decision, so the box is [7, 16]^3 whatever the rounding."""
    from carpetx_amd.tagging import BoxInBoxRegion, boxinbox_setup, error_est, refined_box
    op_args = ("natural", 3, (1, 1, 1))
    it = op_args[2]
    rng = np.random.default_rng(23)
    cdom, fdom = domains(it)
    g = rng.uniform(-1, 1, (NC_CRSE,) + tuple(cdom.size[d] + 2 * CG for d in (2, 1, 0)))
    crse = coarse_level(it, g)
    h = 1.0 / N
    cube = BoxInBoxRegion("cube", num_levels=2, position=(0.5, 0.5, 0.5), radius=[1.0, 0.25])
    total, lo, hi = 0, [1 << 30] * 3, [-(1 << 30)] * 3
    for v in crse.valid:
        err = FArrayBox(v, 1)
        boxinbox_setup([cube], 1, (h / 2,) * 3, (h,) * 3, err)
        count, tb = error_est(err, 1.0)
        if count:
            total += count
            lo = [min(lo[d], tb.lo[d]) for d in range(3)]
            hi = [max(hi[d], tb.hi[d]) for d in range(3)]
    assert total == 10 ** 3 and Box(lo, hi) == Box((7,) * 3, (16,) * 3)
    new = refined_box(Box(lo, hi))
    assert new == Box((12,) * 3, (35,) * 3)
    fine = fine_level(it, rng, boxes=[(new.lo, new.hi)], fill_valid=False)
    originals = [f.t.clone() for f in fine.fabs]
    want = [_prolongated(op_args, it, g, fab) for fab in fine.fabs]
    assert fdom.contains(fine.fabs[0].box)
    wl = WorkList()
    wl.append_new_level(get_interpolator(*op_args), crse, 1, fine, 0, NCOMP)
    _run_all_ways(wl, fine.fabs, originals, want, ("new level from tags", op_args))
    wl.close()
