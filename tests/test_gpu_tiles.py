"""Every tile shape of the separable kernels (one and two row pairs per
thread) and several z-chunks give the oracle's bits, per patch and in
lists: the shapes change only which workgroup computes a point, never the
arithmetic (cxp_tune_set forces each shape)."""
import ctypes

import numpy as np
import pytest
import torch

import oracle
from carpetx_amd import _lib
from carpetx_amd.batched import WorkList
from carpetx_amd.interp import Box, FArrayBox, get_interpolator

pytestmark = pytest.mark.gpu

OPS = [("interpolate", 5, (0, 0, 0)), ("conservative", 1, (1, 1, 1)),
       ("minmod", 1, (1, 1, 1)), ("natural", 3, (1, 0, 1))]


def _i3(v):
    return (ctypes.c_int * 3)(*v)


@pytest.fixture(autouse=True)
def clear_tuning():
    _lib.lib().cxp_tune_clear()
    yield
    _lib.lib().cxp_tune_clear()


def _case(typ, order, it, region, ncomp, seed):
    P = get_interpolator(typ, order, it)
    cb = P.CoarseBox(region)
    rng = np.random.default_rng(seed)
    u = rng.uniform(-1, 1, (ncomp,) + tuple(reversed(cb.size)))
    crse = FArrayBox(cb, ncomp, tensor=torch.from_numpy(u).cuda())
    oop = oracle.lookup(typ, order, it)
    ch = oracle.Fab(cb.lo, cb.hi, ncomp, u)
    fh = oracle.Fab(region.lo, region.hi, ncomp)
    oracle.prolongate(oop, ch, fh, region.lo, region.hi)
    return P, crse, fh.arr


@pytest.mark.parametrize("typ,order,it", OPS)
def test_every_shape_per_patch(typ, order, it):
    L = _lib.lib()
    region = Box((-3, 2, -5), (60, 45, 30))     # ragged, odd corners
    P, crse, expect = _case(typ, order, it, region, 2, 11)
    n = L.cxp_tile_config_count()
    assert n == 9
    for cfg in range(n):
        for zc in (0, 6, 24):
            assert L.cxp_tune_set(P.handle, _i3(region.lo), _i3(region.hi), cfg, zc) == 0
            fine = FArrayBox(region, 2)
            P.interp(crse, 0, fine, 0, 2, region)
            torch.cuda.synchronize()
            assert np.array_equal(fine.t.cpu().numpy(), expect), (cfg, zc)


@pytest.mark.parametrize("typ,order,it", OPS[:2])
def test_every_shape_in_a_list(typ, order, it):
    """A list of one operator (the specialised kernel, two-pair shapes
    included) and a mixed list (the dispatching kernel, which keeps to one
    row pair: two-pair settings fall back to the one-pair shape of the same
    width)."""
    L = _lib.lib()
    regions = [Box((0, 0, 0), (33, 29, 17)), Box((-6, 4, 2), (25, 41, 30))]
    cases = [_case(typ, order, it, r, 1, 20 + i) for i, r in enumerate(regions)]
    other = ("natural", 3, (0, 1, 0))
    oreg = Box((2, -2, 0), (37, 21, 19))
    P2, crse2, expect2 = _case(*other, oreg, 1, 30)
    for mixed in (False, True):
        for cfg in range(L.cxp_tile_config_count()):
            for r in regions:
                assert L.cxp_tune_set(cases[0][0].handle, _i3(r.lo), _i3(r.hi), cfg, 8) == 0
            wl = WorkList()
            fines = []
            for (P, crse, _), r in zip(cases, regions):
                f = FArrayBox(r, 1)
                fines.append(f)
                wl.append(P, crse, 0, f, 0, 1, r)
            if mixed:
                f2 = FArrayBox(oreg, 1)
                wl.append(P2, crse2, 0, f2, 0, 1, oreg)
            wl.upload()
            wl.run()
            torch.cuda.synchronize()
            for f, (_, _, e) in zip(fines, cases):
                assert np.array_equal(f.t.cpu().numpy(), e), (mixed, cfg)
            if mixed:
                assert np.array_equal(f2.t.cpu().numpy(), expect2), cfg
            wl.close()


@pytest.mark.parametrize("typ,order,it", OPS + [("eno", 3, (1, 1, 1))])
def test_regions_in_one_call_equal_region_by_region(typ, order, it):
    """cxp_prolongate_regions (a ghost shell's six slabs; then one whole box,
    whose z-chunks the launch shortens) gives cxp_prolongate's bits region by
    region, also with a two-row-pair shape in the tile table (the one-launch
    kernel takes the one-pair shape of the same width) and for an operator
    that is not separable (region by region inside)."""
    from carpetx_amd.workloads import shell_slabs
    L = _lib.lib()
    P = get_interpolator(typ, order, it)
    interior = Box((4, 2, 6), (43, 37, 29))
    fbox = interior.grow(3)
    cb = P.CoarseBox(fbox)
    rng = np.random.default_rng(13)
    crse = FArrayBox(cb, 2, tensor=torch.from_numpy(
        rng.uniform(-1, 1, (2,) + tuple(reversed(cb.size)))).cuda())
    for regions in (shell_slabs(fbox, interior), [interior]):
        for tuned in (False, True):
            L.cxp_tune_clear()
            if tuned:
                for r in regions:
                    assert L.cxp_tune_set(P.handle, _i3(r.lo), _i3(r.hi), 5, 0) == 0
            want = FArrayBox(fbox, 2)
            for r in regions:
                P.interp(crse, 0, want, 0, 2, r)
            got = FArrayBox(fbox, 2)
            P.interp_regions(crse, 0, got, 0, 2, regions)
            torch.cuda.synchronize()
            assert torch.equal(got.t, want.t), (typ, len(regions), tuned)


@pytest.mark.parametrize("typ,order,it", [("interpolate", 5, (0, 0, 0)),
                                          ("conservative", 1, (1, 1, 1)),
                                          ("natural", 3, (1, 0, 1))])
def test_adjoint_regions_equal_region_by_region(typ, order, it):
    """cxp_prolongate_adjoint_regions over a ghost shell's six slabs (lo/hi
    pairs apart: shared launches), over the shell of a box too thin for
    that (windows that meet: one region after another) and over one whole
    box gives the region-by-region adjoints' sums bit for bit."""
    from carpetx_amd.workloads import shell_slabs
    P = get_interpolator(typ, order, it)
    rng = np.random.default_rng(17)
    for interior in (Box((4, 2, 6), (43, 37, 29)), Box((2, 4, 0), (9, 11, 5))):
        fbox = interior.grow(3)
        cb = P.CoarseBox(fbox)
        fbar = FArrayBox(fbox, 2, tensor=torch.from_numpy(
            rng.uniform(-1, 1, (2,) + tuple(reversed(fbox.size)))).cuda())
        for regions in (shell_slabs(fbox, interior), [interior]):
            want = FArrayBox(cb, 2)
            for r in regions:
                P.interp_adjoint(want, 0, fbar, 0, 2, r)
            got = FArrayBox(cb, 2)
            P.interp_adjoint_regions(got, 0, fbar, 0, 2, regions)
            torch.cuda.synchronize()
            assert torch.equal(got.t, want.t), (typ, interior, len(regions))
