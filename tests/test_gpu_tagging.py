"""GPU box-in-box tags (cxp_boxinbox_tag) and ErrorEst's threshold test
(cxp_error_est) against the oracle (oracle/tagging.py), bit for bit.

Expected values: the oracle evaluates the same IEEE expressions in numpy on
the same parameters; the tags are amrex::TagBox's CLEAR = 0 / SET = 2, so any
difference in a decision
shows.  FAB points outside the tagged box hold a sentinel the kernels must
not touch."""
import numpy as np
import pytest
import torch

from carpetx_amd.interp import Box, FArrayBox
from carpetx_amd.tagging import BoxInBoxRegion, boxinbox_setup, error_est, refined_box
from oracle import tagging as ot

pytestmark = pytest.mark.gpu

SENTINEL = -7.25


def oracle_regions(regions, level):
    return [dict(active=r.active, shape=0 if r.shape == "sphere" else 1,
                 num_levels=r.num_levels, position=r.position,
                 radius=r.radii(level)) for r in regions]


def run_case(regions, level, x0, dx, fab_box, box):
    err = FArrayBox(fab_box, 1, device="cuda",
                    tensor=torch.full((1,) + tuple(reversed(fab_box.size)), SENTINEL,
                                      dtype=torch.float64, device="cuda"))
    boxinbox_setup(regions, level, x0, dx, err, box)
    got = err.t[0].cpu().numpy()
    want = ot.boxinbox_tag(oracle_regions(regions, level), level, x0, dx, box.lo, box.hi)
    sl = tuple(slice(box.lo[d] - fab_box.lo[d], box.hi[d] - fab_box.lo[d] + 1)
               for d in (2, 1, 0))
    assert np.array_equal(got[sl], want)
    outside = np.ones(got.shape, dtype=bool)
    outside[sl] = False
    assert np.all(got[outside] == SENTINEL)
    return err, want


CASES = [
    # random, non-dyadic geometry: three regions, both shapes, per-direction radii
    ("mixed", [BoxInBoxRegion("sphere", num_levels=3, position=(0.31, -0.12, 0.07),
                              radius=[1.0, 0.37, 0.2]),
               BoxInBoxRegion("cube", num_levels=2, position=(-0.4, 0.35, -0.05),
                              radius=[1.0, 0.21], radius_y=[-1.0, 0.13]),
               BoxInBoxRegion("cube", active=False, radius=[1.0, 1.0])],
     1, (-0.981, -1.0, -0.9437), (0.0271, 0.031, 0.0293),
     Box((-3, 2, -4), (70, 66, 71)), Box((0, 4, -1), (67, 63, 68))),
    # level 2: only the sphere applies, with its smaller radius
    ("deeper", [BoxInBoxRegion("sphere", num_levels=3, position=(0.31, -0.12, 0.07),
                               radius=[1.0, 0.37, 0.2]),
                BoxInBoxRegion("cube", num_levels=2, radius=[1.0, 0.21])],
     2, (-0.5, -0.5, -0.5), (1 / 96, 1 / 96, 1 / 96),
     Box((0, 0, 0), (95, 95, 95)), Box((0, 0, 0), (95, 95, 95))),
    # radius of one cell (Rmax = 0, q infinite or NaN) and a thin box
    ("zero_rmax", [BoxInBoxRegion("cube", num_levels=2, position=(0.0, 0.0, 0.0),
                                  radius=[1.0, 0.125])],
     1, (-0.4375, -0.4375, -0.4375), (0.125, 0.125, 0.125),
     Box((0, 0, 0), (7, 7, 7)), Box((0, 3, 0), (7, 3, 7))),
]


@pytest.mark.parametrize("name,regions,level,x0,dx,fab_box,box", CASES,
                         ids=[c[0] for c in CASES])
def test_tags_match_the_oracle(name, regions, level, x0, dx, fab_box, box):
    print("case", name, flush=True)
    err, want = run_case(regions, level, x0, dx, fab_box, box)
    if name != "zero_rmax":
        assert 0 < want.sum() < want.size  # the case decides both ways
    # the threshold test on the same FAB, over the same box
    tags = torch.full(tuple(err.t.shape), 5, dtype=torch.int8, device="cuda")
    count, tb = error_est(err, 1.0, region=box, tags=tags)
    wtags, wcount, wbox = ot.error_est(want, box.lo, 1.0)
    assert count == wcount
    assert (tb is None and wbox is None) or (tb.lo, tb.hi) == wbox
    t = tags[0].cpu().numpy()
    sl = tuple(slice(box.lo[d] - fab_box.lo[d], box.hi[d] - fab_box.lo[d] + 1)
               for d in (2, 1, 0))
    assert np.array_equal(t[sl], wtags)
    outside = np.ones(t.shape, dtype=bool)
    outside[sl] = False
    assert np.all(t[outside] == 5)


def test_error_est_thresholds_and_box_reuse():
    """Random data, several thresholds, one box_out reused across calls (the
    call initialises it each time), the default box_out, a sub-box."""
    g = torch.Generator(device="cpu").manual_seed(7)
    fab_box = Box((-2, -3, 1), (61, 40, 33))
    data = torch.rand((1,) + tuple(reversed(fab_box.size)), generator=g, dtype=torch.float64)
    err = FArrayBox(fab_box, 1, device="cuda", tensor=data.cuda())
    box_out = torch.empty(7, dtype=torch.int32, device="cuda")
    arr = data[0].numpy()
    for thr, sub in [(0.5, fab_box), (0.999, Box((0, 0, 2), (50, 30, 30))),
                     (2.0, fab_box), (0.9995, fab_box)]:
        count, tb = error_est(err, thr, region=sub, box_out=box_out)
        sl = tuple(slice(sub.lo[d] - fab_box.lo[d], sub.hi[d] - fab_box.lo[d] + 1)
                   for d in (2, 1, 0))
        _, wcount, wbox = ot.error_est(arr[sl], sub.lo, thr)
        assert count == wcount, thr
        assert (tb is None and wbox is None) or (tb.lo, tb.hi) == wbox, thr
    assert error_est(err, 0.0)[0] == fab_box.numPts


def test_full_size_cube():
    """256^3 coarse cells (a 134 MB FAB), a cube of radius 1/4 about the
    centre on the dyadic grid: the tagged cells are exactly [65, 190]^3,
    counted and boxed on the device."""
    n = 256
    h = 1.0 / n
    err = FArrayBox(Box((0, 0, 0), (n - 1,) * 3), 1, device="cuda")
    r = BoxInBoxRegion("cube", num_levels=2, radius=[1.0, 0.25])
    boxinbox_setup([r], 1, (-0.5 + h / 2,) * 3, (h,) * 3, err)
    count, tb = error_est(err, 1.0)
    # cell centre (i + 1/2) h - 1/2, Rmax = 1/4 - h: |i + 1/2 - 128| <= 63
    # -> i in [65, 190]
    assert tb == Box((65,) * 3, (190,) * 3)
    assert count == 126 ** 3
    assert float(err.t.sum()) == 126 ** 3
    assert refined_box(tb) == Box((128,) * 3, (383,) * 3)


def test_wave_harness_places_its_fine_level_from_tags():
    from carpetx_amd.wave import TwoLevelWave
    for n in (24, 32, 128):
        w = TwoLevelWave(n=n)
        q = n // 4
        assert w.fine_int == Box((2 * q,) * 3, (6 * q,) * 3), n
        assert w.crse_under == Box((q,) * 3, (3 * q,) * 3), n
    # a smaller cube: radius 1/8 -> coarse cells within 1/8 - h of the centre
    w = TwoLevelWave(n=64, regions=[BoxInBoxRegion("cube", num_levels=2,
                                                   radius=[1.0, 0.125])])
    # |i + 1/2 - 32| <= 7 -> i in [25, 38]; grown -> [24, 39]; fine [48, 80]
    assert w.fine_int == Box((48,) * 3, (80,) * 3)
    sc, sf = w.gaussian()
    sc, sf = w.step(sc, sf)
    assert torch.isfinite(sf).all() and torch.isfinite(sc).all()


def test_error_est_tags_without_result_words():
    """cxp_error_est with tags only (box_out null), through the C ABI."""
    import ctypes
    from carpetx_amd import _lib
    fab_box = Box((3, -2, 0), (40, 29, 17))
    data = torch.rand((1,) + tuple(reversed(fab_box.size)), dtype=torch.float64,
                      generator=torch.Generator().manual_seed(3))
    err = FArrayBox(fab_box, 1, device="cuda", tensor=data.cuda())
    tags = torch.full(tuple(err.t.shape), 7, dtype=torch.int8, device="cuda")
    sub = Box((5, 0, 2), (33, 21, 15))
    st = _lib.lib().cxp_error_est(
        ctypes.c_void_p(err.t.data_ptr()), _lib.i3(fab_box.lo), _lib.i3(fab_box.hi),
        _lib.i3(sub.lo), _lib.i3(sub.hi), 0.25, ctypes.c_void_p(tags.data_ptr()),
        None, None)
    assert st == 0
    torch.cuda.synchronize()
    t = tags[0].cpu().numpy()
    sl = tuple(slice(sub.lo[d] - fab_box.lo[d], sub.hi[d] - fab_box.lo[d] + 1)
               for d in (2, 1, 0))
    want, _, _ = ot.error_est(data[0].numpy()[sl], sub.lo, 0.25)
    assert np.array_equal(t[sl], want)
    # the byte values are amrex::TagBox::SET (2) and CLEAR (0)
    assert set(np.unique(t[sl]).tolist()) == {0, 2}
    outside = np.ones(t.shape, dtype=bool)
    outside[sl] = False
    assert np.all(t[outside] == 7)
