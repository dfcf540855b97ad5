"""Pins of the oracle's box-in-box tags (oracle/tagging.py), CPU only.

The grids are dyadic: dx = 1/8, cell centres x0 + i dx = 1/16 + i/8 and
radii r with r - dx a power of two, so every coordinate, difference and
quotient is exact and the refined cells follow from integer arithmetic
alone: with u = 2 i - 7 (odd), (X - 1/2) / (1/4) = u / 4.  A wrong centre
(i dx without x0, or (i + 1/2) dx), a missing "- dx", a squared radius, a
wrong fold or a transposed axis changes the sets below."""
import itertools

import numpy as np
import pytest

from oracle.tagging import CUBE, SPHERE, boxinbox_tag, error_est

X0, DX = (1 / 16,) * 3, (1 / 8,) * 3
LO, HI = (0, 0, 0), (7, 7, 7)


def region(shape, pos=(0.5, 0.5, 0.5), radius=(0.375, 0.375, 0.375), active=1,
           num_levels=2):
    return dict(active=active, shape=shape, num_levels=num_levels, position=pos,
                radius=radius)


def cells(arr, lo=LO):
    return {(i + lo[0], j + lo[1], k + lo[2])
            for k, j, i in zip(*np.nonzero(arr))}


def test_cube_holds_the_cells_within_radius_minus_dx():
    # Rmax = 3/8 - 1/8 = 1/4: |u| / 4 <= 1 for u = 2i - 7, i.e. i = 2..5
    t = boxinbox_tag([region(CUBE)], 1, X0, DX, LO, HI)
    want = set(itertools.product(range(2, 6), repeat=3))
    assert cells(t) == want and len(want) == 64
    assert set(np.unique(t)) == {0.0, 1.0}


def test_sphere_is_the_integer_ball():
    t = boxinbox_tag([region(SPHERE)], 1, X0, DX, LO, HI)
    want = {(i, j, k) for i, j, k in itertools.product(range(8), repeat=3)
            if (2 * i - 7) ** 2 + (2 * j - 7) ** 2 + (2 * k - 7) ** 2 <= 16}
    # u in {+-1, +-3}: all |u| = 1 (8 cells), or one |u| = 3 (24 cells)
    assert len(want) == 32
    assert cells(t) == want


def test_per_direction_radii_and_a_shifted_box():
    # Rmax = (1/4, 1/8, 3/8): i in 2..5, j in 3..4, k in 1..6; the box from
    # (-2, 1, 0) only sees part of that
    r = region(CUBE, radius=(0.375, 0.25, 0.5))
    lo, hi = (-2, 1, 0), (4, 6, 7)
    t = boxinbox_tag([r], 1, X0, DX, lo, hi)
    assert t.shape == (8, 6, 7)
    want = set(itertools.product(range(2, 5), range(3, 5), range(1, 7)))
    assert cells(t, lo) == want


def test_union_and_regions_that_do_not_apply():
    a = region(CUBE, pos=(0.3125, 0.5, 0.5), radius=(0.25, 0.25, 0.25))
    # a: Rmax 1/8 about 5/16 (= centre of cell 2): |X - 5/16| <= 1/8 -> i = 1..3
    b = region(SPHERE, pos=(0.8125, 0.8125, 0.8125), radius=(0.25, 0.25, 0.25))
    # b: Rmax 1/8 about the centre of cell 6: only cells at distance <= 1/8
    # (the centre and its 6 neighbours)
    inactive = region(CUBE, radius=(1.0, 1.0, 1.0), active=0)
    too_deep = region(CUBE, radius=(1.0, 1.0, 1.0), num_levels=1)
    t = boxinbox_tag([a, b, inactive, too_deep], 1, X0, DX, LO, HI)
    wa = set(itertools.product(range(1, 4), range(3, 5), range(3, 5)))
    wb = {(6, 6, 6)} | {tuple(6 + (d == e) * s for e in range(3))
                        for d in range(3) for s in (-1, 1)}
    assert cells(t) == wa | wb
    # at level 0 the too-deep region applies and covers everything
    t0 = boxinbox_tag([too_deep], 0, X0, DX, LO, HI)
    assert t0.all()


def test_non_positive_radius_only_where_the_region_applies():
    bad = region(CUBE, radius=(0.5, 0.0, 0.5))
    with pytest.raises(ValueError):
        boxinbox_tag([bad], 1, X0, DX, LO, HI)
    bad["active"] = 0
    assert not boxinbox_tag([bad], 1, X0, DX, LO, HI).any()
    bad["active"], bad["num_levels"] = 1, 1
    assert not boxinbox_tag([bad], 1, X0, DX, LO, HI).any()


def test_radius_of_one_cell_refines_nothing():
    # Rmax = 0: q is +-inf, or NaN (0 / 0) at a cell centre on the position;
    # neither passes "<= 1", for the sphere or for max1's NaN propagation
    for shape in (SPHERE, CUBE):
        r = region(shape, pos=(7 / 16, 7 / 16, 7 / 16), radius=(0.125,) * 3)
        assert not boxinbox_tag([r], 1, X0, DX, LO, HI).any()


def test_error_est_threshold_count_and_bounding_box():
    rng = np.random.default_rng(5)
    lo = (-3, 2, 1)
    err = rng.uniform(0, 1, (4, 5, 6))
    err[0, 0, 0] = 0.75  # exactly at the threshold: set (>=)
    tags, count, box = error_est(err, lo, 0.75)
    assert tags.dtype == np.int8 and tags[0, 0, 0] == 2  # amrex::TagBox::SET
    assert set(np.unique(tags)) <= {0, 2}
    assert count == int((err >= 0.75).sum())
    pts = [(i + lo[0], j + lo[1], k + lo[2]) for k, j, i in zip(*np.nonzero(err >= 0.75))]
    assert box == (tuple(min(p[d] for p in pts) for d in range(3)),
                   tuple(max(p[d] for p in pts) for d in range(3)))
    tags, count, box = error_est(err, lo, 2.0)
    assert count == 0 and box is None and not tags.any()
