"""Pins of the oracle's linear-fallback decisions against hand-derived values.

The fixtures in golden/fallback_kinks.json are 1-D coarse profiles along one
cell-centred axis (constant along the two vertex-centred axes) whose expected
fine values were worked out by hand with exact rationals
(prolongate_3d_rf2_impl.hxx:1535-1667 per-var, :1990-2137 per-group).  Each
case separates one decision rule: the slope test's start at r0 + 2, the fixed
s0 = c(1) - c(0) (also read outside a shifted ENO stencil), the range test on
its own, and the per-group exclusion of ENO directions.
"""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = json.load(open(os.path.join(HERE, "golden", "fallback_kinks.json")))["cases"]


def kink_fabs(case, axis, profiles):
    """Coarse FAB holding `profiles` (one per component) along `axis` at coarse
    indices 0..4, constant along the other axes; fine region: the point
    2*2 + off along `axis`, fine [0, 1] along the vertex-centred axes."""
    it = tuple(1 if d == axis else 0 for d in range(3))
    op = oracle.lookup(case["family"], case["order"], it)
    rlo = [0, 0, 0]
    rhi = [1, 1, 1]
    rlo[axis] = rhi[axis] = 4 + case["off"]
    clo, chi = oracle.coarse_box(op, rlo, rhi)
    assert clo[axis] >= 0 and chi[axis] <= 4
    lo = list(clo)
    hi = list(chi)
    lo[axis], hi[axis] = 0, 4
    shape = tuple(hi[d] - lo[d] + 1 for d in (2, 1, 0))
    data = np.empty((len(profiles),) + shape)
    for c, prof in enumerate(profiles):
        vals = np.array([float(Fraction(v)) for v in prof])
        bshape = [1, 1, 1]
        bshape[2 - axis] = 5
        data[c] = np.broadcast_to(vals.reshape(bshape), shape)
    crse = oracle.Fab(lo, hi, len(profiles), data)
    return op, crse, tuple(rlo), tuple(rhi)


def run(case, axis, profiles, per_group):
    op, crse, rlo, rhi = kink_fabs(case, axis, profiles)
    fine = oracle.Fab(rlo, rhi, len(profiles))
    oracle.prolongate(op, crse, fine, rlo, rhi, ncomp=len(profiles), per_group=per_group)
    return fine.arr


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_fallback_kink(case, axis):
    for per_group, key in ((False, "per_var"), (True, "per_group")):
        got = run(case, axis, [case["profile"]], per_group)
        want = float(Fraction(case[key]))
        assert np.all(got == want), (case["name"], axis, key, got.ravel(), want)


def test_fixtures_separate_the_rules():
    """Every rule the fixtures pin is exercised by a case whose two modes or
    whose high- and low-order values differ, so a mutation shows."""
    names = {c["name"]: c for c in CASES}
    assert names["eno_overshoot_falls_back_per_var_only"]["per_var"] != \
        names["eno_overshoot_falls_back_per_var_only"]["per_group"]
    assert names["eno_s0_read_outside_the_shifted_stencil"]["per_var"] != \
        names["eno_s0_read_outside_the_shifted_stencil"]["per_group"]
    # a nonzero shift whose value overshoots: a per-group fallback that took
    # ENO directions would change it (to the unshifted row or to CONS 1)
    c = names["eno_shifted_overshoot_falls_back_per_var_only"]
    assert c["per_var"] != c["per_group"] and c["per_group"] != "-1/8"


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_per_group_slope_test_spans_components(axis):
    """Per-group: one component's slope flip makes every component fall back
    (:2026-2116, need_fallback shared), per-var only that component.
    Component 0 alone keeps CONS 3 = 39/64 (linear 3/4); component 1 falls
    back to 1/4 by its slope test."""
    case = {"family": "poly-cons3lfb", "order": 3, "off": 0}
    comps = [["2", "0", "1", "2", "3"], ["0", "1", "0", "1", "2"]]
    v = run(case, axis, comps, per_group=False)
    assert np.all(v[0] == 39 / 64) and np.all(v[1] == 0.25)
    g = run(case, axis, comps, per_group=True)
    assert np.all(g[0] == 0.75) and np.all(g[1] == 0.25)
