"""The oracle against CarpetX's own TestProlongate golden outputs.

tests/golden/testprolongate_max_error.json holds, for each of the 19
TestProlongate cases (TestProlongate/test/*.par), the operator parameters
and the max_error scalar of every centring gfXYZ at iterations 0..2; it is
written by tests/golden/extract_testprolongate.py.

* Every (case, centring) whose golden max_error is at rounding level is
  re-run here: TestProlongate's own test function (test.cxx:43-73: per
  direction sum_{p<=order} x^p, cell averages in conservative directions)
  with its per-family order mapping (test.cxx:86-150) is sampled on the
  test's coarse grid (30 cells on [-1,1] shifted by 0.1234567 / 0.2345678 /
  0.3456789), prolongated by the oracle with the same family, order,
  centring and per-group flag, and compared with the same function on the
  fine grid.  This pins that each registry entry is exact for exactly the
  polynomials the reference's own test expects.
* test_natural_o7's cell-centred rows are not exact: 2.13623046869e7 =
  (35/16384) * (1e10 - f), the outer weight of the order-7 CC conservative
  stencil times the Dirichlet ghost value 1e10 (TestProlongate/
  interface.ccl:13-20) that the stencil reaches near the boundary.  The
  oracle reproduces that mechanism.
* Not pinned here: the cc_nc_o3 cell-centred rows (~1e-3, from restricting
  point values by averaging and prolongating again on the reference's AMReX
  grid) and the poly-cons3lfb rows (the linear fallback next to the unit
  jump at 0, masked by is_near_jump).  Both depend on the fine boxes AMReX's
  regridding produces, which this image cannot run.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "testprolongate_max_error.json")))
EXACT = 1e-14          # golden rows at rounding level
NC = 30                # coarse cells per direction
DX = 2.0 / NC
XMIN = (-1.0 + 0.1234567, -1.0 + 0.2345678, -1.0 + 0.3456789)


def tp_flags(ptype, order):
    """(vc_cons, cc_cons, cc_eno, cc_lfb, vc_order, cc_order) of
    TestProlongate_Set / _Check (test.cxx:86-150)."""
    if ptype == "interpolate":
        return False, False, False, False, order, order
    if ptype == "conservative":
        return True, True, False, False, order, (0 if order == 0 else order + 1)
    if ptype in ("ddf", "hermite"):
        return False, True, False, False, order, order - 1
    if ptype == "eno":
        return False, True, True, False, order, order - 1
    if ptype == "natural":
        return False, True, False, False, order, order
    if ptype == "poly-cons3lfb":
        return False, True, False, True, order, min(3, order)
    raise ValueError(ptype)


def fun1d(x, dx, cons, order):
    """test.cxx:52-61: sum_{p<=order} x^p, or its average over [x-dx/2,
    x+dx/2] in a conservative direction."""
    res = 0.0
    for p in range(order + 1):
        if not cons:
            res += 1.0 if p == 0 else x ** p
        else:
            res += ((x + dx / 2) ** (p + 1) - (x - dx / 2) ** (p + 1)) / ((p + 1) * dx)
    return res


def axis_values(d, idx, h, cc, cons, order):
    return np.array([fun1d(XMIN[d] + (i + 0.5 * cc) * h, h, cons, order) for i in idx])


def run_case(case, gf, rlo, rhi, ghost_plane=None):
    ptype, order = case["prolongation_type"], case["prolongation_order"]
    per_group = case["prolongate_per_group"] == "yes"
    it = tuple(int(c) for c in gf[2:])
    vc_cons, cc_cons, cc_eno, cc_lfb, vc_order, cc_order = tp_flags(ptype, order)
    op = oracle.lookup(ptype, order, it)
    clo, chi = oracle.coarse_box(op, rlo, rhi)
    crse = oracle.Fab(clo, chi, 1)
    ca, fa = [], []
    for d in range(3):
        cons = cc_cons if it[d] else vc_cons
        o = cc_order if it[d] else vc_order
        ca.append(axis_values(d, range(clo[d], chi[d] + 1), DX, it[d], cons, o))
        fa.append(axis_values(d, range(rlo[d], rhi[d] + 1), DX / 2, it[d], cons, o))
    crse.arr[0] = ca[2][:, None, None] * ca[1][None, :, None] * ca[0][None, None, :]
    if ghost_plane is not None:
        crse.arr[0, ghost_plane - clo[2]] = 1e10
    fine = oracle.Fab(rlo, rhi, 1)
    oracle.prolongate(op, crse, fine, rlo, rhi, per_group=per_group)
    expect = fa[2][:, None, None] * fa[1][None, :, None] * fa[0][None, None, :]
    return fine.arr[0], expect, clo


def exact_rows():
    for name, case in sorted(GOLD.items()):
        for gf, errs in sorted(case["max_error"].items()):
            if errs[2] is not None and errs[2] <= EXACT:
                yield name, gf


def test_fixture_is_complete():
    assert len(GOLD) == 19
    assert all(len(c["max_error"]) == 8 for c in GOLD.values())
    # exact rows: everything except cc_nc_o3 / natural_o7 cell-centred rows
    # and the poly-cons3lfb rows next to the jump
    assert len(list(exact_rows())) == 152 - 7 - 4 - 1 - 5 - 5


@pytest.mark.parametrize("name,gf", list(exact_rows()))
def test_golden_exact_rows_are_exact_in_the_oracle(name, gf):
    """The reference reports max_error at rounding level for this case and
    centring; the oracle must reproduce TestProlongate's function exactly on
    the same grid (a region at x, y, z > 0, clear of the eno/lfb jump at 0
    and of the outer boundary)."""
    # fine indices: x in (0.35, 0.6) etc.; the widest stencil reaches 4
    # coarse cells (0.27) inward, so every coarse value sits at x > 0
    rlo = tuple(int(math.ceil((0.35 - XMIN[d]) / (DX / 2))) for d in range(3))
    rhi = tuple(r + 7 for r in rlo)
    got, expect, _ = run_case(GOLD[name], gf, rlo, rhi)
    err = np.max(np.abs(got - expect))
    assert err < 1e-12, (name, gf, err)


def test_natural_o7_cell_centred_rows_are_the_outer_weight_times_the_ghost():
    """2.13623046869e7 = (35/16384)(1e10 - f): a CC CONS order-7 z stencil
    whose outermost coarse point holds the Dirichlet ghost value 1e10."""
    case = GOLD["test_natural_o7"]
    outer = 35.0 / 16384.0
    assert oracle.coeff(oracle.CC, oracle.CONS, 7, 0, 0) == -outer
    for gf in ("gf001", "gf011", "gf101", "gf111"):
        g = case["max_error"][gf][2]
        assert abs(g - 1e10 * outer) / (1e10 * outer) < 1e-9, (gf, g)
    # the mechanism: put 1e10 on the lowest coarse z-plane of the stencil
    rlo = (30, 30, 30)
    rhi = (37, 37, 37)
    clo, _ = oracle.coarse_box(oracle.lookup("natural", 7, (0, 0, 1)), rlo, rhi)
    got, expect, _ = run_case(case, "gf001", rlo, rhi, ghost_plane=clo[2])
    err = np.max(np.abs(got - expect))
    g = case["max_error"]["gf001"][2]
    assert abs(err - 1e10 * outer) / (1e10 * outer) < 1e-6, err
    assert abs(err - g) / g < 1e-6, (err, g)
