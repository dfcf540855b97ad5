"""Pins the oracle's weight tables and 1-D operators (CPU only).

* tables: three-way agreement between the transcription of
  CarpetX/src/prolongate_3d_rf2_impl.hxx:56-384 (golden fixture), the exact
  rational derivation (carpetx_amd/csrc/gen_coeffs.py) and the oracle's
  compiled tables;
* 1-D operators: the known-answer tests of test_interp1d
  (prolongate_3d_rf2_impl.hxx:882-1174), asserted with exact equality:
  reproduction of x^q for q <= order, exact conservation and exact integral.
  ENO rows are checked for every shift (the reference's own test loops over
  shifts but always evaluates shift 0).
"""
import json
import os
from fractions import Fraction as Fr

import numpy as np
import pytest

import oracle
from oracle import CC, CONS, ENO, MINMOD, POLY, VC
from carpetx_amd.csrc import gen_coeffs

GOLD = os.path.join(os.path.dirname(__file__), "golden", "coeffs_reference.json")


@pytest.fixture(scope="module")
def ref():
    with open(GOLD) as f:
        return json.load(f)


def _fr(s):
    return Fr(s)


@pytest.mark.parametrize("name,cent,intp,orders", [
    ("VC_POLY", VC, POLY, (1, 3, 5, 7)),
    ("CC_POLY", CC, POLY, range(8)),
    ("CC_CONS", CC, CONS, range(8)),
])
def test_tables_three_way(ref, name, cent, intp, orders):
    for p in orders:
        gold = [_fr(s) for s in ref[name][str(p)]]
        derived = gen_coeffs.TABLES[name][p]
        assert derived == gold, (name, p)
        compiled = [oracle.coeff(cent, intp, p, 0, i) for i in range(p + 1)]
        assert compiled == [float(g) for g in gold], (name, p)


def test_eno_tables(ref):
    gold2 = [[_fr(s) for s in row] for row in ref["CC_ENO2"]]
    assert gen_coeffs.cc_eno2_rows() == gold2
    for m in range(3):
        assert [oracle.coeff(CC, ENO, 2, m, i) for i in range(3)] == \
            [float(x) for x in gold2[m]]
    gold4 = [[_fr(s) for s in row] for row in ref["CC_ENO4"]]
    assert gen_coeffs.cc_eno4_rows() == gold4


def test_derivation_is_exact_polynomial_reproduction():
    # independent of the tables: each derived VC/CC POLY row reproduces x^q
    for p in (1, 3, 5, 7):
        w = gen_coeffs.vc_poly(p)
        xs = range(-(p - 1) // 2, (p + 1) // 2 + 1)
        for q in range(p + 1):
            assert sum(wi * Fr(x) ** q for wi, x in zip(w, xs)) == Fr(1, 2) ** q
    for p in range(8):
        w = gen_coeffs.cc_cons(p)
        lo = -((p + 1) // 2)
        for q in range(p + 1):
            # cell averages of (q+1) x^q over [i-1/2, i+1/2]
            avg = [Fr(2 * i + 1, 2) ** (q + 1) - Fr(2 * i - 1, 2) ** (q + 1)
                   for i in range(lo, lo + p + 1)]
            target = (Fr(0) ** (q + 1) - Fr(-1, 2) ** (q + 1)) * 2
            assert sum(wi * a for wi, a in zip(w, avg)) == target


# --- known-answer tests (test_interp1d, prolongate_3d_rf2_impl.hxx:882-1174)

def _poly_kat(cent, p):
    g = oracle.required_ghosts(cent, POLY, p)
    n = 1 + 2 * (g + 1)
    i0 = n // 2
    for q in range(p + 1):
        for off in (0, 1):
            rmin, rmax = oracle.stencil_radius(cent, POLY, p, 0, off)
            assert -g <= rmin <= 0 <= rmax <= g
            ys = np.full(n, np.nan)
            for i in range(-(g + 1), g + 2):
                if rmin <= i <= rmax:
                    ys[i0 + i] = (i + cent / 2.0) ** q
            x = cent / 4.0 + off / 2.0
            y1 = oracle.interp1d(cent, POLY, p, 0, off, ys, i0)
            assert np.isfinite(y1)
            assert y1 == x ** q, (cent, p, q, off)


@pytest.mark.parametrize("p", [1, 3, 5, 7])
def test_kat_vc_poly(p):
    _poly_kat(VC, p)


@pytest.mark.parametrize("p", list(range(8)))
def test_kat_cc_poly(p):
    _poly_kat(CC, p)


def _cons_kat(intp, p, shift):
    g = oracle.required_ghosts(CC, intp, p)
    n = 1 + 2 * (g + 2)
    i0 = n // 2
    for q in range(p + 1):
        def fint(x):
            return x ** (q + 1)
        y1 = []
        for off in (0, 1):
            rmin, rmax = oracle.stencil_radius(CC, intp, p, shift, off)
            assert -g <= rmin and rmax <= g
            ys = np.full(n, np.nan)
            for i in range(-(g + 1), g + 2):
                if rmin <= i <= rmax:
                    x = i + 0.5
                    ys[i0 + i] = fint(x + 0.5) - fint(x - 0.5)
            v = oracle.interp1d(CC, intp, p, shift, off, ys, i0)
            assert np.isfinite(v)
            y1.append(v)
        avg0 = fint(1.0) - fint(0.0)
        assert y1[0] / 2 + y1[1] / 2 == avg0, (intp, p, q, shift)
        assert y1[0] * 0.5 + y1[1] * 0.5 == fint(1.0) - fint(0.0)


@pytest.mark.parametrize("p", list(range(8)))
def test_kat_cc_cons(p):
    _cons_kat(CONS, p, 0)


@pytest.mark.parametrize("shift", [-1, 0, 1])
def test_kat_cc_eno2_every_shift(shift):
    _cons_kat(ENO, 2, shift)


def test_kat_cc_minmod():
    _cons_kat(MINMOD, 1, 0)


def test_minmod_closed_form():
    # minmod slope / 4 (prolongate_3d_rf2_impl.hxx:857-877); textbook cases
    f = oracle.interp1d
    assert f(CC, MINMOD, 1, 0, 0, [1.0, 2.0, 4.0], 1) == 2.0 - 0.25
    assert f(CC, MINMOD, 1, 0, 1, [1.0, 2.0, 4.0], 1) == 2.0 + 0.25
    # extremum: slopes of opposite sign -> no slope
    assert f(CC, MINMOD, 1, 0, 0, [1.0, 2.0, 1.5], 1) == 2.0
    assert f(CC, MINMOD, 1, 0, 1, [3.0, 2.0, 2.0], 1) == 2.0


def test_vc_cons_order0_odd_points_are_zero():
    # the deprecated VC CONS order 0 has an empty coeffs1 (:225, :619-628)
    assert oracle.interp1d(VC, CONS, 0, 0, 1, [5.0, 7.0, 9.0], 1) == 0.0
    assert oracle.interp1d(VC, CONS, 0, 0, 0, [5.0, 7.0, 9.0], 1) == 7.0
    assert oracle.interp1d(VC, CONS, 1, 0, 1, [5.0, 7.0, 9.0], 1) == 8.0


def test_undivided_difference_tie_prefers_centre():
    """ENO picks argmin of penalised max |second difference| (strict <).

    1-D data with |D(0)| = |D(-1)| (1 + 2^-30) = |D(1)| (1 + 2^-30): the
    centred stencil wins only because of the 1 + 2^-26 |s| penalty
    (prolongate_3d_rf2_impl.hxx:1505-1513).  With |D(0)| larger by 2^-24 the
    penalty no longer covers it and the left shift wins, being the first of
    the two tied candidates under strict <."""
    def run(scale):
        # second differences: D(-1) = a, D(0) = -b, D(1) = a
        # construct c(-2..2) with c(-2)=0, c(-1)=0
        a = 1.0
        b = a * scale
        c = [0.0, 0.0]
        c.append(a - c[0] + 2 * c[1])          # c(0): D(-1) = c(-2)-2c(-1)+c(0)
        c.append(-b - c[1] + 2 * c[2])         # c(1): D(0)
        c.append(a - c[2] + 2 * c[3])          # c(2): D(1)
        data = np.broadcast_to(np.array(c), (1, 5, 5, 5)).copy()
        fab = oracle.Fab((-2, -2, -2), (2, 2, 2), 1, data)
        op = oracle.lookup("eno", 3, (1, 0, 0))
        # fine x = 0 (ic = 0, off = 0); VC in y/z so only x varies
        fine = oracle.Fab((0, 0, 0), (0, 0, 0), 1)
        oracle.prolongate(op, fab, fine, (0, 0, 0), (0, 0, 0))
        return fine.arr[0, 0, 0, 0], c
    v, c = run(1.0 + 2.0 ** -30)
    rows = gen_coeffs.cc_eno2_rows()
    centred = sum(float(w) * c[1 + i] for i, w in enumerate(rows[1]))
    assert v == centred
    v, c = run(1.0 + 2.0 ** -24)
    left = sum(float(w) * c[0 + i] for i, w in enumerate(rows[0]))
    assert v == left
