#!/usr/bin/env python3
"""Exact-rational derivation of the rf=2 prolongation stencil weights.

Derives, from first principles, every 1-D weight table the kernels use and
writes them to ``cxp_coeffs.inc`` as exact dyadic rationals.

The construction follows the one described in
CarpetX/stencils/stencils.jl:36-277 (Lagrange solve for polynomial stencils,
antiderivative trick for conservative ones); the code is our own, written
with Python's ``fractions`` instead of Julia's BigInt rationals.

Conventions (coarse index 0 is the point/cell that owns the fine point):

* vertex-centred (VC): coarse vertex i sits at x = i; the fine point with
  offset 1 sits at x = 1/2 (offset 0 coincides with vertex 0 and is a copy);
* cell-centred (CC): coarse cell i spans [i - 1/2, i + 1/2]; the fine
  sub-cell with offset 0 spans [-1/2, 0], offset 1 spans [0, 1/2].

Table layouts match how the 1-D operators index them
(CarpetX/src/prolongate_3d_rf2_impl.hxx:396-842):

* ``VC POLY p``  (p odd): weights for coarse offsets -(p-1)/2 .. (p+1)/2 at
  fine offset 1.
* ``CC POLY p``: weights for coarse offsets -(p+1)//2 - ... i.e. the
  "lean left" stencil [-(p//2) - 1 + (p%2 == 0), ...]; precisely offsets
  -((p+1)//2) .. p//2 evaluated at the centre of the left sub-cell x=-1/4.
* ``CC CONS p``: cell averages over offsets -((p+1)//2) .. p//2, giving the
  average over the left sub-cell [-1/2, 0].
* ``CC ENO 2``: row m is the conservative 3-cell stencil over offsets
  [m-2, m] for the left sub-cell.
"""
from fractions import Fraction as Fr
import os
import sys


def lagrange_weights(nodes, x):
    """Weights w_j with sum_j w_j f(nodes[j]) == p(x) for the interpolant."""
    w = []
    for j, xj in enumerate(nodes):
        num = Fr(1)
        den = Fr(1)
        for m, xm in enumerate(nodes):
            if m != j:
                num *= (x - xm)
                den *= (xj - xm)
        w.append(num / den)
    return w


def vc_poly(p):
    """VC polynomial stencil of odd order p at x = 1/2."""
    assert p % 2 == 1
    nodes = [Fr(i) for i in range(-(p - 1) // 2, (p + 1) // 2 + 1)]
    return lagrange_weights(nodes, Fr(1, 2))


def cc_poly(p):
    """CC polynomial stencil of order p at the left sub-cell centre x=-1/4.

    Cell centres sit at integers; offsets -((p+1)//2) .. p//2.
    """
    lo = -((p + 1) // 2)
    nodes = [Fr(i) for i in range(lo, lo + p + 1)]
    return lagrange_weights(nodes, Fr(-1, 4))


def cc_cons_range(lo, hi, xl, xr):
    """Conservative weights: average over [xl, xr] from cell averages lo..hi.

    The primitive F(e) = sum of averages of cells left of edge e is known
    exactly at the edges lo-1/2 .. hi+1/2; interpolate F with the degree
    (hi-lo+1) polynomial through those edges and difference it.
    """
    edges = [Fr(2 * k - 1, 2) for k in range(lo, hi + 2)]
    wl = lagrange_weights(edges, xl)
    wr = lagrange_weights(edges, xr)
    ncell = hi - lo + 1
    w = [Fr(0)] * ncell
    for k in range(len(edges)):
        d = (wr[k] - wl[k]) / (xr - xl)
        # F at edge k is the sum of the averages of cells lo .. lo+k-1
        for c in range(k):
            w[c] += d
    return w


def cc_cons(p):
    """CC conservative stencil of order p for the left sub-cell [-1/2, 0]."""
    lo = -((p + 1) // 2)
    hi = p // 2
    return cc_cons_range(lo, hi, Fr(-1, 2), Fr(0))


def cc_eno2_rows():
    """Three shifted 3-cell conservative stencils for the left sub-cell."""
    return [cc_cons_range(m - 2, m, Fr(-1, 2), Fr(0)) for m in range(3)]


def cc_eno4_rows():
    """Five shifted 5-cell conservative stencils (unused by any family)."""
    return [cc_cons_range(m - 4, m, Fr(-1, 2), Fr(0)) for m in range(5)]


TABLES = {
    "VC_POLY": {p: vc_poly(p) for p in (1, 3, 5, 7)},
    "CC_POLY": {p: cc_poly(p) for p in range(0, 8)},
    "CC_CONS": {p: cc_cons(p) for p in range(0, 8)},
    "CC_ENO2": {0: [w for row in cc_eno2_rows() for w in row]},
}


def _lit(fr):
    return "(%d.0 / %d.0)" % (fr.numerator, fr.denominator)


def emit(path):
    lines = [
        "// Generated by gen_coeffs.py: exact rf=2 stencil weights (do not edit).",
        "#pragma once",
    ]
    for name, table in TABLES.items():
        for p, ws in table.items():
            arr = ", ".join(_lit(w) for w in ws)
            lines.append(
                "static constexpr double k%s_%d[%d] = {%s};" % (name, p, len(ws), arr))
    text = "\n".join(lines) + "\n"
    with open(path, "w") as f:
        f.write(text)


if __name__ == "__main__":
    out = sys.argv[1] if len(sys.argv) > 1 else os.path.join(
        os.path.dirname(os.path.abspath(__file__)), "cxp_coeffs.inc")
    emit(out)
    for name, table in TABLES.items():
        for p, ws in table.items():
            print(name, p, [str(w) for w in ws])
