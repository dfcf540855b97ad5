"""Pins of the oracle's boundary conditions (oracle/boundary.py) against
closed forms: each condition must reproduce the data it is exact for, so a
wrong source index, mirror point, sign, parity or quotient fails here.

Domain [0, 5]^3 in a FAB box [-3, 8]^3 (three boundary points per face)."""
import math

import numpy as np
import pytest

from oracle.boundary import (DIRICHLET, LINEAR_EXTRAPOLATION, NEUMANN, NONE,
                             REFLECTION, ROBIN, apply_bc)

LO, HI = -3, 8
DOM_LO, DOM_HI = (0, 0, 0), (5, 5, 5)
AX = np.arange(LO, HI + 1)


def grid():
    z, y, x = np.meshgrid(AX, AX, AX, indexing="ij")
    return x, y, z


def outside_dirs():
    x, y, z = grid()
    return [(a < DOM_LO[d]) | (a > DOM_HI[d]) for d, a in enumerate((x, y, z))]


def fab(fs, fill=np.nan):
    """ncomp arrays of closed-form values inside the domain, fill outside."""
    o = outside_dirs()
    out = o[0] | o[1] | o[2]
    arr = np.stack([np.where(out, fill, f) for f in fs]).astype(np.float64)
    return arr, out


def run(arr, faces, cell=(1, 1, 1), **kw):
    return apply_bc(arr, (LO, LO, LO), DOM_LO, DOM_HI, faces, cell, **kw)


def test_dirichlet_wins_and_takes_no_parity():
    rng = np.random.default_rng(1)
    f = rng.uniform(-1, 1, (2,) + (HI - LO + 1,) * 3)
    arr, out = fab(list(f))
    faces = [[DIRICHLET, NEUMANN, REFLECTION], [NEUMANN, DIRICHLET, REFLECTION]]
    run(arr, faces, dirichlet=[7.5, -2.25], parities=[(-1, -1, -1)] * 2)
    x, y, z = grid()
    o = outside_dirs()
    # every point outside through a Dirichlet face, corners included
    dmask = (x < 0) | (y > 5)
    for c, v in enumerate((7.5, -2.25)):
        assert np.all(arr[c][dmask] == v)
    # Neumann alone in y (lower face): the edge value
    m = o[1] & ~o[0] & ~o[2] & (y < 0)
    src = arr[:, :, 3, :]  # y = 0 plane
    for c in range(2):
        assert np.array_equal(arr[c][m], np.broadcast_to(src[c][:, None, :], arr[c].shape)[m])


def test_neumann_is_the_clamped_point():
    rng = np.random.default_rng(2)
    f = rng.uniform(-1, 1, (HI - LO + 1,) * 3)
    arr, out = fab([f])
    run(arr, [[NEUMANN] * 3, [NEUMANN] * 3])
    x, y, z = grid()
    cx, cy, cz = (np.clip(a, 0, 5) - LO for a in (x, y, z))
    assert np.array_equal(arr[0], f[cz, cy, cx])


def lin(x, y, z):
    return 3.0 + 2.0 * x - 1.0 * y + 5.0 * z


def test_linear_extrapolation_is_exact_for_linear_data():
    x, y, z = grid()
    arr, out = fab([lin(x, y, z)])
    run(arr, [[LINEAR_EXTRAPOLATION] * 3, [LINEAR_EXTRAPOLATION] * 3])
    o = outside_dirs()
    nout = o[0].astype(int) + o[1] + o[2]
    want = lin(x, y, z)
    # faces, and edges and corners on the diagonal (equal depth in every
    # outside direction): sqrt(S / T) is the depth, the value is linear
    depth = [np.where(o[d], np.maximum(DOM_LO[d] - a, a - DOM_HI[d]), 0)
             for d, a in enumerate((x, y, z))]
    dmax = np.maximum(np.maximum(depth[0], depth[1]), depth[2])
    diag = np.ones_like(out)
    for d in range(3):
        diag &= ~o[d] | (depth[d] == dmax)
    m = out & diag
    assert np.array_equal(arr[0][m], want[m])
    assert m.sum() > 200 and (nout[m] == 3).any()
    # off the diagonal the reference divides the int sums: at (-1, -2, 2),
    # src = (0, 0, 2), delta = (1, 1, 0), S = 5, T = 2, S / T = 2, so
    # val = f(0,0,2) + sqrt(2) (f(0,0,2) - f(1,1,2)) = 13 - sqrt(2)
    v = arr[0, 2 - LO, -2 - LO, -1 - LO]
    assert v == 13.0 - math.sqrt(2.0)
    assert v != 13.0 - math.sqrt(2.5)


@pytest.mark.parametrize("cell", [(1, 1, 1), (0, 0, 0), (1, 0, 1)])
def test_reflection_reproduces_data_of_the_given_parity(cell):
    """Lower faces reflect about -1/2 (cell) or 0 (vertex), upper faces about
    5 + 1/2 or 5; data odd/even about the mirror in each direction, so the
    reflected value is the closed form itself, corners included."""
    x, y, z = grid()
    for face in (0, 1):
        m = [(-0.5 * cell[d] if face == 0 else 5 + 0.5 * cell[d]) for d in range(3)]
        odd = [x - m[0], y - m[1], z - m[2]]
        f0 = odd[0] * odd[1] ** 2 * odd[2]          # parities (-1, +1, -1)
        f1 = odd[0] ** 2 * odd[1] * odd[2] ** 2     # parities (+1, -1, +1)
        arr, out = fab([f0, f1])
        faces = [[DIRICHLET] * 3, [DIRICHLET] * 3]
        faces[face] = [REFLECTION] * 3
        run(arr, faces, cell=cell, dirichlet=[0.0, 0.0],
            parities=[(-1, 1, -1), (1, -1, 1)])
        o = outside_dirs()
        side = [(a < 0) if face == 0 else (a > 5) for a in (x, y, z)]
        m_ = out.copy()
        for d in range(3):
            m_ &= ~o[d] | side[d]  # outside only through the reflecting faces
        assert m_.sum() > 300
        assert np.array_equal(arr[0][m_], f0[m_]), (cell, face)
        assert np.array_equal(arr[1][m_], f1[m_]), (cell, face)


def test_robin_reproduces_one_over_r_on_faces():
    """f = finf + C / r is what the Robin condition extends exactly on faces;
    on an edge with two Robin directions the factor ri / rb applies twice,
    finf + C ri / rb^2."""
    finf, C = 2.0, 0.75
    xmin, dx = (1.0, 0.6, 0.8), (0.1, 0.1, 0.1)
    x, y, z = grid()
    X = [xmin[0] + x * dx[0], xmin[1] + y * dx[1], xmin[2] + z * dx[2]]
    r = np.sqrt(X[0] ** 2 + X[1] ** 2 + X[2] ** 2)
    arr, out = fab([finf + C / r])
    run(arr, [[ROBIN] * 3, [ROBIN] * 3], robin=[finf], xmin=xmin, dx=dx)
    o = outside_dirs()
    nout = o[0].astype(int) + o[1] + o[2]
    face = out & (nout == 1)
    np.testing.assert_allclose(arr[0][face], (finf + C / r)[face], rtol=1e-14)
    # edge points: src clamps both outside coordinates
    cs = [np.clip(a, 0, 5) for a in (x, y, z)]
    ri = np.sqrt(sum((xmin[d] + cs[d] * dx[d]) ** 2 for d in range(3)))
    edge = out & (nout == 2)
    np.testing.assert_allclose(arr[0][edge], (finf + C * ri / r ** 2)[edge], rtol=1e-14)
    assert not np.allclose(arr[0][edge], (finf + C / r)[edge], rtol=1e-6)


def test_faces_without_condition_are_left_alone():
    rng = np.random.default_rng(4)
    f = rng.uniform(-1, 1, (HI - LO + 1,) * 3)
    arr, out = fab([f], fill=9.0)
    run(arr, [[NONE, NEUMANN, NONE], [NONE, NONE, NONE]])
    x, y, z = grid()
    o = outside_dirs()
    untouched = out & ~(o[1] & (y < 0))
    assert np.all(arr[0][untouched] == 9.0)
    # y-lower Neumann with x outside through a face without a condition
    # reads the (unset) point beside it: the reference treats x as interior
    m = (y < 0) & o[0] & ~o[2]
    assert np.all(arr[0][m] == 9.0)
    m = (y < 0) & ~o[0] & ~o[2]
    assert np.array_equal(arr[0][m], f[(z - LO)[m], 3, (x - LO)[m]])


def test_periodic_images_are_the_points_one_period_away():
    """periodic_fill: data (i mod p) + 100 (k mod p) with p the period (the
    domain's extent, one less for vertex centring, whose first and last
    points coincide) is reproduced at every point outside the domain only
    in the periodic directions; points also outside in another direction,
    and the domain itself, are left alone; periodic faces are 'none' for
    the boundary conditions."""
    from oracle.boundary import PERIODIC, periodic_fill
    x, y, z = grid()
    for cell in (1, 0):
        p = 6 - (1 - cell)  # domain [0, 5]: 6 cells, or 5 vertex intervals
        f = (x % p) + 100.0 * (z % p)
        arr, out = fab([f], fill=np.nan)
        periodic_fill(arr, (LO, LO, LO), DOM_LO, DOM_HI, (0, 2), (cell, 0, cell))
        o = outside_dirs()
        m = (o[0] | o[2]) & ~o[1]
        assert np.array_equal(arr[0][m], f[m]), cell
        assert np.all(np.isnan(arr[0][o[1]]))
        assert np.array_equal(arr[0][~out], f[~out])
        # the y faces' condition reads the images (a periodic x face is none)
        run(arr, [[PERIODIC, NEUMANN, PERIODIC], [PERIODIC, NEUMANN, PERIODIC]], cell=(cell, 0, cell))
        yc = np.clip(y, 0, 5)
        assert np.array_equal(arr[0], np.take_along_axis(arr[0], (yc - LO), axis=1))
