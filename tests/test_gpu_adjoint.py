"""GPU adjoints through the C ABI.

Checks (no reference counterpart exists; these pin the transpose from the
forward operator):
* dot-product identity <P u, v> == <u, P^T v> with the GPU forward, every
  linear operator family and centring;
* an explicit dense P^T on a small sub-box: columns of P are the oracle's
  forward applied to coarse unit vectors;
* ENO and linear-fallback operators, per-var and per-group: the dense
  Jacobian of the oracle at u, taken by central differences (exact inside
  one branch: every fine point's value is linear in the coarse data for a
  fixed ENO shift and fallback decision), transposed; and the dot-product
  identity with the GPU forward's directional derivative;
* MINMOD: the oracle's Jacobian by exact central differences, the
  host build of the same record and gather routines bit for bit, the
  dot-product identity with P(u) = J(u) u, and the recorded path.
"""
import numpy as np
import pytest
import torch

import oracle
from carpetx_amd.interp import Box, FArrayBox, get_interpolator

pytestmark = pytest.mark.gpu


def _fabs(op, rlo, rhi, ncomp, rng):
    P = get_interpolator(*op)
    cb = P.CoarseBox(Box(rlo, rhi))
    u = rng.uniform(-1, 1, size=(ncomp,) + tuple(reversed(cb.size)))
    fbox = Box(rlo, rhi)
    v = rng.uniform(-1, 1, size=(ncomp,) + tuple(reversed(fbox.size)))
    return P, cb, fbox, u, v


LINEAR = [("interpolate", 1), ("interpolate", 3), ("interpolate", 5), ("interpolate", 7),
          ("natural", 1), ("natural", 3), ("natural", 5), ("natural", 7),
          ("ddf", 3), ("ddf", 5), ("conservative", 0), ("conservative", 1), ("hermite", 5)]


@pytest.mark.parametrize("fam,order", LINEAR)
def test_dot_product_identity(fam, order):
    rng = np.random.default_rng(order)
    for c in range(8):
        it = ((c >> 2) & 1, (c >> 1) & 1, c & 1)
        P, cb, fbox, u, v = _fabs((fam, order, it), (-3, 2, -1), (40, 21, 17), 2, rng)
        crse = FArrayBox(cb, 2, tensor=torch.from_numpy(u).cuda())
        fine = FArrayBox(fbox, 2)
        P.interp(crse, 0, fine, 0, 2, fbox)
        vb = FArrayBox(fbox, 2, tensor=torch.from_numpy(v).cuda())
        cbar = FArrayBox(cb, 2)
        P.interp_adjoint(cbar, 0, vb, 0, 2, fbox)
        torch.cuda.synchronize()
        lhs = float((fine.t * vb.t).sum())
        rhs = float((crse.t * cbar.t).sum())
        scale = float((fine.t.abs() * vb.t.abs()).sum())
        assert abs(lhs - rhs) <= 1e-13 * scale, (fam, order, it, lhs, rhs)


def _dense_P(op_o, cb, rlo, rhi):
    """Columns of P from the oracle applied to coarse unit vectors."""
    ncoarse = int(np.prod(cb.size))
    fshape = tuple(rhi[d] - rlo[d] + 1 for d in (2, 1, 0))
    cols = []
    for n in range(ncoarse):
        e = np.zeros(ncoarse)
        e[n] = 1.0
        ch = oracle.Fab(cb.lo, cb.hi, 1, e.reshape((1,) + tuple(reversed(cb.size))))
        fh = oracle.Fab(rlo, rhi, 1)
        oracle.prolongate(op_o, ch, fh, rlo, rhi)
        cols.append(fh.arr.reshape(-1))
    return np.stack(cols, axis=1), fshape


@pytest.mark.parametrize("fam,order,it", [("natural", 5, (0, 0, 0)), ("natural", 3, (1, 0, 1)),
                                          ("interpolate", 7, (0, 1, 0)), ("ddf", 5, (1, 1, 1))])
def test_dense_transpose_small_box(fam, order, it):
    rng = np.random.default_rng(3)
    rlo, rhi = (1, 0, -1), (8, 7, 6)   # an 8^3 fine sub-box
    op_o = oracle.lookup(fam, order, it)
    P, cb, fbox, u, v = _fabs((fam, order, it), rlo, rhi, 1, rng)
    A, fshape = _dense_P(op_o, cb, rlo, rhi)
    expect = (A.T @ v.reshape(-1)).reshape((1,) + tuple(reversed(cb.size)))
    vb = FArrayBox(fbox, 1, tensor=torch.from_numpy(v).cuda())
    cbar = FArrayBox(cb, 1, tensor=torch.full((1,) + tuple(reversed(cb.size)), 0.25,
                                               dtype=torch.float64, device="cuda"))
    P.interp_adjoint(cbar, 0, vb, 0, 1, fbox)
    torch.cuda.synchronize()
    got = cbar.t.cpu().numpy() - 0.25   # the adjoint accumulates
    assert np.max(np.abs(got - expect)) <= 1e-13 * max(1.0, np.max(np.abs(expect)))


NONLINEAR = [("eno", 3, (1, 1, 1), False), ("eno", 3, (1, 0, 1), True),
             ("poly-eno3lfb", 3, (0, 1, 1), False), ("poly-eno3lfb", 3, (1, 1, 0), True),
             ("poly-cons3lfb", 3, (1, 1, 1), False), ("poly-cons3lfb", 5, (1, 0, 0), True),
             ("poly-cons3lfb", 1, (0, 1, 1), False)]


@pytest.mark.parametrize("fam,order,it,pg", NONLINEAR)
def test_nonlinear_adjoint_against_dense_jacobian(fam, order, it, pg):
    """Jacobian of the oracle at u by central differences (h = 1e-7),
    transposed, against the device adjoint linearised at u."""
    rng = np.random.default_rng(order * 10 + it[0] + 2 * it[1] + 4 * it[2])
    rlo, rhi = (0, 1, -2), (7, 6, 4)
    nc = 2
    op_o = oracle.lookup(fam, order, it)
    P, cb, fbox, u, v = _fabs((fam, order, it), rlo, rhi, nc, rng)
    csz = int(np.prod(cb.size))
    h = 1e-7

    def fwd(x):
        ch = oracle.Fab(cb.lo, cb.hi, nc, x.reshape((nc,) + tuple(reversed(cb.size))))
        fh = oracle.Fab(rlo, rhi, nc)
        oracle.prolongate(op_o, ch, fh, rlo, rhi, ncomp=nc, per_group=pg)
        return fh.arr.reshape(-1)
    u0 = u.reshape(-1)
    J = np.zeros((nc * int(np.prod(fbox.size)), nc * csz))
    for n in range(nc * csz):
        e = np.zeros(nc * csz)
        e[n] = h
        J[:, n] = (fwd(u0 + e) - fwd(u0 - e)) / (2 * h)
    expect = (J.T @ v.reshape(-1)).reshape((nc,) + tuple(reversed(cb.size)))
    crse = FArrayBox(cb, nc, tensor=torch.from_numpy(u).cuda())
    vb = FArrayBox(fbox, nc, tensor=torch.from_numpy(v).cuda())
    cbar = FArrayBox(cb, nc)
    P.interp_adjoint(cbar, 0, vb, 0, nc, fbox, crse_in=crse, per_group=pg)
    torch.cuda.synchronize()
    got = cbar.t.cpu().numpy()
    assert np.max(np.abs(got - expect)) <= 1e-6 * max(1.0, np.max(np.abs(expect)))


@pytest.mark.parametrize("fam,order,it,pg,nc", [x + (2,) for x in NONLINEAR[:4]] +
                         [("poly-cons3lfb", 5, (1, 0, 1), True, 12)])
def test_nonlinear_adjoint_dot_product(fam, order, it, pg, nc):
    """<J du, v> == <du, J^T v> on a larger region, J du from the GPU
    forward by central differences (12 per-group components: the branch
    record no longer fits in LDS and runs one thread per point)."""
    rng = np.random.default_rng(21)
    rlo, rhi = (-4, 3, 0), (37, 22, 13)
    P, cb, fbox, u, v = _fabs((fam, order, it), rlo, rhi, nc, rng)
    du = rng.uniform(-1, 1, size=u.shape)
    eps = 1e-7

    def fwd(x):
        crse = FArrayBox(cb, nc, tensor=torch.from_numpy(x).cuda())
        fine = FArrayBox(fbox, nc)
        P.interp(crse, 0, fine, 0, nc, fbox, per_group=pg)
        return fine.t
    jdu = (fwd(u + eps * du) - fwd(u - eps * du)) / (2 * eps)
    vb = FArrayBox(fbox, nc, tensor=torch.from_numpy(v).cuda())
    cbar = FArrayBox(cb, nc)
    crse = FArrayBox(cb, nc, tensor=torch.from_numpy(u).cuda())
    P.interp_adjoint(cbar, 0, vb, 0, nc, fbox, crse_in=crse, per_group=pg)
    torch.cuda.synchronize()
    lhs = float((jdu * vb.t).sum())
    rhs = float((torch.from_numpy(du).cuda() * cbar.t).sum())
    scale = float((jdu.abs() * vb.t.abs()).sum())
    assert abs(lhs - rhs) <= 1e-6 * scale, (lhs, rhs)


@pytest.mark.parametrize("fam,order,it,pg,nc", [x + (2,) for x in NONLINEAR] +
                         [("eno", 3, (1, 1, 1), True, 12), ("natural", 5, (0, 0, 0), False, 2)])
def test_branch_record_matches_recomputed_branches(fam, order, it, pg, nc):
    """cxp_prolongate_record writes the same fine values as cxp_prolongate,
    and the adjoint from its branch record equals the adjoint that recomputes
    the branches from crse_in, bit for bit (per-var and per-group; 12
    per-group components run the one-thread-per-point kernels).  A linear
    operator (poly-cons3lfb of order 1 included: its fallback is inactive)
    records nothing and its recorded adjoint is its adjoint."""
    rng = np.random.default_rng(31)
    rlo, rhi = (-4, 3, 0), (29, 22, 13)
    P, cb, fbox, u, v = _fabs((fam, order, it), rlo, rhi, nc, rng)
    crse = FArrayBox(cb, nc, tensor=torch.from_numpy(u).cuda())
    f1, f2 = FArrayBox(fbox, nc), FArrayBox(fbox, nc)
    P.interp(crse, 0, f1, 0, nc, fbox, per_group=pg)
    nb = P.branch_bytes(fbox, nc, per_group=pg)
    # ENO shifts, or a fallback that is active (a CONS direction of order > 1)
    branches = fam in ("eno", "poly-eno3lfb") or (fam == "poly-cons3lfb" and order > 1)
    assert nb == (fbox.numPts * (1 if pg else nc) if branches else 0)
    br = torch.empty(nb, dtype=torch.uint8, device="cuda") if nb else None
    P.interp_record(crse, 0, f2, 0, nc, fbox, br, per_group=pg)
    torch.cuda.synchronize()
    assert torch.equal(f1.t, f2.t)
    vb = FArrayBox(fbox, nc, tensor=torch.from_numpy(v).cuda())
    a1, a2 = FArrayBox(cb, nc), FArrayBox(cb, nc)
    P.interp_adjoint(a1, 0, vb, 0, nc, fbox, crse_in=crse, per_group=pg)
    P.interp_adjoint_recorded(br, a2, 0, vb, 0, nc, fbox, per_group=pg)
    torch.cuda.synchronize()
    assert torch.equal(a1.t, a2.t)


def _dyadic(rng, shape):
    return rng.integers(-(1 << 24), 1 << 24, size=shape).astype(np.float64) / float(1 << 24)


@pytest.mark.parametrize("it", [(1, 1, 1), (1, 0, 1), (0, 1, 0)])
def test_minmod_adjoint_against_dense_jacobian(it):
    """The oracle's Jacobian at u by central differences, transposed.  The
    data are dyadic (k / 2^24) and h = 2^-28, so the differences are exact
    inside one limiter branch and no branch flips within +-h."""
    from test_minmod_adjoint_host import oracle_jt
    rng = np.random.default_rng(4)
    rlo, rhi = (0, 1, -2), (7, 6, 4)
    P = get_interpolator("minmod", 3, it)
    cb = P.CoarseBox(Box(rlo, rhi))
    u = _dyadic(rng, tuple(reversed(cb.size)))
    v = _dyadic(rng, tuple(reversed(Box(rlo, rhi).size)))
    expect = oracle_jt(3, it, u, cb.lo, cb.hi, v, rlo, rhi)
    crse = FArrayBox(cb, 1, tensor=torch.from_numpy(u[None]).cuda())
    vb = FArrayBox(Box(rlo, rhi), 1, tensor=torch.from_numpy(v[None]).cuda())
    cbar = FArrayBox(cb, 1)
    P.interp_adjoint(cbar, 0, vb, 0, 1, Box(rlo, rhi), crse_in=crse)
    torch.cuda.synchronize()
    got = cbar.t.cpu().numpy()[0]
    assert np.max(np.abs(got - expect)) <= 1e-12 * max(1.0, np.max(np.abs(expect)))


@pytest.mark.parametrize("order", [1, 3, 5])
def test_minmod_adjoint_equals_its_host_build(order):
    """Every centring with a MINMOD direction: the device kernels give the
    bits of the same record and gather routines built for the host with
    bounds-checked accesses (carpetx_amd/host/minmod_adjoint_host.cpp),
    which the CPU tests pin to the oracle's Jacobian; CCC against the tiled
    gather's host build (several tiles and z-chunks), the others against
    the per-point gather.  VVV has no MINMOD direction and runs the linear
    adjoint, which sums in another order."""
    from test_minmod_adjoint_host import host_adjoint, host_adjoint_tiled
    rlo, rhi = (-5, 2, -3), (40, 37, 12)
    for c in range(8):
        it = ((c >> 2) & 1, (c >> 1) & 1, c & 1)
        rng = np.random.default_rng(50 + 8 * order + c)
        P = get_interpolator("minmod", order, it)
        cb = P.CoarseBox(Box(rlo, rhi))
        u = rng.uniform(-1, 1, tuple(reversed(cb.size)))
        v = rng.uniform(-1, 1, tuple(reversed(Box(rlo, rhi).size)))
        if all(it):
            bad, want = host_adjoint_tiled(order, it, u, cb.lo, cb.hi, v, rlo, rhi, 5, 3)
        else:
            bad, want, _ = host_adjoint(order, it, u, cb.lo, cb.hi, v, rlo, rhi)
        assert bad == 0
        crse = FArrayBox(cb, 1, tensor=torch.from_numpy(u[None]).cuda())
        vb = FArrayBox(Box(rlo, rhi), 1, tensor=torch.from_numpy(v[None]).cuda())
        cbar = FArrayBox(cb, 1)
        P.interp_adjoint(cbar, 0, vb, 0, 1, Box(rlo, rhi), crse_in=crse)
        torch.cuda.synchronize()
        got = cbar.t.cpu().numpy()[0]
        if any(it):
            assert np.array_equal(got, want), (order, it)
        else:
            assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want)), (order, it)


def test_minmod_dot_product_with_linearisation():
    """<J u, v> == <u, J^T v> where J u = P(u) for a positively homogeneous
    limiter (minmod(a u) = a minmod(u) for a > 0, so P(u) = J(u) u)."""
    rng = np.random.default_rng(8)
    rlo, rhi = (-4, 3, 0), (37, 22, 19)
    P, cb, fbox, u, v = _fabs(("minmod", 5, (1, 1, 1)), rlo, rhi, 2, rng)
    crse = FArrayBox(cb, 2, tensor=torch.from_numpy(u).cuda())
    fine = FArrayBox(fbox, 2)
    P.interp(crse, 0, fine, 0, 2, fbox)
    vb = FArrayBox(fbox, 2, tensor=torch.from_numpy(v).cuda())
    cbar = FArrayBox(cb, 2)
    P.interp_adjoint(cbar, 0, vb, 0, 2, fbox, crse_in=crse)
    torch.cuda.synchronize()
    lhs = float((fine.t * vb.t).sum())
    rhs = float((crse.t * cbar.t).sum())
    assert abs(lhs - rhs) <= 1e-12 * float((fine.t.abs() * vb.t.abs()).sum())


@pytest.mark.parametrize("it", [(1, 1, 1), (0, 1, 1), (1, 0, 0)])
def test_minmod_record_and_recorded_adjoint(it):
    """cxp_prolongate_record writes the plain forward's values and the
    MINMOD branch codes; the adjoint from them equals the recomputing
    adjoint bit for bit (2 components, codes per component)."""
    rng = np.random.default_rng(12)
    rlo, rhi = (-3, 0, 2), (26, 19, 17)
    P, cb, fbox, u, v = _fabs(("minmod", 5, it), rlo, rhi, 2, rng)
    crse = FArrayBox(cb, 2, tensor=torch.from_numpy(u).cuda())
    f1, f2 = FArrayBox(fbox, 2), FArrayBox(fbox, 2)
    P.interp(crse, 0, f1, 0, 2, fbox)
    nb = P.branch_bytes(fbox, 2)
    assert nb > 0
    br = torch.full((nb,), 255, dtype=torch.uint8, device="cuda")
    P.interp_record(crse, 0, f2, 0, 2, fbox, br)
    torch.cuda.synchronize()
    assert torch.equal(f1.t, f2.t)
    assert int(br.max()) <= 2   # every code written, each one of 3 branches
    vb = FArrayBox(fbox, 2, tensor=torch.from_numpy(v).cuda())
    a1, a2 = FArrayBox(cb, 2), FArrayBox(cb, 2)
    P.interp_adjoint(a1, 0, vb, 0, 2, fbox, crse_in=crse)
    P.interp_adjoint_recorded(br, a2, 0, vb, 0, 2, fbox)
    torch.cuda.synchronize()
    assert torch.equal(a1.t, a2.t)


@pytest.mark.parametrize("it", [(1, 1, 1), (1, 0, 1), (0, 1, 1)])
def test_minmod_device_codes_equal_the_host_record(it):
    """The device record (tiled for CCC, by columns else) writes, per
    component, exactly the codes of the per-point host record
    (cxp_host_minmod_codes mode 0), byte for byte."""
    from test_minmod_adjoint_host import host_codes
    rng = np.random.default_rng(31)
    rlo, rhi = (-5, 2, -3), (40, 37, 12)
    P, cb, fbox, u, v = _fabs(("minmod", 3, it), rlo, rhi, 2, rng)
    crse = FArrayBox(cb, 2, tensor=torch.from_numpy(u).cuda())
    fine = FArrayBox(fbox, 2)
    nb = P.branch_bytes(fbox, 2)
    br = torch.full((nb,), 255, dtype=torch.uint8, device="cuda")
    P.interp_record(crse, 0, fine, 0, 2, fbox, br)
    torch.cuda.synchronize()
    got = br.cpu().numpy()
    for comp in range(2):
        bad, want = host_codes(3, it, np.ascontiguousarray(u[comp]), cb.lo, cb.hi, rlo, rhi,
                               0, 1)
        assert bad == 0 and 2 * len(want) == nb
        assert np.array_equal(got[comp * len(want):(comp + 1) * len(want)], want), comp


def test_restrict_adjoint_dot_product():
    from carpetx_amd.interp import restrict, restrict_adjoint
    rng = np.random.default_rng(2)
    for it in [(0, 0, 0), (1, 0, 0), (1, 1, 0), (1, 1, 1)]:
        fbox = Box((-4, -2, 0), (21, 17, 15))
        creg = Box((-2, -1, 0), tuple((fbox.hi[d] - it[d]) // 2 for d in range(3)))
        f = FArrayBox(fbox, 1, tensor=torch.from_numpy(rng.uniform(-1, 1, (1,) + tuple(reversed(fbox.size)))).cuda())
        c = FArrayBox(creg, 1)
        restrict(f, c, it, crse_region=creg)
        w = FArrayBox(creg, 1, tensor=torch.from_numpy(rng.uniform(-1, 1, (1,) + tuple(reversed(creg.size)))).cuda())
        fb = FArrayBox(fbox, 1)
        restrict_adjoint(w, fb, it, 1, creg)
        torch.cuda.synchronize()
        lhs = float((c.t * w.t).sum())
        rhs = float((f.t * fb.t).sum())
        assert abs(lhs - rhs) < 1e-12 * float((c.t.abs() * w.t.abs()).sum()), it
