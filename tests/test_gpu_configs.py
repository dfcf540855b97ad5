"""GPU checks of the benchmark configurations at their configured sizes.

* One AMR level: 64 fine boxes of 32^3 on a 128^3 coarse level, 4 grid
  functions, CCC CONS-2 and VVV POLY-5 (the batched list): every box and
  component against the oracle, bit for bit.
* Adjoint check: coarse 96^3 -> fine 192^3 with L = 1/2 ||P u - t||^2, u and
  t uniform[-1, 1] (seeds 4 and 5): the gradient P^T (P u - t) against the
  central difference of L along a seeded direction (h = 1e-6), the dot-product
  identity at full size, and an explicit dense P^T on an 8^3 fine sub-box of
  the same run; VVV POLY-5 and CCC MINMOD.
* Differentiable evolution: the two-level wave at 128^3 / 128^3 over 16
  steps with checkpointing every 4 (gradient against the plain backward and
  the central difference), and its solution against the exact spherical
  wave at two resolutions.
"""
import math

import numpy as np
import pytest
import torch

import oracle
from carpetx_amd import workloads
from carpetx_amd.batched import WorkList
from carpetx_amd.interp import Box, FArrayBox, get_interpolator

pytestmark = pytest.mark.gpu


def test_level64_every_box_and_component_equals_the_oracle():
    lv = workloads.level_64()
    for name, d in lv.items():
        wl = WorkList()
        for fab, reg in zip(d["fines"], d["regions"]):
            wl.append(d["op"], d["crse"], 0, fab, 0, d["ncomp"], reg)
        wl.upload()
        wl.run()
        torch.cuda.synchronize()
        op = d["op"]
        oop = oracle.lookup(op.prolongation_type, op.prolongation_order, op.indextype)
        ch = oracle.Fab(d["crse"].box.lo, d["crse"].box.hi, d["ncomp"], d["crse"].t.cpu().numpy())
        items, outs = [], []
        for reg in d["regions"]:
            fh = oracle.Fab(reg.lo, reg.hi, d["ncomp"])
            items.append((oop, ch, fh, reg.lo, reg.hi, d["ncomp"]))
            outs.append(fh)
        oracle.prolongate_many(items, kind="strict")
        assert len(outs) == 64
        for n, (fab, fh) in enumerate(zip(d["fines"], outs)):
            assert np.array_equal(fab.t.cpu().numpy(), fh.arr), (name, n)
        wl.close()


REGION = Box((0, 0, 0), (191, 191, 191))


def _uniform(shape, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    return torch.rand(shape, dtype=torch.float64, device="cuda", generator=g) * 2 - 1


def c4_setup(typ, order, it):
    P = get_interpolator(typ, order, it)
    cb = P.CoarseBox(REGION)
    u = FArrayBox(cb, 1, tensor=_uniform((1,) + tuple(reversed(cb.size)), 4))
    t = _uniform((1,) + tuple(reversed(REGION.size)), 5)
    return P, cb, u, t


def forward(P, u, cb):
    f = FArrayBox(REGION, 1)
    P.interp(u, 0, f, 0, 1, REGION)
    return f


def residual(P, u, cb, t):
    return forward(P, u, cb).t - t


def c4_gradient_checks(P, cb, u, t, fd_rtol):
    r = residual(P, u, cb, t)
    rf = FArrayBox(REGION, 1, tensor=r.contiguous())
    grad = FArrayBox(cb, 1)
    P.interp_adjoint(grad, 0, rf, 0, 1, REGION, crse_in=u)
    d = _uniform(tuple(u.t.shape), 6)
    h = 1e-6
    up = FArrayBox(cb, 1, tensor=u.t + h * d)
    um = FArrayBox(cb, 1, tensor=u.t - h * d)
    rp = residual(P, up, cb, t)
    rm = residual(P, um, cb, t)
    # central difference of L = 1/2 sum r^2, differenced per point before the
    # sum so that the O(1e6) loss does not cancel in the subtraction
    fd = float(((rp - rm) * (rp + rm)).sum()) / (4 * h)
    ad = float((grad.t * d).sum())
    assert abs(fd - ad) <= fd_rtol * abs(ad), (fd, ad)
    return r, grad


@pytest.mark.parametrize("typ,order,it,fd_rtol", [
    ("interpolate", 5, (0, 0, 0), 1e-7),
    # MINMOD is piecewise linear: with ~6e6 branch decisions a few lie within
    # h of a kink and bend the central difference by O(1) each, against a
    # directional derivative of O(sqrt(7e6)); hence the looser bound
    ("minmod", 1, (1, 1, 1), 5e-3)])
def test_c4_adjoint_at_size(typ, order, it, fd_rtol):
    P, cb, u, t = c4_setup(typ, order, it)
    assert cb.size == ((103, 103, 103) if typ == "interpolate" else (98, 98, 98))
    r, grad = c4_gradient_checks(P, cb, u, t, fd_rtol=fd_rtol)
    # dot-product identity, full size: <P v, w> = <v, P^T w> (for MINMOD at
    # v, since P(v) = J(v) v for a positively homogeneous limiter)
    v = FArrayBox(cb, 1, tensor=_uniform(tuple(u.t.shape), 7))
    w = FArrayBox(REGION, 1, tensor=_uniform((1,) + tuple(reversed(REGION.size)), 8))
    pv = forward(P, v, cb)
    ptw = FArrayBox(cb, 1)
    P.interp_adjoint(ptw, 0, w, 0, 1, REGION, crse_in=v)
    lhs = float((pv.t * w.t).sum())
    rhs = float((v.t * ptw.t).sum())
    assert abs(lhs - rhs) <= 1e-12 * float((pv.t.abs() * w.t.abs()).sum())
    # dense P^T on an 8^3 fine sub-box of the same run: the oracle's forward
    # on coarse unit vectors (linear), or its Jacobian at u by central
    # differences with h = 2^-30 (MINMOD)
    sub = Box((92, 93, 94), (99, 100, 101))
    sb = P.CoarseBox(sub)
    rs = FArrayBox(REGION, 1, tensor=r.contiguous())
    gsub = FArrayBox(cb, 1)
    P.interp_adjoint(gsub, 0, rs, 0, 1, sub, crse_in=u)
    torch.cuda.synchronize()
    rsub = r[0, sub.lo[2]:sub.hi[2] + 1, sub.lo[1]:sub.hi[1] + 1,
             sub.lo[0]:sub.hi[0] + 1].cpu().numpy()
    if typ == "minmod":
        from test_minmod_adjoint_host import oracle_jt
        usub = u.view(sb)[0].cpu().numpy().copy()
        want = oracle_jt(order, it, usub, sb.lo, sb.hi, rsub.copy(), sub.lo, sub.hi,
                         h=2.0 ** -30)
        tol = 1e-6
    else:
        op = oracle.lookup(typ, order, it)
        ncol = sb.numPts
        fshape = tuple(reversed(sub.size))
        A = np.empty((int(np.prod(fshape)), ncol))
        for n in range(ncol):
            e = np.zeros(ncol)
            e[n] = 1.0
            ch = oracle.Fab(sb.lo, sb.hi, 1, e.reshape((1,) + tuple(reversed(sb.size))))
            fh = oracle.Fab(sub.lo, sub.hi, 1)
            oracle.prolongate(op, ch, fh, sub.lo, sub.hi)
            A[:, n] = fh.arr.reshape(-1)
        want = (A.T @ rsub.reshape(-1)).reshape(tuple(reversed(sb.size)))
        tol = 1e-13
    got = gsub.view(sb)[0].cpu().numpy()
    assert np.max(np.abs(got - want)) <= tol * max(1.0, np.max(np.abs(want)))
    # nothing outside CoarseBox(sub) was touched
    rest = gsub.t.clone()
    rest[(slice(None),) + tuple(slice(sb.lo[d] - cb.lo[d], sb.hi[d] - cb.lo[d] + 1)
                                for d in (2, 1, 0))] = 0
    assert float(rest.abs().max()) == 0.0


# ---- differentiable evolution: coarse 128^3, fine 128^3 over the central
# half, 16 coarse steps (4 RK4 substeps each, the fine level subcycled),
# gradient w.r.t. the initial data with checkpointing every 4 steps

def _loss_difference(ap, bp, am, bm):
    """L(+) - L(-) of carpetx_amd.wave.loss, differenced per point before the
    sums so that the loss itself does not cancel in the subtraction."""
    dc = ((ap[0] - am[0]) * (ap[0] + am[0])).sum()
    fi = (slice(2, -2),) * 3
    df = ((bp[0][fi] - bm[0][fi]) * (bp[0][fi] + bm[0][fi])).sum()
    return 0.5 * float(dc) + 0.5 * float(df) / 8.0


def test_c5_wave_gradient_at_size():
    from carpetx_amd.wave import TwoLevelWave, loss
    w = TwoLevelWave(n=128, native=True)
    assert w.fine_int.size == (129, 129, 129)
    sc0, sf0 = w.gaussian(width=0.05)
    steps = 16
    grads = {}
    for ck in (4, 0):
        xc = sc0.clone().requires_grad_(True)
        xf = sf0.clone().requires_grad_(True)
        a, b = w.evolve(xc, xf, steps, checkpoint_every=ck)
        loss(a, b).backward()
        grads[ck] = (xc.grad.detach().clone(), xf.grad.detach().clone())
        del xc, xf, a, b
    for g4, g0 in zip(grads[4], grads[0]):
        assert float(g0.abs().max()) > 0
        assert float((g4 - g0).abs().max()) <= 1e-12 * float(g0.abs().max())
    g = torch.Generator(device="cuda").manual_seed(9)
    vc = torch.rand(sc0.shape, dtype=torch.float64, device="cuda", generator=g) - 0.5
    vf = torch.rand(sf0.shape, dtype=torch.float64, device="cuda", generator=g) - 0.5
    h = 1e-6
    with torch.no_grad():
        ap, bp = w.evolve(sc0 + h * vc, sf0 + h * vf, steps)
        am, bm = w.evolve(sc0 - h * vc, sf0 - h * vf, steps)
    fd = _loss_difference(ap, bp, am, bm) / (2 * h)
    gc, gf = grads[4]
    ad = float((gc * vc).sum() + (gf * vf).sum())
    assert abs(fd - ad) <= 1e-7 * abs(ad), (fd, ad)


def _spherical_wave(x, y, z, t, width):
    """u(r, t) of u_tt = Laplace u with u(r, 0) = exp(-r^2 / (2 width^2)),
    u_t(r, 0) = 0 in free space: r u solves the 1-D wave equation, so
    u = ((r - t) g(r - t) + (r + t) g(r + t)) / (2 r), and at r = 0 the limit
    g(t) (1 - t^2 / width^2)."""
    r = torch.sqrt(x * x + y * y + z * z)
    g = lambda s: torch.exp(-s * s / (2 * width * width))
    rs = torch.where(r > 0, r, torch.ones_like(r))
    u = ((r - t) * g(r - t) + (r + t) * g(r + t)) / (2 * rs)
    u0 = math.exp(-t * t / (2 * width * width)) * (1 - t * t / (width * width))
    return torch.where(r > 0, u, torch.full_like(r, u0))


def test_c5_wave_converges_to_the_spherical_wave():
    """The two-level evolution against the exact solution at t = 1/16, for
    n = 64 and 128 (16 and 32 coarse steps).  Inside |x|_inf <= 0.15 no
    signal from the coarse/fine interface (at 0.25) has arrived by then, and
    the fine level converges at 4th order like its stencils.  Over the whole
    fine level the ghost values, interpolated linearly in time between
    coarse steps, limit the order to 1 at the interface: the test pins that
    too, as the harness's known behaviour."""
    from carpetx_amd.wave import TwoLevelWave
    width, T = 0.1, 1.0 / 16
    err_in, err_all = {}, {}
    for n in (64, 128):
        w = TwoLevelWave(n=n)
        sc, sf = w.gaussian(width=width)
        steps = int(round(T / w.dt))
        assert abs(steps * w.dt - T) < 1e-15
        with torch.no_grad():
            for _ in range(steps):
                sc, sf = w.step(sc, sf)
        fi, fb = w.fine_int, w.fine_box
        ax = [torch.arange(fi.lo[d], fi.hi[d] + 1, dtype=torch.float64, device="cuda")
              * (w.h / 2) - 0.5 for d in range(3)]
        Z, Y, X = torch.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
        sl = tuple(slice(fi.lo[d] - fb.lo[d], fi.hi[d] - fb.lo[d] + 1) for d in (2, 1, 0))
        e = (sf[0][sl] - _spherical_wave(X, Y, Z, T, width)).abs()
        inner = (X.abs() <= 0.15) & (Y.abs() <= 0.15) & (Z.abs() <= 0.15)
        err_in[n] = float(e[inner].max())
        err_all[n] = float(e.max())
    assert err_in[64] / err_in[128] >= 12.0, err_in
    assert err_in[128] <= 1e-6, err_in
    assert err_all[64] / err_all[128] >= 1.6, err_all
    assert err_all[128] <= 1e-4, err_all
