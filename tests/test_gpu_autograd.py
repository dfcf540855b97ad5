"""GPU autograd harness: torch.autograd.Function wrappers whose backward is
the hand-written adjoint, checked with central finite differences (h=1e-6)
and an explicit dense transpose, and the differentiable two-level wave
evolution with checkpointing."""
import numpy as np
import pytest
import torch

from carpetx_amd.autograd import fill_ghosts, prolongate, restrict
from carpetx_amd.interp import Box, get_interpolator

pytestmark = pytest.mark.gpu


def _fd_directional(f, x, v, h=1e-6):
    with torch.no_grad():
        return (f(x + h * v) - f(x - h * v)) / (2 * h)


@pytest.mark.parametrize("fam,order,it", [("natural", 5, (0, 0, 0)), ("natural", 3, (1, 1, 1)),
                                          ("conservative", 1, (1, 0, 1)),
                                          ("minmod", 3, (1, 1, 1)), ("minmod", 5, (1, 0, 1))])
def test_prolongate_gradient_fd(fam, order, it):
    """L = 0.5 ||P u - t||^2: autograd gradient vs central FD (h=1e-6).
    MINMOD is piecewise linear; for these seeds no branch decision lies
    within h of its kink, so the difference is exact to rounding there too."""
    torch.manual_seed(4)
    P = get_interpolator(fam, order, it)
    region = Box((0, 0, 0), (23, 19, 17))
    cb = P.CoarseBox(region)
    u = (torch.rand((1,) + tuple(reversed(cb.size)), dtype=torch.float64, device="cuda") * 2 - 1)
    t = (torch.rand((1,) + tuple(reversed(region.size)), dtype=torch.float64, device="cuda") * 2 - 1)

    def L(x):
        return 0.5 * ((prolongate(P, x, cb, region) - t) ** 2).sum()
    x = u.clone().requires_grad_(True)
    L(x).backward()
    for seed in range(3):
        g = torch.Generator(device="cuda").manual_seed(seed)
        v = torch.rand(u.shape, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
        fd = float(_fd_directional(L, u, v))
        ad = float((x.grad * v).sum())
        assert abs(fd - ad) <= 1e-7 * max(1.0, abs(ad)), (fd, ad)


@pytest.mark.parametrize("fam,order,it", [("natural", 5, (0, 0, 0)), ("eno", 3, (1, 1, 1))])
def test_fill_ghosts_equals_its_slabs(fam, order, it):
    """The ghost-shell fill as one autograd node: values bit for bit those
    of the six slabs' prolongations copied in; the coarse gradient the sum
    of the slabs' gradients (one buffer, slab after slab, so equal to
    rounding); the fine gradient passes where the fill did not write and is
    zero on the shell."""
    from carpetx_amd.workloads import shell_slabs
    torch.manual_seed(6)
    P = get_interpolator(fam, order, it)
    interior = Box((4, 2, 6), (27, 21, 19))
    fbox = interior.grow(3)
    cb = P.CoarseBox(fbox)
    u = torch.rand((2,) + tuple(reversed(cb.size)), dtype=torch.float64, device="cuda") * 2 - 1
    f = torch.rand((2,) + tuple(reversed(fbox.size)), dtype=torch.float64, device="cuda") * 2 - 1
    g = torch.rand_like(f) * 2 - 1
    x = u.clone().requires_grad_(True)
    y = f.clone().requires_grad_(True)
    out = fill_ghosts(P, x, cb, y, fbox, interior)
    out.backward(g)

    def sl(b):
        return (slice(None),) + tuple(slice(b.lo[d] - fbox.lo[d], b.hi[d] - fbox.lo[d] + 1)
                                      for d in (2, 1, 0))
    x2 = u.clone().requires_grad_(True)
    want = f.clone()
    parts = []
    for slab in shell_slabs(fbox, interior):
        v = prolongate(P, x2, cb, slab)
        want[sl(slab)] = v.detach()
        parts.append((v, g[sl(slab)]))
    assert torch.equal(out.detach(), want)
    torch.autograd.backward([p for p, _ in parts], [q for _, q in parts])
    err = float((x.grad - x2.grad).abs().max())
    assert err <= 1e-14 * max(1.0, float(x2.grad.abs().max())), err
    gf = g.clone()
    for slab in shell_slabs(fbox, interior):
        gf[sl(slab)] = 0.0
    assert torch.equal(y.grad, gf)


def test_restrict_gradient_fd():
    torch.manual_seed(5)
    fbox = Box((0, 0, 0), (17, 15, 13))
    for it in [(0, 0, 0), (1, 1, 1), (1, 0, 0)]:
        creg = Box((0, 0, 0), tuple((fbox.hi[d] - it[d]) // 2 for d in range(3)))
        f = torch.rand((2,) + tuple(reversed(fbox.size)), dtype=torch.float64, device="cuda")
        t = torch.rand((2,) + tuple(reversed(creg.size)), dtype=torch.float64, device="cuda")

        def L(x):
            return 0.5 * ((restrict(x, fbox, creg, it) - t) ** 2).sum()
        x = f.clone().requires_grad_(True)
        L(x).backward()
        v = torch.rand_like(f)
        fd = float(_fd_directional(L, f, v))
        ad = float((x.grad * v).sum())
        assert abs(fd - ad) <= 1e-7 * max(1.0, abs(ad)), (it, fd, ad)


def test_wave_gradient_fd_and_checkpointing():
    """The two-level evolution is linear in the initial data, so the loss is
    quadratic and the central difference equals the directional derivative
    up to rounding; checkpointed and plain backward agree."""
    from carpetx_amd.wave import TwoLevelWave, loss
    w = TwoLevelWave(n=24)
    sc0, sf0 = w.gaussian(width=0.08)

    def run(sc, sf, ck):
        a, b = w.evolve(sc, sf, 3, checkpoint_every=ck)
        return loss(a, b)
    xc = sc0.clone().requires_grad_(True)
    xf = sf0.clone().requires_grad_(True)
    run(xc, xf, 0).backward()
    gc, gf = xc.grad.clone(), xf.grad.clone()
    yc = sc0.clone().requires_grad_(True)
    yf = sf0.clone().requires_grad_(True)
    run(yc, yf, 2).backward()
    assert torch.allclose(yc.grad, gc, rtol=1e-12, atol=1e-14)
    assert torch.allclose(yf.grad, gf, rtol=1e-12, atol=1e-14)
    g = torch.Generator(device="cuda").manual_seed(1)
    vc = torch.rand(sc0.shape, dtype=torch.float64, device="cuda", generator=g) - 0.5
    vf = torch.rand(sf0.shape, dtype=torch.float64, device="cuda", generator=g) - 0.5
    h = 1e-6
    with torch.no_grad():
        fd = (run(sc0 + h * vc, sf0 + h * vf, 0) - run(sc0 - h * vc, sf0 - h * vf, 0)) / (2 * h)
    ad = (gc * vc).sum() + (gf * vf).sum()
    assert abs(float(fd) - float(ad)) <= 1e-7 * max(1.0, abs(float(ad))), (float(fd), float(ad))


def test_wave_energy_bounded():
    """Sanity of the evolution: a Gaussian pulse keeps its energy bounded
    and the fine and coarse solutions agree under the fine level."""
    from carpetx_amd.wave import TwoLevelWave
    w = TwoLevelWave(n=32)
    sc, sf = w.gaussian(width=0.06)
    with torch.no_grad():
        e0 = float((sc[0] ** 2).sum())
        sc, sf = w.evolve(sc, sf, 4)
    assert torch.isfinite(sc).all() and torch.isfinite(sf).all()
    assert float((sc[0] ** 2).sum()) < 2 * e0


def test_wave_graphed_step_replays_match_eager():
    """A whole coarse step captured as one CUDA graph (every RK4 stage's ghost
    fill and the restriction inside it), replayed step after step with eager
    work and synchronisations in between, reproduces the eager evolution bit
    for bit at every step."""
    from carpetx_amd.wave import TwoLevelWave
    w = TwoLevelWave(n=32)
    sc, sf = w.gaussian(width=0.06)
    advance, c, f = w.graphed_step(sc, sf)
    ec, ef = sc.clone(), sf.clone()
    for step in range(4):
        with torch.no_grad():
            ec, ef = w.step(ec, ef)
        torch.cuda.synchronize()
        advance()
        torch.cuda.synchronize()
        assert torch.equal(c, ec), step
        assert torch.equal(f, ef), step


def test_wave_native_step_matches_step():
    """NativeStep (the fused native RK4 stage on both levels, batched ghost
    fills) advances the two-level wave exactly as TwoLevelWave.step does:
    three steps eagerly, and three replays of its CUDA graph, bit for
    bit."""
    from carpetx_amd.wave import NativeStep, TwoLevelWave
    w = TwoLevelWave(n=32)
    sc, sf = w.gaussian(width=0.08)
    with torch.no_grad():
        c, f = sc.clone(), sf.clone()
        ref = []
        for _ in range(3):
            c, f = w.step(c, f)
            ref.append((c.clone(), f.clone()))
        ns = NativeStep(w, sc, sf)
        for k in range(3):
            ns.run()
            torch.cuda.synchronize()
            assert torch.equal(ns.c, ref[k][0]), ("coarse", k)
            assert torch.equal(ns.f, ref[k][1]), ("fine", k)
        ns.close()
        ng = NativeStep(w, sc, sf)
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            ng.run()  # set-up outside capture
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        ng.c.copy_(sc)
        ng.f.copy_(sf)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            ng.run()
        torch.cuda.synchronize()
        for k in range(3):
            graph.replay()
            torch.cuda.synchronize()
            assert torch.equal(ng.c, ref[k][0]), ("graph coarse", k)
            assert torch.equal(ng.f, ref[k][1]), ("graph fine", k)
        ng.close()


def test_wave_native_autograd_matches_torch():
    """TwoLevelWave(native=True): the native stage and its gather-form
    adjoint under autograd.  Two checkpointed steps give the torch path's
    loss bit for bit (the forward is identical) and its gradient with
    respect to both levels' initial data to 1e-11 (the adjoint sums in
    another order)."""
    from carpetx_amd.wave import TwoLevelWave, loss
    out = []
    for native in (False, True):
        w = TwoLevelWave(n=32, native=native)
        sc, sf = w.gaussian(width=0.08)
        xc = sc.clone().requires_grad_(True)
        xf = sf.clone().requires_grad_(True)
        L = loss(*w.evolve(xc, xf, 2, checkpoint_every=1))
        L.backward()
        out.append((L.detach().clone(), xc.grad.clone(), xf.grad.clone()))
    (lt, gct, gft), (ln, gcn, gfn) = out
    assert torch.equal(lt, ln)
    for a, b in ((gct, gcn), (gft, gfn)):
        scale = a.abs().max()
        assert scale > 0
        assert (a - b).abs().max() <= 1e-11 * scale
