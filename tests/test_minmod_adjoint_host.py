"""The MINMOD adjoint's record and gather routines, built for the host with
every access bounds-checked (carpetx_amd/host/minmod_adjoint_host.cpp runs
the per-point functions of carpetx_amd/csrc/minmod_adjoint.hpp that the
device kernels run), against the transpose of the oracle's Jacobian.

The forward is piecewise linear in the coarse data, so central differences
give the Jacobian exactly inside one limiter branch.  The data are dyadic
(k / 2^24), the step h = 2^-28, so the differences are exact and no
branch flips within +-h; ties, where the Jacobian is undefined, do not
occur for these seeds."""
import ctypes
import os

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "carpetx_amd", "host", "libminmod_adjoint_host.so")

I3 = ctypes.c_int * 3
PD = ctypes.POINTER(ctypes.c_double)


def host_lib():
    if not os.path.exists(LIB):
        pytest.fail("%s is not built (run __graft_entry__.build())" % LIB)
    L = ctypes.CDLL(LIB)
    L.cxp_host_minmod_adjoint.restype = ctypes.c_longlong
    L.cxp_host_minmod_adjoint.argtypes = [ctypes.c_int, I3, PD, I3, I3, PD, I3, I3, PD,
                                          ctypes.POINTER(ctypes.c_longlong)]
    return L


def _ptr(a):
    return a.ctypes.data_as(PD)


def host_adjoint(order, it, crse, clo, chi, fbar, rlo, rhi):
    cbar = np.zeros_like(crse)
    n = ctypes.c_longlong()
    bad = host_lib().cxp_host_minmod_adjoint(order, I3(*it), _ptr(crse), I3(*clo), I3(*chi),
                                             _ptr(fbar), I3(*rlo), I3(*rhi), _ptr(cbar),
                                             ctypes.byref(n))
    return bad, cbar, n.value


def oracle_jt(order, it, crse, clo, chi, fbar, rlo, rhi, h=2.0 ** -28):
    op = oracle.lookup("minmod", order, it)
    shape = crse.shape

    def fwd(x):
        c = oracle.Fab(clo, chi, 1, x.reshape((1,) + shape))
        f = oracle.Fab(rlo, rhi, 1)
        oracle.prolongate(op, c, f, rlo, rhi)
        return f.arr.reshape(-1)
    u = crse.reshape(-1)
    v = fbar.reshape(-1)
    out = np.empty(u.size)
    for n in range(u.size):
        e = np.zeros(u.size)
        e[n] = h
        out[n] = np.dot((fwd(u + e) - fwd(u - e)) / (2 * h), v)
    return out.reshape(shape)


def dyadic(rng, shape):
    return rng.integers(-(1 << 24), 1 << 24, size=shape).astype(np.float64) / float(1 << 24)


@pytest.mark.parametrize("order", [1, 3, 5])
@pytest.mark.parametrize("c", range(8))
def test_host_minmod_adjoint_equals_the_jacobian_transpose(order, c):
    it = ((c >> 2) & 1, (c >> 1) & 1, c & 1)
    rlo, rhi = (-3, 1, -2), (8, 6, 5)
    op = oracle.lookup("minmod", order, it)
    clo, chi = oracle.coarse_box(op, rlo, rhi)
    rng = np.random.default_rng(1000 + 10 * order + c)
    crse = dyadic(rng, tuple(chi[d] - clo[d] + 1 for d in (2, 1, 0)))
    fbar = dyadic(rng, tuple(rhi[d] - rlo[d] + 1 for d in (2, 1, 0)))
    bad, got, ncodes = host_adjoint(order, it, crse, clo, chi, fbar, rlo, rhi)
    assert bad == 0
    # codes only for the MINMOD (cell-centred) directions
    assert (ncodes > 0) == any(it)
    want = oracle_jt(order, it, crse, clo, chi, fbar, rlo, rhi)
    err = np.max(np.abs(got - want))
    assert err <= 1e-12 * max(1.0, np.max(np.abs(want))), (order, it, err)


def test_host_minmod_adjoint_larger_box_with_a_coarse_margin():
    """A coarse FAB larger than CoarseBox(region): the adjoint touches only
    the output window and leaves the margin at zero."""
    order, it = 3, (1, 0, 1)
    rlo, rhi = (-5, 2, 1), (13, 9, 10)
    op = oracle.lookup("minmod", order, it)
    cb_lo, cb_hi = oracle.coarse_box(op, rlo, rhi)
    clo = tuple(x - 2 for x in cb_lo)
    chi = tuple(x + 1 for x in cb_hi)
    rng = np.random.default_rng(77)
    crse = dyadic(rng, tuple(chi[d] - clo[d] + 1 for d in (2, 1, 0)))
    fbar = dyadic(rng, tuple(rhi[d] - rlo[d] + 1 for d in (2, 1, 0)))
    bad, got, _ = host_adjoint(order, it, crse, clo, chi, fbar, rlo, rhi)
    assert bad == 0
    want = oracle_jt(order, it, crse, clo, chi, fbar, rlo, rhi)
    assert np.max(np.abs(got - want)) <= 1e-12 * max(1.0, np.max(np.abs(want)))
    inner = tuple(slice(cb_lo[d] - clo[d], cb_hi[d] - clo[d] + 1) for d in (2, 1, 0))
    margin = np.ones(got.shape, dtype=bool)
    margin[inner] = False
    assert np.all(got[margin] == 0.0)


def host_adjoint_tiled(order, it, crse, clo, chi, fbar, rlo, rhi, zc, kc):
    L = host_lib()
    fn = L.cxp_host_minmod_adjoint_tiled
    fn.restype = ctypes.c_longlong
    fn.argtypes = [ctypes.c_int, I3, PD, I3, I3, PD, I3, I3, PD, ctypes.c_int, ctypes.c_int,
                   ctypes.POINTER(ctypes.c_longlong)]
    cbar = np.zeros_like(crse)
    n = ctypes.c_longlong()
    bad = fn(order, I3(*it), _ptr(crse), I3(*clo), I3(*chi), _ptr(fbar), I3(*rlo), I3(*rhi),
             _ptr(cbar), zc, kc, ctypes.byref(n))
    return bad, cbar


@pytest.mark.parametrize("order", [1, 3, 5])
@pytest.mark.parametrize("c", range(8))
def test_host_tiled_routines_equal_the_per_point_gather(order, c):
    """The kernels' column record (z-chunks of 3) and, for CCC, the tiled
    gather (several 16 x 16 tiles, z-chunks of 2), run on the host with
    every access checked: no out-of-range access, every code written once,
    and the result of the per-point gather (pinned to the oracle's
    Jacobian above) to rounding."""
    it = ((c >> 2) & 1, (c >> 1) & 1, c & 1)
    rlo, rhi = (-5, 2, -3), (40, 37, 8)
    op = oracle.lookup("minmod", order, it)
    clo, chi = oracle.coarse_box(op, rlo, rhi)
    rng = np.random.default_rng(300 + 10 * order + c)
    crse = rng.uniform(-1, 1, tuple(chi[d] - clo[d] + 1 for d in (2, 1, 0)))
    fbar = rng.uniform(-1, 1, tuple(rhi[d] - rlo[d] + 1 for d in (2, 1, 0)))
    bad0, want, _ = host_adjoint(order, it, crse, clo, chi, fbar, rlo, rhi)
    assert bad0 == 0
    for zc, kc in ((3, 2), (64, 16)):
        bad, got = host_adjoint_tiled(order, it, crse, clo, chi, fbar, rlo, rhi, zc, kc)
        assert bad == 0, (zc, kc)
        err = np.max(np.abs(got - want))
        assert err <= 1e-13 * max(1.0, np.max(np.abs(want))), (order, it, zc, kc, err)


def test_host_tiled_gather_against_the_jacobian():
    rlo, rhi = (1, -2, 0), (34, 9, 7)
    it = (1, 1, 1)
    op = oracle.lookup("minmod", 3, it)
    clo, chi = oracle.coarse_box(op, rlo, rhi)
    rng = np.random.default_rng(77)
    crse = dyadic(rng, tuple(chi[d] - clo[d] + 1 for d in (2, 1, 0)))
    fbar = dyadic(rng, tuple(rhi[d] - rlo[d] + 1 for d in (2, 1, 0)))
    bad, got = host_adjoint_tiled(3, it, crse, clo, chi, fbar, rlo, rhi, 2, 3)
    assert bad == 0
    want = oracle_jt(3, it, crse, clo, chi, fbar, rlo, rhi)
    assert np.max(np.abs(got - want)) <= 1e-12 * max(1.0, np.max(np.abs(want)))


def host_adjoint_fused(order, it, crse, clo, chi, fbar, rlo, rhi, kc):
    L = host_lib()
    fn = L.cxp_host_minmod_adjoint_fused
    fn.restype = ctypes.c_longlong
    fn.argtypes = [ctypes.c_int, I3, PD, I3, I3, PD, I3, I3, PD, ctypes.c_int]
    cbar = np.zeros_like(crse)
    bad = fn(order, I3(*it), _ptr(crse), I3(*clo), I3(*chi), _ptr(fbar), I3(*rlo), I3(*rhi),
             _ptr(cbar), kc)
    return bad, cbar


@pytest.mark.parametrize("order", [1, 3, 5])
@pytest.mark.parametrize("rlo,rhi", [((-5, 2, -3), (40, 37, 8)),      # ragged, odd corners
                                      ((0, 0, 0), (31, 31, 15)),      # tile-aligned
                                      ((3, -7, 1), (4, -6, 2))])      # 2 x 2 x 2 fine points
def test_host_fused_adjoint_equals_record_and_tiled_gather(order, rlo, rhi):
    """The fused CCC march (branches recomputed per workgroup, rings of
    codes in LDS): no failed check (every LDS read sees the value of its
    plane written in an earlier phase, no element written where it is read
    or twice in a phase, no access out of range) and the bits of record +
    tiled gather, for z-chunks of 1, 2, 5 and 16 planes."""
    it = (1, 1, 1)
    op = oracle.lookup("minmod", order, it)
    clo, chi = oracle.coarse_box(op, rlo, rhi)
    rng = np.random.default_rng(900 + order)
    crse = rng.uniform(-1, 1, tuple(chi[d] - clo[d] + 1 for d in (2, 1, 0)))
    fbar = rng.uniform(-1, 1, tuple(rhi[d] - rlo[d] + 1 for d in (2, 1, 0)))
    bad0, want = host_adjoint_tiled(order, it, crse, clo, chi, fbar, rlo, rhi, 4, 3)
    assert bad0 == 0
    for kc in (1, 2, 5, 16):
        bad, got = host_adjoint_fused(order, it, crse, clo, chi, fbar, rlo, rhi, kc)
        assert bad == 0, kc
        assert np.array_equal(got, want), kc


def test_host_fused_adjoint_against_the_jacobian():
    rlo, rhi = (1, -2, 0), (34, 9, 7)
    it = (1, 1, 1)
    op = oracle.lookup("minmod", 3, it)
    clo, chi = oracle.coarse_box(op, rlo, rhi)
    rng = np.random.default_rng(78)
    crse = dyadic(rng, tuple(chi[d] - clo[d] + 1 for d in (2, 1, 0)))
    fbar = dyadic(rng, tuple(rhi[d] - rlo[d] + 1 for d in (2, 1, 0)))
    bad, got = host_adjoint_fused(3, it, crse, clo, chi, fbar, rlo, rhi, 3)
    assert bad == 0
    want = oracle_jt(3, it, crse, clo, chi, fbar, rlo, rhi)
    assert np.max(np.abs(got - want)) <= 1e-12 * max(1.0, np.max(np.abs(want)))
    assert host_adjoint_fused(3, (1, 0, 1), crse, clo, chi, fbar, rlo, rhi, 3)[0] == -1


def host_codes(order, it, crse, clo, chi, rlo, rhi, mode, kc):
    L = host_lib()
    fn = L.cxp_host_minmod_codes
    fn.restype = ctypes.c_longlong
    fn.argtypes = [ctypes.c_int, I3, PD, I3, I3, I3, I3, ctypes.c_int, ctypes.c_int,
                   ctypes.c_void_p, ctypes.c_longlong, ctypes.POINTER(ctypes.c_longlong)]
    cap = 8 * np.prod([rhi[d] - rlo[d] + 8 for d in range(3)])
    out = np.full(int(cap), 255, np.uint8)
    n = ctypes.c_longlong()
    bad = fn(order, I3(*it), _ptr(crse), I3(*clo), I3(*chi), I3(*rlo), I3(*rhi), mode, kc,
             out.ctypes.data, int(cap), ctypes.byref(n))
    return bad, out[:n.value]


@pytest.mark.parametrize("rlo,rhi", [((-5, 2, -3), (40, 37, 8)), ((0, 0, 0), (27, 27, 13)),
                                      ((3, -7, 1), (4, -6, 2))])
@pytest.mark.parametrize("c", range(8))
def test_host_code_records_agree(rlo, rhi, c):
    """The column record (every centring) and the tiled record (CCC) write
    every code exactly once, with no out-of-range or unsynchronised access,
    and give the per-point record's codes, for several z-chunks."""
    it = ((c >> 2) & 1, (c >> 1) & 1, c & 1)
    op = oracle.lookup("minmod", 3, it)
    clo, chi = oracle.coarse_box(op, rlo, rhi)
    rng = np.random.default_rng(40 + c)
    crse = rng.uniform(-1, 1, tuple(chi[d] - clo[d] + 1 for d in (2, 1, 0)))
    bad, want = host_codes(3, it, crse, clo, chi, rlo, rhi, 0, 1)
    assert bad == 0 and (len(want) > 0) == any(it)
    assert np.all(want <= 2)
    for mode, kcs in ((1, (1, 5, 12)), (2, (1, 2, 5, 16))):
        if mode == 2 and not all(it):
            continue
        for kc in kcs:
            bad, got = host_codes(3, it, crse, clo, chi, rlo, rhi, mode, kc)
            assert bad == 0, (mode, kc)
            assert np.array_equal(got, want), (mode, kc)
