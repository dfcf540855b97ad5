"""The CCC ENO routines (carpetx_amd/csrc/eno_ccc.hpp), built for the host
with every access checked (carpetx_amd/host/eno_ccc_host.cpp runs the
kernel's workgroups and phases one after another): the forward against the
oracle bit for bit, and the adjoint from the forward's branch codes against
the transpose of the oracle's Jacobian by central differences (exact
within one shift: the values are linear in the coarse data for fixed
shifts, and h = 2^-28 on dyadic data flips none for these seeds)."""
import ctypes
import os

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "carpetx_amd", "host", "libeno_ccc_host.so")
I3 = ctypes.c_int * 3
PD = ctypes.POINTER(ctypes.c_double)
PU = ctypes.POINTER(ctypes.c_ubyte)
IT = (1, 1, 1)


def host_lib():
    if not os.path.exists(LIB):
        pytest.fail("%s is not built (run __graft_entry__.build())" % LIB)
    L = ctypes.CDLL(LIB)
    L.cxp_host_eno_ccc_forward.restype = ctypes.c_longlong
    L.cxp_host_eno_ccc_forward.argtypes = [PD, I3, I3, PD, PU, I3, I3]
    L.cxp_host_eno_ccc_adjoint.restype = ctypes.c_longlong
    L.cxp_host_eno_ccc_adjoint.argtypes = [PU, PD, I3, I3, PD, I3, I3]
    return L


def _shape(lo, hi):
    return tuple(hi[d] - lo[d] + 1 for d in (2, 1, 0))


def host_forward(crse, clo, chi, rlo, rhi):
    fine = np.full(_shape(rlo, rhi), np.nan)
    codes = np.full(_shape(rlo, rhi), 255, np.uint8)
    bad = host_lib().cxp_host_eno_ccc_forward(
        crse.ctypes.data_as(PD), I3(*clo), I3(*chi), fine.ctypes.data_as(PD),
        codes.ctypes.data_as(PU), I3(*rlo), I3(*rhi))
    return bad, fine, codes


def host_adjoint(codes, fbar, rlo, rhi, clo, chi):
    cbar = np.zeros(_shape(clo, chi))
    bad = host_lib().cxp_host_eno_ccc_adjoint(
        codes.ctypes.data_as(PU), fbar.ctypes.data_as(PD), I3(*rlo), I3(*rhi),
        cbar.ctypes.data_as(PD), I3(*clo), I3(*chi))
    return bad, cbar


def oracle_forward(crse, clo, chi, rlo, rhi):
    op = oracle.lookup("eno", 3, IT)
    c = oracle.Fab(clo, chi, 1, crse.reshape((1,) + crse.shape))
    f = oracle.Fab(rlo, rhi, 1)
    oracle.prolongate(op, c, f, rlo, rhi)
    return f.arr[0]


def dyadic(rng, shape):
    return rng.integers(-(1 << 24), 1 << 24, size=shape).astype(np.float64) / float(1 << 24)


@pytest.mark.parametrize("rlo,rhi", [((0, 0, 0), (15, 15, 7)),          # one tile
                                      ((-5, 2, -3), (40, 37, 12)),       # ragged, odd
                                      ((3, -7, 1), (4, -6, 2)),          # 2 x 2 x 2
                                      ((-1, -1, -1), (30, 17, 20))])
def test_host_forward_equals_the_oracle(rlo, rhi):
    op = oracle.lookup("eno", 3, IT)
    clo, chi = oracle.coarse_box(op, rlo, rhi)
    rng = np.random.default_rng(abs(sum(rlo)) + 7 * abs(sum(rhi)))
    crse = rng.uniform(-1, 1, _shape(clo, chi))
    # a few kinks and flat stretches, where shifts change and scores tie
    crse[:, :, ::5] = np.abs(crse[:, :, ::5])
    crse[::3] = 0.25
    bad, fine, codes = host_forward(crse, clo, chi, rlo, rhi)
    assert bad == 0
    assert np.array_equal(fine, oracle_forward(crse, clo, chi, rlo, rhi))
    assert codes.max() <= 26


def test_host_adjoint_equals_the_jacobian_transpose():
    rlo, rhi = (-3, 1, -2), (8, 6, 5)
    op = oracle.lookup("eno", 3, IT)
    clo, chi = oracle.coarse_box(op, rlo, rhi)
    rng = np.random.default_rng(17)
    crse = dyadic(rng, _shape(clo, chi))
    fbar = dyadic(rng, _shape(rlo, rhi))
    bad, _, codes = host_forward(crse, clo, chi, rlo, rhi)
    assert bad == 0
    bad, got = host_adjoint(codes, fbar, rlo, rhi, clo, chi)
    assert bad == 0
    h = 2.0 ** -28
    u = crse.reshape(-1)
    want = np.empty(u.size)
    for n in range(u.size):
        e = np.zeros(u.size)
        e[n] = h
        d = (oracle_forward((u + e).reshape(crse.shape), clo, chi, rlo, rhi) -
             oracle_forward((u - e).reshape(crse.shape), clo, chi, rlo, rhi)) / (2 * h)
        want[n] = np.dot(d.reshape(-1), fbar.reshape(-1))
    want = want.reshape(crse.shape)
    assert np.max(np.abs(got - want)) <= 1e-12 * max(1.0, np.max(np.abs(want)))


def host_adjoint_tiled(codes, fbar, rlo, rhi, clo, chi, zc):
    L = host_lib()
    L.cxp_host_eno_ccc_adjoint_tiled.restype = ctypes.c_longlong
    L.cxp_host_eno_ccc_adjoint_tiled.argtypes = [PU, PD, I3, I3, PD, I3, I3, ctypes.c_int]
    cbar = np.zeros(_shape(clo, chi))
    bad = L.cxp_host_eno_ccc_adjoint_tiled(
        codes.ctypes.data_as(PU), fbar.ctypes.data_as(PD), I3(*rlo), I3(*rhi),
        cbar.ctypes.data_as(PD), I3(*clo), I3(*chi), zc)
    return bad, cbar


@pytest.mark.parametrize("rlo,rhi,zc", [((-3, 1, -2), (8, 6, 5), 1),
                                         ((-5, 2, -3), (40, 37, 12), 4),
                                         ((3, -7, 1), (4, -6, 2), 4),
                                         ((0, 0, 0), (31, 33, 29), 7),
                                         ((-1, -1, -1), (30, 17, 20), 100)])
def test_host_tiled_adjoint_equals_the_per_point_sum(rlo, rhi, zc):
    """The tiled gather (LDS partial sums per parent, sums per z-shift, a
    five-plane register window per output column) against the
    per-coarse-point sum: the same terms grouped by parent plane and z-shift,
    so equal to rounding; every access checked."""
    op = oracle.lookup("eno", 3, IT)
    clo, chi = oracle.coarse_box(op, rlo, rhi)
    rng = np.random.default_rng(abs(sum(rlo)) + 3 * abs(sum(rhi)) + zc)
    crse = rng.uniform(-1, 1, _shape(clo, chi))
    crse[:, :, ::5] = np.abs(crse[:, :, ::5])
    crse[::3] = 0.25
    bad, _, codes = host_forward(crse, clo, chi, rlo, rhi)
    assert bad == 0
    # every shift occurs
    assert len(np.unique(codes)) == 27 or min(_shape(rlo, rhi)) < 8
    fbar = rng.uniform(-1, 1, _shape(rlo, rhi))
    bad, want = host_adjoint(codes, fbar, rlo, rhi, clo, chi)
    assert bad == 0
    bad, got = host_adjoint_tiled(codes, fbar, rlo, rhi, clo, chi, zc)
    assert bad == 0
    assert np.max(np.abs(got - want)) <= 1e-14 * max(1.0, np.max(np.abs(want)))


def test_host_tiled_adjoint_catches_a_dropped_parent_plane():
    """Sanity of the comparison: the per-point sum with one parent plane's
    fine adjoint zeroed differs from the tiled gather of the full input."""
    rlo, rhi = (0, 0, 0), (15, 15, 15)
    op = oracle.lookup("eno", 3, IT)
    clo, chi = oracle.coarse_box(op, rlo, rhi)
    rng = np.random.default_rng(3)
    crse = rng.uniform(-1, 1, _shape(clo, chi))
    bad, _, codes = host_forward(crse, clo, chi, rlo, rhi)
    fbar = rng.uniform(-1, 1, _shape(rlo, rhi))
    f2 = fbar.copy()
    f2[6:8] = 0.0
    _, want = host_adjoint(codes, f2, rlo, rhi, clo, chi)
    _, got = host_adjoint_tiled(codes, fbar, rlo, rhi, clo, chi, 4)
    assert np.max(np.abs(got - want)) > 1e-3
