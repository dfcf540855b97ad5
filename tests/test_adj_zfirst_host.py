"""The z-first linear adjoint (carpetx_amd/csrc/adj_zfirst.hpp), built for
the host with every access checked (carpetx_amd/host/adj_zfirst_host.cpp
runs the kernel's workgroups and phases one after another): crse_bar +=
P^T fine_bar against the exact transpose of the oracle's prolongation.

P = Pz (x) Py (x) Px for these operators, and each 1-D factor is read off the
oracle: prolongating a coarse field that is a unit spike along direction d
and 1 along the others gives (Pd e_c) (x) 1 (x) 1, since the other
directions reproduce constants (at even fine points for VC CONS 0, whose
odd points are 0).  So the expected adjoint is the einsum of fine_bar with
the three factors, not a restatement of the kernel's sums."""
import ctypes
import os

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "carpetx_amd", "host", "libadj_zfirst_host.so")
I3 = ctypes.c_int * 3
PD = ctypes.POINTER(ctypes.c_double)


def host_lib():
    if not os.path.exists(LIB):
        pytest.fail("%s is not built (run __graft_entry__.build())" % LIB)
    L = ctypes.CDLL(LIB)
    L.cxp_host_adj_zfirst.restype = ctypes.c_longlong
    L.cxp_host_adj_zfirst.argtypes = [ctypes.c_int, PD, I3, I3, PD, I3, I3, ctypes.c_int]
    return L


def _shape(lo, hi):
    return tuple(hi[d] - lo[d] + 1 for d in (2, 1, 0))


def sid_of(op):
    """Stencil id of an operator whose three directions agree (stencil1d.hpp)."""
    dirs = op.dirs()
    assert dirs[0] == dirs[1] == dirs[2]
    cent, intp, order = dirs[0]
    if intp in (oracle.POLY, oracle.HERMITE):
        return {1: 0, 3: 1, 5: 2, 7: 3}[order] if cent == oracle.VC else 4 + order
    assert intp == oracle.CONS
    return 12 + order if cent == oracle.VC else 14 + order


def factor(op, d, rlo, rhi, clo, chi):
    """The 1-D prolongation matrix of direction d: fine index x coarse index."""
    nf, nc = rhi[d] - rlo[d] + 1, chi[d] - clo[d] + 1
    P = np.zeros((nf, nc))
    # a fine point with even coordinates in the other directions
    pick = [rlo[e] + (rlo[e] & 1) if e != d else None for e in range(3)]
    for c in range(nc):
        u = np.zeros(_shape(clo, chi))  # the plane c across direction d
        idx = [slice(None)] * 3
        idx[2 - d] = c
        u[tuple(idx)] = 1.0
        cf = oracle.Fab(clo, chi, 1, u.reshape((1,) + u.shape))
        ff = oracle.Fab(rlo, rhi, 1)
        oracle.prolongate(op, cf, ff, rlo, rhi)
        sel = [pick[e] - rlo[e] if e != d else slice(None) for e in range(3)]
        P[:, c] = ff.arr[0][sel[2], sel[1], sel[0]]
    return P


CASES = [("interpolate", 1, (0, 0, 0)), ("interpolate", 3, (0, 0, 0)),
         ("interpolate", 5, (0, 0, 0)), ("interpolate", 7, (0, 0, 0)),
         ("conservative", 0, (1, 1, 1)), ("conservative", 1, (1, 1, 1)),
         ("conservative", 0, (0, 0, 0)), ("interpolate", 5, (1, 1, 1)),
         ("hermite", 3, (0, 0, 0))]


@pytest.mark.parametrize("typ,order,it", CASES)
@pytest.mark.parametrize("rlo,rhi,kchunk", [((-3, 1, -2), (20, 17, 25), 8),
                                             ((5, -4, 3), (40, 9, 6), 3),
                                             ((1, 2, -5), (30, 27, -3), 2),
                                             ((-4, 0, 0), (-2, 21, 17), 2),
                                             ((0, 0, 0), (33, 35, 31), 100)])
def test_host_zfirst_adjoint_equals_the_transpose(typ, order, it, rlo, rhi, kchunk):
    op = oracle.lookup(typ, order, it)
    sid = sid_of(op)
    clo, chi = oracle.coarse_box(op, rlo, rhi)
    rng = np.random.default_rng(abs(hash((typ, order, it, rlo, kchunk))) % (1 << 32))
    fbar = rng.uniform(-1, 1, _shape(rlo, rhi))
    cbar0 = rng.uniform(-1, 1, _shape(clo, chi))
    got = cbar0.copy()
    bad = host_lib().cxp_host_adj_zfirst(sid, fbar.ctypes.data_as(PD), I3(*rlo), I3(*rhi),
                                         got.ctypes.data_as(PD), I3(*clo), I3(*chi), kchunk)
    assert bad == 0
    Px, Py, Pz = (factor(op, d, rlo, rhi, clo, chi) for d in range(3))
    want = cbar0 + np.einsum("zyx,zk,yj,xi->kji", fbar, Pz, Py, Px)
    assert np.max(np.abs(got - want)) <= 1e-13 * max(1.0, np.max(np.abs(want)))


def test_host_zfirst_catches_a_wrong_factor():
    """The comparison is not vacuous: the expected value built with one
    factor of another operator differs from the kernel's."""
    rlo, rhi = (0, 0, 0), (19, 15, 13)
    op, other = oracle.lookup("interpolate", 5, (0, 0, 0)), oracle.lookup("interpolate", 3, (0, 0, 0))
    clo, chi = oracle.coarse_box(op, rlo, rhi)
    fbar = np.random.default_rng(1).uniform(-1, 1, _shape(rlo, rhi))
    got = np.zeros(_shape(clo, chi))
    assert host_lib().cxp_host_adj_zfirst(sid_of(op), fbar.ctypes.data_as(PD), I3(*rlo), I3(*rhi),
                                          got.ctypes.data_as(PD), I3(*clo), I3(*chi), 8) == 0
    Px = factor(other, 0, rlo, rhi, clo, chi)
    Py, Pz = factor(op, 1, rlo, rhi, clo, chi), factor(op, 2, rlo, rhi, clo, chi)
    want = np.einsum("zyx,zk,yj,xi->kji", fbar, Pz, Py, Px)
    assert np.max(np.abs(got - want)) > 1e-3
