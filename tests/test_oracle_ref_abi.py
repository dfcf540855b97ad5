"""The oracle behind the C interface's own signatures (oracle/cxp_ref.c):
the same operator registry, CoarseBox, argument checks and status codes as
the device library (compared call by call, on the CPU: the device library
checks its arguments before any device work), and the oracle's results."""
import ctypes

import numpy as np

import oracle
from carpetx_amd import _lib
from carpetx_amd._lib import i3


def _fab(lo, hi, ncomp=1, arr=None):
    f = _lib.CxpFab()
    if arr is None:
        f.data = ctypes.c_void_p(0x1000)
    else:
        f.data = ctypes.c_void_p(arr.ctypes.data)
    f.lo, f.hi = i3(lo), i3(hi)
    f.comp_stride = (hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1)
    f.ncomp = ncomp
    return f


def _ref_lookup(t, o, it):
    h = ctypes.c_void_p()
    st = oracle.lib().cxp_ref_op_lookup(t.encode(), o, i3(it), ctypes.byref(h))
    return st, h


def _dev_lookup(t, o, it):
    h = ctypes.c_void_p()
    st = _lib.lib().cxp_op_lookup(t.encode(), o, i3(it), ctypes.byref(h))
    return st, h


def test_registry_and_coarse_box_agree_with_the_device_library():
    R, D = oracle.lib(), _lib.lib()
    for fam in list(oracle.FAMILIES) + ["no-such-type"]:
        for order in range(-1, 9):
            for c in range(8):
                it = ((c >> 2) & 1, (c >> 1) & 1, c & 1)
                sr, hr = _ref_lookup(fam, order, it)
                sd, hd = _dev_lookup(fam, order, it)
                assert sr == sd, (fam, order, it, sr, sd)
                if sr:
                    continue
                for flo, fhi in [((0, 0, 0), (63, 63, 63)), ((-3, 1, -5), (4, 9, 0))]:
                    a, b = i3((0, 0, 0)), i3((0, 0, 0))
                    x, y = i3((0, 0, 0)), i3((0, 0, 0))
                    assert R.cxp_ref_coarse_box(hr, i3(flo), i3(fhi), a, b) == 0
                    assert D.cxp_coarse_box(hd, i3(flo), i3(fhi), x, y) == 0
                    assert (tuple(a), tuple(b)) == (tuple(x), tuple(y))
    assert _ref_lookup("natural", 3, (2, 0, 0))[0] == _dev_lookup("natural", 3, (2, 0, 0))[0]


def test_argument_errors_match_the_device_library():
    """Each bad call returns the device library's status (no data is read:
    every case fails its checks first)."""
    R, D = oracle.lib(), _lib.lib()
    hr = _ref_lookup("natural", 3, (0, 0, 0))[1]
    hd = _dev_lookup("natural", 3, (0, 0, 0))[1]
    crse = _fab((-4, -4, -4), (20, 20, 20))
    fine = _fab((0, 0, 0), (15, 15, 15))
    small = _fab((0, 0, 0), (9, 9, 9))
    big_c = _fab((-4, -4, -4), (20, 20, 20), 17)
    big_f = _fab((0, 0, 0), (15, 15, 15), 17)
    cases = [
        dict(rlo=(0, 0, 0), rhi=(16, 15, 15)),
        dict(rlo=(0, 0, 0), rhi=(15, 15, 15), crse=small),
        dict(rlo=(0, 0, 0), rhi=(15, 15, 15), ncomp=2),
        dict(rlo=(0, 0, 0), rhi=(15, 15, 15), ncomp=17, per_group=1, crse=big_c, fine=big_f),
        dict(rlo=(3, 0, 0), rhi=(2, 15, 15)),
        dict(rlo=(0, 0, 0), rhi=(15, 15, 15), fine_comp=-1),
    ]
    seen = set()
    for c in cases:
        args = lambda h: (h, ctypes.byref(c.get("crse", crse)), 0,
                          ctypes.byref(c.get("fine", fine)), c.get("fine_comp", 0),
                          c.get("ncomp", 1), i3(c["rlo"]), i3(c["rhi"]),
                          c.get("per_group", 0), None)
        sr = R.cxp_ref_prolongate(*args(hr))
        sd = D.cxp_prolongate(*args(hd))
        assert sr == sd and sr < 0, (c, sr, sd)
        seen.add(sr)
    assert seen >= {-3, -4, -5, -6, -9}
    # restriction
    it = i3((0, 1, 0))
    for rlo, rhi, fc, want in [((0, 0, 0), (9, 7, 7), 0, -5), ((0, 0, 0), (21, 7, 7), 0, -4),
                               ((0, 0, 0), (7, 7, 7), 1, -9)]:
        f16 = _fab((0, 0, 0), (16, 15, 15))
        c8 = _fab((0, 0, 0), (20, 7, 7))
        sr = R.cxp_ref_restrict(it, ctypes.byref(f16), fc, ctypes.byref(c8), 0, 1,
                                i3(rlo), i3(rhi), None)
        sd = D.cxp_restrict(it, ctypes.byref(f16), fc, ctypes.byref(c8), 0, 1,
                            i3(rlo), i3(rhi), None)
        assert sr == sd == want, (rlo, rhi, sr, sd)


def test_results_are_the_oracles():
    """cxp_ref_prolongate / cxp_ref_restrict on host arrays equal the
    oracle's own calls bit for bit, per-var and per-group."""
    rng = np.random.default_rng(5)
    for fam, order, it, pg in [("natural", 5, (0, 0, 0), 0), ("eno", 3, (1, 1, 1), 1),
                               ("poly-cons3lfb", 3, (1, 0, 1), 0)]:
        hr = _ref_lookup(fam, order, it)[1]
        op = oracle.lookup(fam, order, it)
        rlo, rhi = (-3, 1, 2), (12, 14, 9)
        clo, chi = oracle.coarse_box(op, rlo, rhi)
        carr = rng.uniform(-1, 1, (2,) + tuple(chi[d] - clo[d] + 1 for d in (2, 1, 0)))
        farr = np.full((2,) + tuple(rhi[d] - rlo[d] + 1 for d in (2, 1, 0)), 7.0)
        want = oracle.Fab(rlo, rhi, 2)
        oracle.prolongate(op, oracle.Fab(clo, chi, 2, carr.copy()), want, rlo, rhi,
                          per_group=bool(pg))
        st = oracle.lib().cxp_ref_prolongate(
            hr, ctypes.byref(_fab(clo, chi, 2, carr)), 0, ctypes.byref(_fab(rlo, rhi, 2, farr)),
            0, 2, i3(rlo), i3(rhi), pg, None)
        assert st == 0
        assert np.array_equal(farr, want.arr), fam
    it = (1, 0, 1)
    fine = rng.uniform(-1, 1, (1, 16, 16, 17))
    crse = np.zeros((1, 8, 9, 8))
    st = oracle.lib().cxp_ref_restrict(i3(it), ctypes.byref(_fab((0, 0, 0), (16, 15, 15), 1, fine)), 0,
                                       ctypes.byref(_fab((0, 0, 0), (7, 8, 7), 1, crse)), 0, 1,
                                       i3((0, 0, 0)), i3((7, 7, 7)), None)
    assert st == 0
    want = oracle.Fab((0, 0, 0), (7, 8, 7), 1)
    oracle.restrict(it, oracle.Fab((0, 0, 0), (16, 15, 15), 1, fine.copy()), want, (0, 0, 0), (7, 7, 7))
    assert np.array_equal(crse, want.arr)
