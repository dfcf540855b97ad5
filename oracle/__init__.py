"""ctypes binding of the CPU oracle (oracle/cxp_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the measured or
shipped path.  The product (carpetx_amd) never imports this package.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

VC, CC = 0, 1
POLY, HERMITE, CONS, ENO, MINMOD = 0, 1, 2, 3, 4
FB_NONE, FB_LINEAR = 0, 1

FAMILIES = ("interpolate", "conservative", "ddf", "eno", "minmod", "hermite",
            "natural", "poly-cons3lfb", "poly-eno3lfb")
FAMILY_ORDERS = {
    "interpolate": (1, 3, 5, 7), "conservative": (0, 1), "ddf": (1, 3, 5, 7),
    "eno": (3, 5), "minmod": (1, 3, 5), "hermite": (1, 3, 5),
    "natural": (1, 3, 5, 7), "poly-cons3lfb": (1, 3, 5),
    "poly-eno3lfb": (1, 3, 5),
}


class OrOp(ctypes.Structure):
    _fields_ = [("cent", ctypes.c_int * 3), ("intp", ctypes.c_int * 3),
                ("order", ctypes.c_int * 3), ("fb", ctypes.c_int)]

    def dirs(self):
        return [(self.cent[d], self.intp[d], self.order[d]) for d in range(3)]


class OrFab(ctypes.Structure):
    _fields_ = [("data", ctypes.POINTER(ctypes.c_double)),
                ("lo", ctypes.c_int * 3), ("hi", ctypes.c_int * 3),
                ("comp_stride", ctypes.c_int64)]


class OrItem(ctypes.Structure):
    _fields_ = [("op", ctypes.POINTER(OrOp)), ("crse", OrFab),
                ("crse_comp", ctypes.c_int), ("fine", OrFab),
                ("fine_comp", ctypes.c_int), ("ncomp", ctypes.c_int),
                ("rlo", ctypes.c_int * 3), ("rhi", ctypes.c_int * 3)]


_libs = {}


def lib(kind="strict"):
    """Load liboracle_{strict,fast}.so (built by oracle/Makefile)."""
    if kind not in _libs:
        path = os.path.join(_HERE, "liboracle_%s.so" % kind)
        if not os.path.exists(path):
            raise RuntimeError("oracle library %s is not built; run "
                               "`make -C oracle`" % path)
        L = ctypes.CDLL(path)
        L.or_lookup.argtypes = [ctypes.c_char_p, ctypes.c_int,
                                ctypes.POINTER(ctypes.c_int), ctypes.POINTER(OrOp)]
        L.or_required_ghosts.argtypes = [ctypes.c_int] * 3
        L.or_stencil_radius.argtypes = [ctypes.c_int] * 5 + [ctypes.POINTER(ctypes.c_int)]
        L.or_coarse_box.argtypes = [ctypes.POINTER(OrOp)] + [ctypes.POINTER(ctypes.c_int)] * 4
        L.or_interp1d.argtypes = [ctypes.c_int] * 5 + [ctypes.POINTER(ctypes.c_double), ctypes.c_ssize_t]
        L.or_interp1d.restype = ctypes.c_double
        L.or_coeff.argtypes = [ctypes.c_int] * 5
        L.or_coeff.restype = ctypes.c_double
        L.or_prolongate.argtypes = [ctypes.POINTER(OrOp), ctypes.POINTER(OrFab), ctypes.c_int,
                                    ctypes.POINTER(OrFab), ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                    ctypes.c_int]
        L.or_prolongate_many.argtypes = [ctypes.c_int, ctypes.POINTER(OrItem), ctypes.c_int]
        L.or_restrict.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(OrFab), ctypes.c_int,
                                  ctypes.POINTER(OrFab), ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.or_num_threads.restype = ctypes.c_int
        L.or_set_num_threads.argtypes = [ctypes.c_int]
        # the C interface's own signatures (cxp_ref.h); cxp_fab structs are
        # laid out as in include/cxp.h
        P = ctypes.POINTER
        L.cxp_ref_op_lookup.argtypes = [ctypes.c_char_p, ctypes.c_int, P(ctypes.c_int),
                                        P(ctypes.c_void_p)]
        L.cxp_ref_coarse_box.argtypes = [ctypes.c_void_p] + [P(ctypes.c_int)] * 4
        L.cxp_ref_prolongate.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                         ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                         P(ctypes.c_int), P(ctypes.c_int), ctypes.c_int,
                                         ctypes.c_void_p]
        L.cxp_ref_restrict.argtypes = [P(ctypes.c_int), ctypes.c_void_p, ctypes.c_int,
                                       ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                       P(ctypes.c_int), P(ctypes.c_int), ctypes.c_void_p]
        _libs[kind] = L
    return _libs[kind]


def _i3(v):
    return (ctypes.c_int * 3)(*[int(x) for x in v])


class Fab:
    """Host FArrayBox: numpy array (ncomp, nz, ny, nx), x fastest."""

    def __init__(self, lo, hi, ncomp=1, data=None):
        self.lo = tuple(int(x) for x in lo)
        self.hi = tuple(int(x) for x in hi)
        self.ncomp = ncomp
        shape = (ncomp,) + tuple(self.hi[d] - self.lo[d] + 1 for d in (2, 1, 0))
        if data is None:
            data = np.zeros(shape, dtype=np.float64)
        assert data.shape == shape and data.dtype == np.float64
        self.arr = np.ascontiguousarray(data)

    @property
    def shape3(self):
        return tuple(self.hi[d] - self.lo[d] + 1 for d in range(3))

    def c(self):
        f = OrFab()
        f.data = self.arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        f.lo = _i3(self.lo)
        f.hi = _i3(self.hi)
        f.comp_stride = int(self.arr[0].size)
        return f

    def region(self, lo, hi, comp=None):
        """numpy view of [lo, hi] (inclusive), axes (comp,) z, y, x."""
        sl = tuple(slice(lo[d] - self.lo[d], hi[d] - self.lo[d] + 1) for d in (2, 1, 0))
        if comp is None:
            return self.arr[(slice(None),) + sl]
        return self.arr[(comp,) + sl]

    def coords(self, d):
        return np.arange(self.lo[d], self.hi[d] + 1)


def lookup(type_, order, indextype, kind="strict"):
    op = OrOp()
    e = lib(kind).or_lookup(type_.encode(), int(order), _i3(indextype), ctypes.byref(op))
    if e:
        raise ValueError("or_lookup(%s, %d): status %d" % (type_, order, e))
    return op


def required_ghosts(cent, intp, order):
    return lib().or_required_ghosts(cent, intp, order)


def stencil_radius(cent, intp, order, shift, off):
    r = (ctypes.c_int * 2)()
    lib().or_stencil_radius(cent, intp, order, shift, off, r)
    return (r[0], r[1])


def coarse_box(op, flo, fhi):
    clo, chi = _i3((0, 0, 0)), _i3((0, 0, 0))
    lib().or_coarse_box(ctypes.byref(op), _i3(flo), _i3(fhi), clo, chi)
    return tuple(clo), tuple(chi)


def interp1d(cent, intp, order, shift, off, values, center):
    """Apply a 1-D operator to ``values`` with coarse offset 0 at ``center``."""
    v = np.ascontiguousarray(values, dtype=np.float64)
    ptr = v.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    base = ctypes.cast(ctypes.addressof(ptr.contents) + 8 * center,
                       ctypes.POINTER(ctypes.c_double))
    return lib().or_interp1d(cent, intp, order, shift, off, base, 1)


def coeff(cent, intp, order, row, i):
    return lib().or_coeff(cent, intp, order, row, i)


def prolongate(op, crse, fine, rlo, rhi, ncomp=None, per_group=False,
               crse_comp=0, fine_comp=0, kind="strict"):
    ncomp = crse.ncomp if ncomp is None else ncomp
    c, f = crse.c(), fine.c()
    e = lib(kind).or_prolongate(ctypes.byref(op), ctypes.byref(c), crse_comp,
                                ctypes.byref(f), fine_comp, ncomp, _i3(rlo),
                                _i3(rhi), int(bool(per_group)))
    if e:
        raise ValueError("or_prolongate: status %d" % e)


def prolongate_many(items, per_group=False, kind="fast"):
    """items: list of (op, crse Fab, fine Fab, rlo, rhi, ncomp)."""
    arr = (OrItem * len(items))()
    keep = []
    for n, (op, crse, fine, rlo, rhi, ncomp) in enumerate(items):
        keep.append(op)
        arr[n].op = ctypes.pointer(op)
        arr[n].crse = crse.c()
        arr[n].crse_comp = 0
        arr[n].fine = fine.c()
        arr[n].fine_comp = 0
        arr[n].ncomp = ncomp
        arr[n].rlo = _i3(rlo)
        arr[n].rhi = _i3(rhi)
    e = lib(kind).or_prolongate_many(len(items), arr, int(bool(per_group)))
    if e:
        raise ValueError("or_prolongate_many: status %d" % e)


def restrict(indextype, fine, crse, rlo, rhi, ncomp=None, kind="strict"):
    ncomp = fine.ncomp if ncomp is None else ncomp
    f, c = fine.c(), crse.c()
    e = lib(kind).or_restrict(_i3(indextype), ctypes.byref(f), 0, ctypes.byref(c), 0,
                              ncomp, _i3(rlo), _i3(rhi))
    if e:
        raise ValueError("or_restrict: status %d" % e)


def num_threads(kind="fast"):
    return lib(kind).or_num_threads()


def set_num_threads(n, kind="fast"):
    """Threads of the later parallel regions started from this thread."""
    lib(kind).or_set_num_threads(int(n))
