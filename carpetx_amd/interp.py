"""Host-side mirror of CarpetX's prolongation/restriction call surface.

The names and argument meanings follow the reference so that callers (and
the tests) read like the reference code:

* ``get_interpolator(prolongation_type, prolongation_order, indextype)``
  -- CarpetX/src/driver.cxx:593-642;
* ``Prolongate3DRF2.interp(crse, crse_comp, fine, fine_comp, ncomp,
  fine_region, ratio, ...)``, ``interp_face`` and ``CoarseBox`` -- the
  amrex::Interpolater overrides of CarpetX/src/prolongate_3d_rf2.hxx:76-120;
* ``restrict(fine, crse, ...)`` -- the average_down dispatch of Restrict(),
  CarpetX/src/schedule.cxx:2791-2816.

Device memory comes from torch (HIP); every computation runs in libcxp.so.
Errors that the reference turns into aborts raise ``CxpError``.
"""
import ctypes

import torch

from . import _lib
from ._lib import CxpFab, check, i3

VC, CC = 0, 1


class Box:
    """Inclusive index box [lo, hi] (amrex::Box without index type)."""

    def __init__(self, lo, hi):
        self.lo = tuple(int(x) for x in lo)
        self.hi = tuple(int(x) for x in hi)

    @property
    def size(self):
        return tuple(self.hi[d] - self.lo[d] + 1 for d in range(3))

    @property
    def numPts(self):
        n = 1
        for s in self.size:
            n *= s
        return n

    def grow(self, g):
        g = (g, g, g) if isinstance(g, int) else g
        return Box([self.lo[d] - g[d] for d in range(3)],
                   [self.hi[d] + g[d] for d in range(3)])

    def contains(self, other):
        return all(self.lo[d] <= other.lo[d] and other.hi[d] <= self.hi[d]
                   for d in range(3))

    def __eq__(self, other):
        return isinstance(other, Box) and self.lo == other.lo and self.hi == other.hi

    def __repr__(self):
        return "Box(%s, %s)" % (self.lo, self.hi)


class FArrayBox:
    """A device FAB: float64 tensor (ncomp, nz, ny, nx), x fastest."""

    def __init__(self, box, ncomp=1, device="cuda", tensor=None):
        self.box = box
        self.ncomp = ncomp
        shape = (ncomp,) + tuple(reversed(box.size))
        if tensor is None:
            tensor = torch.zeros(shape, dtype=torch.float64, device=device)
        if tuple(tensor.shape) != shape or tensor.dtype != torch.float64:
            raise ValueError("FArrayBox tensor must be float64 %s, got %s %s"
                             % (shape, tuple(tensor.shape), tensor.dtype))
        if not tensor.is_contiguous():
            raise ValueError("FArrayBox tensor must be contiguous")
        self.t = tensor

    def c(self):
        f = CxpFab()
        f.data = ctypes.c_void_p(self.t.data_ptr())
        f.lo = i3(self.box.lo)
        f.hi = i3(self.box.hi)
        f.comp_stride = self.box.numPts
        f.ncomp = self.ncomp
        return f

    def view(self, region, comp=None):
        """Tensor view of ``region`` (axes comp, z, y, x)."""
        sl = tuple(slice(region.lo[d] - self.box.lo[d],
                         region.hi[d] - self.box.lo[d] + 1) for d in (2, 1, 0))
        if comp is None:
            return self.t[(slice(None),) + sl]
        return self.t[(comp,) + sl]


def _stream(stream):
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


class Prolongate3DRF2:
    """One prolongation operator (an immutable registry entry)."""

    def __init__(self, handle, prolongation_type, prolongation_order, indextype):
        self.handle = handle
        self.prolongation_type = prolongation_type
        self.prolongation_order = prolongation_order
        self.indextype = tuple(indextype)
        cent, intp, order, ghosts = (ctypes.c_int * 3)(), (ctypes.c_int * 3)(), \
            (ctypes.c_int * 3)(), (ctypes.c_int * 3)()
        fb = ctypes.c_int()
        check("cxp_op_describe", _lib.lib().cxp_op_describe(
            handle, cent, intp, order, ghosts, ctypes.byref(fb)))
        self.centering = tuple(cent)
        self.interpolation = tuple(intp)
        self.order = tuple(order)
        self.required_ghosts = tuple(ghosts)
        self.fallback = fb.value

    def CoarseBox(self, fine, ratio=2):
        r = (ratio,) * 3 if isinstance(ratio, int) else tuple(ratio)
        if r != (2, 2, 2):
            raise _lib.CxpError("CoarseBox", -7, "refinement ratio must be 2")
        clo, chi = i3((0, 0, 0)), i3((0, 0, 0))
        check("cxp_coarse_box", _lib.lib().cxp_coarse_box(
            self.handle, i3(fine.lo), i3(fine.hi), clo, chi))
        return Box(tuple(clo), tuple(chi))

    def interp(self, crse, crse_comp, fine, fine_comp, ncomp, fine_region,
               ratio=(2, 2, 2), crse_geom=None, fine_geom=None, bcr=None,
               actual_comp=0, actual_state=0, gpu_or_cpu=None,
               per_group=False, stream=None):
        """prolongate_3d_rf2::interp (prolongate_3d_rf2_impl.hxx:2167-2188)."""
        r = (ratio,) * 3 if isinstance(ratio, int) else tuple(ratio)
        if r != (2, 2, 2):
            raise _lib.CxpError("interp", -7, "refinement ratio must be 2")
        if actual_comp != 0 or actual_state != 0:
            raise _lib.CxpError("interp", -3, "actual_comp/actual_state must be 0")
        c, f = crse.c(), fine.c()
        check("cxp_prolongate", _lib.lib().cxp_prolongate(
            self.handle, ctypes.byref(c), crse_comp, ctypes.byref(f), fine_comp,
            ncomp, i3(fine_region.lo), i3(fine_region.hi), int(bool(per_group)),
            _stream(stream)))

    def interp_regions(self, crse, crse_comp, fine, fine_comp, ncomp, regions,
                       per_group=False, stream=None):
        """interp over several regions of one FAB pair (the six slabs of a
        ghost shell, say), in one launch for separable operators
        (cxp_prolongate_regions; no reference counterpart)."""
        c, f = crse.c(), fine.c()
        n = len(regions)
        lo = (ctypes.c_int * (3 * max(n, 1)))(*[v for r in regions for v in r.lo])
        hi = (ctypes.c_int * (3 * max(n, 1)))(*[v for r in regions for v in r.hi])
        check("cxp_prolongate_regions", _lib.lib().cxp_prolongate_regions(
            self.handle, ctypes.byref(c), crse_comp, ctypes.byref(f), fine_comp,
            ncomp, n, lo, hi, int(bool(per_group)), _stream(stream)))

    def interp_adjoint_regions(self, crse_bar, crse_comp, fine_bar, fine_comp, ncomp,
                               regions, per_group=False, stream=None):
        """crse_bar += P^T fine_bar over several regions, for operators
        without branches (cxp_prolongate_adjoint_regions; no reference
        counterpart)."""
        cb, fb = crse_bar.c(), fine_bar.c()
        n = len(regions)
        lo = (ctypes.c_int * (3 * max(n, 1)))(*[v for r in regions for v in r.lo])
        hi = (ctypes.c_int * (3 * max(n, 1)))(*[v for r in regions for v in r.hi])
        check("cxp_prolongate_adjoint_regions", _lib.lib().cxp_prolongate_adjoint_regions(
            self.handle, ctypes.byref(cb), crse_comp, ctypes.byref(fb), fine_comp,
            ncomp, n, lo, hi, int(bool(per_group)), _stream(stream)))

    def interp_face(self, crse, crse_comp, fine, fine_comp, ncomp, fine_region,
                    ratio=(2, 2, 2), solve_mask=None, crse_geom=None,
                    fine_geom=None, bcr=None, bccomp=0, gpu_or_cpu=None,
                    per_group=False, stream=None):
        """interp_face ignores solve_mask and forwards (impl.hxx:2190-2208)."""
        if bccomp != 0:
            raise _lib.CxpError("interp_face", -3, "bccomp must be 0")
        self.interp(crse, crse_comp, fine, fine_comp, ncomp, fine_region, ratio,
                    per_group=per_group, stream=stream)

    def interp_adjoint(self, crse_bar, crse_comp, fine_bar, fine_comp, ncomp,
                       fine_region, crse_in=None, per_group=False, stream=None):
        """crse_bar += P^T fine_bar (no reference counterpart)."""
        cb, fb = crse_bar.c(), fine_bar.c()
        ci = ctypes.byref(crse_in.c()) if crse_in is not None else None
        check("cxp_prolongate_adjoint", _lib.lib().cxp_prolongate_adjoint(
            self.handle, ci, ctypes.byref(cb), crse_comp, ctypes.byref(fb),
            fine_comp, ncomp, i3(fine_region.lo), i3(fine_region.hi),
            int(bool(per_group)), _stream(stream)))

    def branch_bytes(self, fine_region, ncomp, per_group=False):
        """Size of the branch record of a forward over ``fine_region`` (0 for
        operators without ENO shifts or fallback)."""
        n = ctypes.c_int64()
        check("cxp_branch_bytes", _lib.lib().cxp_branch_bytes(
            self.handle, i3(fine_region.lo), i3(fine_region.hi), int(ncomp),
            int(bool(per_group)), ctypes.byref(n)))
        return n.value

    def interp_record(self, crse, crse_comp, fine, fine_comp, ncomp, fine_region,
                      branches, per_group=False, stream=None):
        """interp that also records every fine point's branch in ``branches``
        (a uint8 device tensor of branch_bytes bytes, or None when 0)."""
        c, f = crse.c(), fine.c()
        ptr = ctypes.c_void_p(branches.data_ptr()) if branches is not None else None
        check("cxp_prolongate_record", _lib.lib().cxp_prolongate_record(
            self.handle, ctypes.byref(c), crse_comp, ctypes.byref(f), fine_comp,
            ncomp, i3(fine_region.lo), i3(fine_region.hi), int(bool(per_group)),
            ptr, _stream(stream)))

    def interp_adjoint_recorded(self, branches, crse_bar, crse_comp, fine_bar,
                                fine_comp, ncomp, fine_region, per_group=False,
                                stream=None):
        """crse_bar += P^T fine_bar with the branches of interp_record."""
        cb, fb = crse_bar.c(), fine_bar.c()
        ptr = ctypes.c_void_p(branches.data_ptr()) if branches is not None else None
        check("cxp_prolongate_adjoint_recorded",
              _lib.lib().cxp_prolongate_adjoint_recorded(
                  self.handle, ptr, ctypes.byref(cb), crse_comp, ctypes.byref(fb),
                  fine_comp, ncomp, i3(fine_region.lo), i3(fine_region.hi),
                  int(bool(per_group)), _stream(stream)))

    def __repr__(self):
        return "Prolongate3DRF2(%s, %d, %s)" % (
            self.prolongation_type, self.prolongation_order, self.indextype)


FB_KEEP, FB_NONE, FB_LINEAR = -1, 0, 1


def get_interpolator(prolongation_type, prolongation_order, indextype, fallback=FB_KEEP):
    """get_interpolator (CarpetX/src/driver.cxx:593-642).  fallback other
    than FB_KEEP replaces the family's linear fallback (FB_NONE / FB_LINEAR;
    no reference counterpart, cxp_op_lookup_fallback)."""
    h = ctypes.c_void_p()
    if fallback == FB_KEEP:
        check("cxp_op_lookup", _lib.lib().cxp_op_lookup(
            prolongation_type.encode(), int(prolongation_order), i3(indextype),
            ctypes.byref(h)))
    else:
        check("cxp_op_lookup_fallback", _lib.lib().cxp_op_lookup_fallback(
            prolongation_type.encode(), int(prolongation_order), i3(indextype),
            int(fallback), ctypes.byref(h)))
    return Prolongate3DRF2(h, prolongation_type, prolongation_order, indextype)


def get_group_interpolator(tags, prolongation_type, prolongation_order, indextype):
    """get_interpolator(groupdata) (CarpetX/src/driver.cxx:558-591, 628-642):
    the group's tags string may set prolongation_type / prolongation_order;
    the driver parameters prolongation_type / prolongation_order fill in the
    rest."""
    h = ctypes.c_void_p()
    check("cxp_group_op_lookup", _lib.lib().cxp_group_op_lookup(
        tags.encode() if tags is not None else None, prolongation_type.encode(),
        int(prolongation_order), i3(indextype), ctypes.byref(h)))
    P = Prolongate3DRF2(h, prolongation_type, prolongation_order, indextype)
    P.tags = tags
    return P


def restrict(fine, crse, indextype, ncomp=None, crse_region=None, fine_comp=0,
             crse_comp=0, stream=None):
    """Restrict()'s average_down family (schedule.cxx:2791-2816)."""
    ncomp = fine.ncomp if ncomp is None else ncomp
    if crse_region is None:
        raise ValueError("crse_region is required")
    f, c = fine.c(), crse.c()
    check("cxp_restrict", _lib.lib().cxp_restrict(
        i3(indextype), ctypes.byref(f), fine_comp, ctypes.byref(c), crse_comp,
        ncomp, i3(crse_region.lo), i3(crse_region.hi), _stream(stream)))


def restrict_adjoint(crse_bar, fine_bar, indextype, ncomp, crse_region,
                     crse_comp=0, fine_comp=0, stream=None):
    """fine_bar += R^T crse_bar."""
    c, f = crse_bar.c(), fine_bar.c()
    check("cxp_restrict_adjoint", _lib.lib().cxp_restrict_adjoint(
        i3(indextype), ctypes.byref(c), crse_comp, ctypes.byref(f), fine_comp,
        ncomp, i3(crse_region.lo), i3(crse_region.hi), _stream(stream)))


def coarsened_region(fine_region, indextype):
    """Coarse points restriction can fill from a fine region: VC directions
    need the fine point 2i, CC directions both children 2i and 2i+1."""
    lo, hi = [], []
    for d in range(3):
        l, h = fine_region.lo[d], fine_region.hi[d]
        if indextype[d]:
            lo.append(-((-l) // 2))        # ceil(l / 2)
            hi.append((h - 1) // 2)
        else:
            lo.append(-((-l) // 2))
            hi.append(h // 2)
    return Box(lo, hi)
