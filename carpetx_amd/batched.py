"""Batched prolongation: a device-resident work list and hipGraph capture.

Replaces the reference's loop of FillPatchInterp -> interp() calls over
fine patches (CarpetX/src/fillpatch.cxx:116-119,179-180,213-216) by one
persistent kernel over every (patch, component, tile) of a level.
"""
import ctypes
import struct

import torch

from . import _lib
from ._lib import check, i3
from .interp import _stream


# CarpetX's FP64 poison pattern (CarpetX/src/valid.hxx:249-256), a NaN.
POISON_BITS = 0xfff80000deadbeef
POISON = struct.unpack("<d", POISON_BITS.to_bytes(8, "little"))[0]


class Level:
    """A refinement level for ``WorkList.append_level``: FArrayBoxes (valid
    region plus ghosts), the valid Box of each, and the level's domain (the
    index range of its valid points, in the FABs' index type)."""

    def __init__(self, fabs, valid, domain):
        if len(fabs) != len(valid):
            raise ValueError("one valid box per FAB")
        self.fabs, self.valid, self.domain = list(fabs), list(valid), domain

    def c(self):
        n = len(self.fabs)
        arr = (_lib.CxpFab * max(n, 1))(*[f.c() for f in self.fabs])
        val = (ctypes.c_int * max(6 * n, 1))(
            *[x for v in self.valid for x in tuple(v.lo) + tuple(v.hi)])
        lv = _lib.CxpLevel()
        lv.nfabs = n
        lv.fabs = ctypes.cast(arr, ctypes.POINTER(_lib.CxpFab))
        lv.valid = ctypes.cast(val, ctypes.POINTER(ctypes.c_int))
        lv.domain_lo = i3(self.domain.lo)
        lv.domain_hi = i3(self.domain.hi)
        return lv, (arr, val)


# Symmetry and boundary condition types (CXP_BC_* of the C interface):
# CarpetX's boundary_t values, its reflection symmetry and periodicity
# (CarpetX/src/boundaries.hxx, boundaries_impl.hxx:150-188).
BC_TYPES = {"none": 0, "dirichlet": 1, "linear_extrapolation": 2,
            "neumann": 3, "robin": 4, "reflection": 5, "periodic": 6}


class BoundaryConditions:
    """The coarse level's conditions for ``WorkList.append_level``.

    ``faces[f][d]`` names the condition of face f (0 lower, 1 upper) in
    direction d, a key of BC_TYPES.  ``dirichlet_values`` and
    ``robin_values`` hold one value per appended component and
    ``parities`` three (+1 or -1, one per direction) per component, as the
    reference's GroupData does; the Robin condition's coordinates are
    x = xmin + i dx (BoundaryCondition's xmin and dx, boundaries.cxx:34-56).
    """

    def __init__(self, faces, dirichlet_values=None, robin_values=None,
                 parities=None, xmin=(0.0, 0.0, 0.0), dx=(1.0, 1.0, 1.0)):
        self.faces = [[BC_TYPES[faces[f][d]] for d in range(3)] for f in range(2)]
        self.dirichlet_values = dirichlet_values
        self.robin_values = robin_values
        self.parities = parities
        self.xmin, self.dx = tuple(xmin), tuple(dx)

    def c(self):
        b = _lib.CxpBc()
        for f in range(2):
            for d in range(3):
                b.type[f][d] = self.faces[f][d]
        keep = []

        def arr(ct, vals):
            if vals is None:
                return None
            a = (ct * len(vals))(*vals)
            keep.append(a)
            return ctypes.cast(a, ctypes.POINTER(ct))

        b.dirichlet_values = arr(ctypes.c_double, self.dirichlet_values)
        b.robin_values = arr(ctypes.c_double, self.robin_values)
        b.parities = arr(ctypes.c_int, None if self.parities is None else
                         [int(x) for p in self.parities for x in p])
        b.xmin = (ctypes.c_double * 3)(*self.xmin)
        b.dx = (ctypes.c_double * 3)(*self.dx)
        return b, keep


class WorkList:
    def __init__(self):
        self.h = ctypes.c_void_p()
        check("cxp_worklist_create", _lib.lib().cxp_worklist_create(ctypes.byref(self.h)))
        self._keep = []  # FArrayBoxes referenced by the list stay alive

    def close(self):
        if self.h:
            _lib.lib().cxp_worklist_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def append(self, op, crse, crse_comp, fine, fine_comp, ncomp, region):
        c, f = crse.c(), fine.c()
        check("cxp_worklist_append", _lib.lib().cxp_worklist_append(
            self.h, op.handle, ctypes.byref(c), crse_comp, ctypes.byref(f),
            fine_comp, ncomp, i3(region.lo), i3(region.hi)))
        self._keep += [crse, fine]

    def append_shell(self, op, crse, crse_comp, fine, fine_comp, ncomp, interior):
        """Ghost shell of ``fine``: its box minus ``interior``, as 6 slabs."""
        c, f = crse.c(), fine.c()
        check("cxp_worklist_append_shell", _lib.lib().cxp_worklist_append_shell(
            self.h, op.handle, ctypes.byref(c), crse_comp, ctypes.byref(f),
            fine_comp, ncomp, i3(interior.lo), i3(interior.hi)))
        self._keep += [crse, fine]

    def append_level(self, op, crse, crse_comp, fine, fine_comp, ncomp, bc=None):
        """Ghost fill of the fine Level from the coarse Level: every fine FAB
        point in no fine valid box (FillPatch_ProlongateGhosts,
        CarpetX/src/fillpatch.cxx:42-138); coarse patches that no single
        coarse FAB covers are gathered on the device at each run.  With
        ``bc`` (BoundaryConditions), coarse points outside the coarse domain
        take the boundary values computed on the gathered patch, as the
        reference applies them to its coarse patch (fillpatch.cxx:107);
        without it they come from the coarse FABs' ghost points."""
        c, keep_c = crse.c()
        f, keep_f = fine.c()
        if bc is None:
            check("cxp_worklist_append_level", _lib.lib().cxp_worklist_append_level(
                self.h, op.handle, ctypes.byref(c), crse_comp, ctypes.byref(f),
                fine_comp, ncomp))
        else:
            b, keep_b = bc.c()
            check("cxp_worklist_append_level_bc",
                  _lib.lib().cxp_worklist_append_level_bc(
                      self.h, op.handle, ctypes.byref(c), crse_comp,
                      ctypes.byref(f), fine_comp, ncomp, ctypes.byref(b)))
        self._keep += list(crse.fabs) + list(fine.fabs)

    def append_new_level(self, op, crse, crse_comp, fine, fine_comp, ncomp, bc=None):
        """Fill of a new fine Level (FillPatch_NewLevel, CarpetX/src/
        fillpatch.cxx:140-181): every fine FAB's whole box within the fine
        domain, prolongated from the coarse Level; ``bc`` as for
        append_level."""
        c, keep_c = crse.c()
        f, keep_f = fine.c()
        b, keep_b = (None, None) if bc is None else bc.c()
        check("cxp_worklist_append_new_level", _lib.lib().cxp_worklist_append_new_level(
            self.h, op.handle, ctypes.byref(c), crse_comp, ctypes.byref(f),
            fine_comp, ncomp, None if b is None else ctypes.byref(b)))
        self._keep += list(crse.fabs) + list(fine.fabs)

    def append_remake_level(self, op, crse, crse_comp, old_fine, fine, fine_comp,
                            ncomp, bc=None):
        """Fill of a regridded fine Level (FillPatch_RemakeLevel,
        fillpatch.cxx:183-227): the old fine Level's valid data where it
        exists, prolongation from the coarse Level elsewhere within the
        fine domain grown by the ghost width."""
        c, keep_c = crse.c()
        o, keep_o = old_fine.c()
        f, keep_f = fine.c()
        b, keep_b = (None, None) if bc is None else bc.c()
        check("cxp_worklist_append_remake_level",
              _lib.lib().cxp_worklist_append_remake_level(
                  self.h, op.handle, ctypes.byref(c), crse_comp, ctypes.byref(o),
                  ctypes.byref(f), fine_comp, ncomp,
                  None if b is None else ctypes.byref(b)))
        self._keep += list(crse.fabs) + list(old_fine.fabs) + list(fine.fabs)

    def append_boundaries(self, level, indextype, comp, ncomp, bc):
        """The Level's own boundary conditions, in place, after the list's
        fills (apply_boundary_conditions, CarpetX/src/boundaries_impl.hxx:
        150-437)."""
        lv, keep = level.c()
        b, keep_b = bc.c()
        check("cxp_worklist_append_boundaries", _lib.lib().cxp_worklist_append_boundaries(
            self.h, ctypes.byref(lv), i3(indextype), comp, ncomp, ctypes.byref(b)))
        self._keep += list(level.fabs)

    def set_per_group(self, per_group):
        """The reference's prolongate_per_group parameter for this list's
        ENO and fallback patches (takes effect at the next upload)."""
        check("cxp_worklist_set_per_group",
              _lib.lib().cxp_worklist_set_per_group(self.h, 1 if per_group else 0))

    def upload(self, stream=None):
        check("cxp_worklist_upload", _lib.lib().cxp_worklist_upload(self.h, _stream(stream)))

    def stats(self):
        v = [ctypes.c_int64() for _ in range(4)]
        check("cxp_worklist_stats", _lib.lib().cxp_worklist_stats(
            self.h, *[ctypes.byref(x) for x in v]))
        return dict(patches=v[0].value, tiles=v[1].value, fine_points=v[2].value,
                    algo_bytes=v[3].value)

    def run(self, stream=None):
        check("cxp_prolongate_batched", _lib.lib().cxp_prolongate_batched(
            self.h, _stream(stream)))

    def run_per_patch(self, stream=None):
        check("cxp_prolongate_per_patch", _lib.lib().cxp_prolongate_per_patch(
            self.h, _stream(stream)))

    def poison(self, value=POISON, stream=None):
        """Write ``value`` into every fine point of every patch (requires an
        uploaded list)."""
        check("cxp_worklist_poison", _lib.lib().cxp_worklist_poison(
            self.h, float(value), _stream(stream)))

    def autotune(self, reps=3, stream=None):
        check("cxp_autotune", _lib.lib().cxp_autotune(self.h, reps, _stream(stream)))


class Graph:
    """A captured fill sequence (hipGraph), replayed with ``launch``."""

    def __init__(self, fn, stream):
        self.h = ctypes.c_void_p()
        s = _stream(stream)
        check("cxp_graph_capture_begin", _lib.lib().cxp_graph_capture_begin(s))
        try:
            fn()
        except Exception:
            g = ctypes.c_void_p()
            _lib.lib().cxp_graph_capture_end(s, ctypes.byref(g))
            if g:
                _lib.lib().cxp_graph_destroy(g)
            raise
        check("cxp_graph_capture_end", _lib.lib().cxp_graph_capture_end(s, ctypes.byref(self.h)))

    def launch(self, stream=None):
        check("cxp_graph_launch", _lib.lib().cxp_graph_launch(self.h, _stream(stream)))

    def close(self):
        if self.h:
            _lib.lib().cxp_graph_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tune_save(path):
    check("cxp_tune_save", _lib.lib().cxp_tune_save(path.encode()))


def tune_load(path):
    check("cxp_tune_load", _lib.lib().cxp_tune_load(path.encode()))
