"""Reflux on the device (libcxp.so, cxp_reflux_* in include/cxp.h).

Mirrors CarpetX's Reflux(cctkGH, level) (CarpetX/src/schedule.cxx:2632-2716)
for one group: the fine level's flux register, made for its cell-centred
boxes with ratio 2 (driver.cxx:918-924), takes the coarse fluxes
(CrseInit, mult -1), the fine fluxes (FineAdd, mult +1) and corrects the
coarse state (Reflux, scale 1, the coarse geometry's cell volume).

``FluxRegister(crse, crse_fluxes, fine, fine_fluxes, dx)`` builds the plan
once (the register faces of the fine boxes, split by the coarse FABs that
hold their fluxes and target cells); ``reflux()`` runs it, asynchronously
on the stream and without allocating, so a graph can capture it.
"""
import ctypes

import torch

from . import _lib
from ._lib import CxpFab, check
from .interp import _stream


class FluxRegister:
    """crse, fine: carpetx_amd.batched.Level objects (cell-centred state
    FABs, valid boxes, domain); crse_fluxes[d] / fine_fluxes[d]: lists of
    FArrayBox, one per FAB of the level, face-centred in direction d.
    crse_mult, fine_mult, scale default to the reference's -1, +1, 1."""

    def __init__(self, crse, crse_fluxes, fine, fine_fluxes, dx, comp=0, ncomp=None,
                 periodic=(0, 0, 0), crse_mult=-1.0, fine_mult=1.0, scale=1.0):
        if ncomp is None:
            ncomp = crse.fabs[0].ncomp
        self.h = ctypes.c_void_p()
        cl, ck = crse.c()
        fl, fk = fine.c()
        keep = [ck, fk]

        def fab_arrays(per_dir, n):
            out = (ctypes.POINTER(CxpFab) * 3)()
            for d in range(3):
                if len(per_dir[d]) != n:
                    raise ValueError("one flux FAB per level FAB in direction %d" % d)
                arr = (CxpFab * n)(*[f.c() for f in per_dir[d]])
                keep.append(arr)
                out[d] = ctypes.cast(arr, ctypes.POINTER(CxpFab))
            return out
        cf = fab_arrays(crse_fluxes, len(crse.fabs))
        ff = fab_arrays(fine_fluxes, len(fine.fabs))
        check("cxp_reflux_plan_create", _lib.lib().cxp_reflux_plan_create(
            ctypes.byref(cl), cf, ctypes.byref(fl), ff, int(comp), int(ncomp),
            (ctypes.c_double * 3)(*dx), (ctypes.c_int * 3)(*periodic),
            float(crse_mult), float(fine_mult), float(scale), ctypes.byref(self.h)))

    def reflux(self, stream=None):
        check("cxp_reflux_run", _lib.lib().cxp_reflux_run(self.h, _stream(stream)))

    def stats(self):
        f, b = ctypes.c_int64(), ctypes.c_int64()
        check("cxp_reflux_plan_stats", _lib.lib().cxp_reflux_plan_stats(
            self.h, ctypes.byref(f), ctypes.byref(b)))
        return {"faces": f.value, "bytes": b.value}

    def close(self):
        if self.h:
            torch.cuda.synchronize()  # the plan's launches must have completed
            _lib.lib().cxp_reflux_plan_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def reflux(crse, crse_fluxes, fine, fine_fluxes, dx, **kw):
    """One Reflux call (the reference's per-step use): plan, run, free."""
    stream = kw.pop("stream", None)
    fr = FluxRegister(crse, crse_fluxes, fine, fine_fluxes, dx, **kw)
    fr.reflux(stream)
    fr.close()
