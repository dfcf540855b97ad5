"""Box-in-box refinement tags on the device (libcxp.so, include/cxp_tag.h).

``BoxInBoxRegion`` holds one region's parameters as BoxInBox/param.ccl
names them (shape_N, active_N, num_levels_N, position_{x,y,z}_N,
radius_N[level] and radius_{x,y,z}_N[level]).  ``boxinbox_setup`` is
BoxInBox_Setup (BoxInBox/src/boxinbox.cxx:102-174) for one box of a level:
it sets regrid_error to 1 in the cells a region refines and 0 elsewhere.
``error_est`` is the threshold test of CactusAmrCore::ErrorEst
(CarpetX/src/driver.cxx:1025-1080): tags where regrid_error >= threshold,
with their number and bounding box.  ``refined_box`` turns such a bounding
box into the region a new level covers.
"""
import ctypes

import torch

from ._lib import CxpBibRegion, P_DOUBLE, check, i3, lib
from .interp import Box, FArrayBox, _stream

MAX_REGIONS = 3
SHAPES = {"sphere": 0, "cube": 1}
TAG_CLEAR, TAG_SET = 0, 2  # amrex::TagBox::CLEAR / SET (include/cxp_tag.h)
MAX_LEVELS = 30  # max_num_levels, boxinbox.cxx:17


class BoxInBoxRegion:
    """One refined region.  radius and radius_x/y/z are per-level lists (at
    most 30 entries, like the parameter arrays); an entry of radius_x/y/z
    that is negative (or missing) defers to radius (boxinbox.cxx:132-141)."""

    def __init__(self, shape="cube", active=True, num_levels=1,
                 position=(0.0, 0.0, 0.0), radius=(), radius_x=(), radius_y=(),
                 radius_z=()):
        if shape not in SHAPES:
            raise ValueError("shape must be 'sphere' or 'cube', not %r" % (shape,))
        self.shape = shape
        self.active = bool(active)
        self.num_levels = int(num_levels)
        self.position = tuple(float(v) for v in position)
        self.radius = [float(v) for v in radius]
        self.radius_xyz = [[float(v) for v in r] for r in (radius_x, radius_y, radius_z)]
        for r in [self.radius] + self.radius_xyz:
            if len(r) > MAX_LEVELS:
                raise ValueError("at most %d levels" % MAX_LEVELS)

    def radii(self, level):
        """Per-direction radius at `level` (-1 where no parameter is set,
        the parameters' default)."""
        def entry(lst):
            return lst[level] if 0 <= level < len(lst) else -1.0
        out = []
        for d in range(3):
            rd = entry(self.radius_xyz[d])
            out.append(entry(self.radius) if rd < 0 else rd)
        return tuple(out)

    def c(self, level):
        r = CxpBibRegion()
        r.active = int(self.active)
        r.shape = SHAPES[self.shape]
        r.num_levels = self.num_levels
        r.position = (ctypes.c_double * 3)(*self.position)
        r.radius = (ctypes.c_double * 3)(*self.radii(level))
        return r


def boxinbox_setup(regions, level, x0, dx, err, region=None, stream=None):
    """regrid_error of the cells `region` (default: err's whole box) of one
    box: `level` is the index of the level being decided (one more than
    err's level), x0 the coordinate of cell 0's centre and dx the cell size
    of err's level.  err is an FArrayBox with one component."""
    if len(regions) > MAX_REGIONS:
        raise ValueError("at most %d regions" % MAX_REGIONS)
    if not isinstance(err, FArrayBox) or err.ncomp != 1:
        raise ValueError("err must be a one-component FArrayBox")
    region = err.box if region is None else region
    arr = (CxpBibRegion * max(1, len(regions)))(*[r.c(level) for r in regions])
    x0a = (ctypes.c_double * 3)(*[float(v) for v in x0])
    dxa = (ctypes.c_double * 3)(*[float(v) for v in dx])
    check("cxp_boxinbox_tag", lib().cxp_boxinbox_tag(
        arr, len(regions), int(level), ctypes.cast(x0a, P_DOUBLE), ctypes.cast(dxa, P_DOUBLE),
        ctypes.c_void_p(err.t.data_ptr()), i3(err.box.lo), i3(err.box.hi),
        i3(region.lo), i3(region.hi), _stream(stream)))


def error_est(err, threshold, region=None, tags=None, box_out=None, stream=None,
              read=True):
    """ErrorEst's tags over `region` (default: err's whole box).

    tags: optional int8 tensor of err's shape, set to TAG_SET (2) / TAG_CLEAR
    (0), amrex::TagBox::SET / CLEAR, in `region`.
    box_out: optional int32 device tensor of 7 entries; by default one is
    allocated.  Returns (count, Box of the set cells or None); reading them
    synchronises with the stream.  With read=False nothing is read back and
    None is returned (the call stays asynchronous)."""
    region = err.box if region is None else region
    if tags is not None and (tags.dtype != torch.int8 or tuple(tags.shape) != tuple(err.t.shape)
                             or not tags.is_contiguous()):
        raise ValueError("tags must be a contiguous int8 tensor shaped like err")
    if box_out is None:
        box_out = torch.empty(7, dtype=torch.int32, device=err.t.device)
    if box_out.dtype != torch.int32 or box_out.numel() < 7:
        raise ValueError("box_out must hold 7 int32")
    check("cxp_error_est", lib().cxp_error_est(
        ctypes.c_void_p(err.t.data_ptr()), i3(err.box.lo), i3(err.box.hi),
        i3(region.lo), i3(region.hi), float(threshold),
        ctypes.c_void_p(tags.data_ptr() if tags is not None else 0),
        ctypes.c_void_p(box_out.data_ptr()), _stream(stream)))
    if not read:
        return None
    if stream is not None:
        stream.synchronize()  # the copy below is ordered on the current stream only
    v = box_out[:7].tolist()
    if v[0] == 0:
        return 0, None
    return v[0], Box(v[1:4], v[4:7])


def refined_box(tag_box, buffer=1):
    """The fine cells covering a tag bounding box grown by `buffer` coarse
    cells (AMReX grows tags by its error buffer before clustering, the
    reason BoxInBox shrinks each radius by one cell, boxinbox.cxx:164-166),
    refined by 2.  For a box-shaped set of tags this is the box AMReX's
    clustering returns; other tag sets are beyond this helper."""
    g = tag_box.grow(buffer)
    return Box([2 * g.lo[d] for d in range(3)], [2 * g.hi[d] + 1 for d in range(3)])
