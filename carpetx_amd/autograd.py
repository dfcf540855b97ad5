"""torch.autograd wrappers around the rf=2 transfer operators.

Forward and backward both run in libcxp.so: the forward is cxp_prolongate /
cxp_restrict, the backward is the hand-written adjoint kernel
(cxp_prolongate_adjoint / cxp_restrict_adjoint).  Tensors are FABs laid out
(ncomp, nz, ny, nx) over an index box, x fastest.

    fine = prolongate(P, crse, crse_box, fine_region)
    crse = restrict(fine, fine_box, crse_region, indextype)
"""
import torch

from .interp import Box, FArrayBox, restrict as _restrict, \
    restrict_adjoint as _restrict_adjoint


def _fab(t, box):
    return FArrayBox(box, t.shape[0], tensor=t)


class _Prolongate(torch.autograd.Function):
    """Operators with branches (ENO shifts, linear fallback) record them in
    the forward, so the backward is the adjoint's gather alone."""

    @staticmethod
    def forward(ctx, crse, P, crse_box, fine_region):
        crse = crse.contiguous()
        ncomp = crse.shape[0]
        fine = torch.empty((ncomp,) + tuple(reversed(fine_region.size)),
                           dtype=torch.float64, device=crse.device)
        nb = P.branch_bytes(fine_region, ncomp)
        branches = None
        if nb > 0:
            branches = torch.empty(nb, dtype=torch.uint8, device=crse.device)
            P.interp_record(_fab(crse, crse_box), 0, _fab(fine, fine_region), 0,
                            ncomp, fine_region, branches)
        else:
            P.interp(_fab(crse, crse_box), 0, _fab(fine, fine_region), 0, ncomp,
                     fine_region)
        ctx.P, ctx.crse_box, ctx.fine_region = P, crse_box, fine_region
        ctx.shape = crse.shape
        ctx.branches = branches
        return fine

    @staticmethod
    def backward(ctx, gfine):
        gfine = gfine.contiguous()
        gcrse = torch.zeros(ctx.shape, dtype=torch.float64, device=gfine.device)
        ctx.P.interp_adjoint_recorded(ctx.branches, _fab(gcrse, ctx.crse_box), 0,
                                      _fab(gfine, ctx.fine_region), 0, ctx.shape[0],
                                      ctx.fine_region)
        return gcrse, None, None, None


def prolongate(P, crse, crse_box, fine_region):
    """Differentiable prolongation of ``crse`` (covering ``crse_box``) onto
    ``fine_region``; returns a tensor shaped like ``fine_region``."""
    return _Prolongate.apply(crse, P, crse_box, fine_region)


class _Restrict(torch.autograd.Function):
    @staticmethod
    def forward(ctx, fine, fine_box, crse_region, indextype):
        fine = fine.contiguous()
        ncomp = fine.shape[0]
        crse = torch.empty((ncomp,) + tuple(reversed(crse_region.size)),
                           dtype=torch.float64, device=fine.device)
        _restrict(_fab(fine, fine_box), _fab(crse, crse_region), indextype,
                  ncomp=ncomp, crse_region=crse_region)
        ctx.fine_box, ctx.crse_region, ctx.indextype = fine_box, crse_region, indextype
        ctx.shape = fine.shape
        return crse

    @staticmethod
    def backward(ctx, gcrse):
        gcrse = gcrse.contiguous()
        gfine = torch.zeros(ctx.shape, dtype=torch.float64, device=gcrse.device)
        _restrict_adjoint(_fab(gcrse, ctx.crse_region), _fab(gfine, ctx.fine_box),
                          ctx.indextype, gcrse.shape[0], ctx.crse_region)
        return gfine, None, None, None


def restrict(fine, fine_box, crse_region, indextype):
    """Differentiable restriction of ``fine`` (covering ``fine_box``) onto
    ``crse_region``."""
    return _Restrict.apply(fine, fine_box, crse_region, tuple(indextype))


def region_view(t, box, region):
    """Slice of a FAB tensor ``t`` over ``box`` restricted to ``region``."""
    sl = tuple(slice(region.lo[d] - box.lo[d], region.hi[d] - box.lo[d] + 1)
               for d in (2, 1, 0))
    return t[(slice(None),) + sl]


class _FillGhosts(torch.autograd.Function):
    """The six slabs of a ghost shell in one node.  The forward writes each
    slab's prolongation into a copy of ``fine`` (recording the branches of
    operators that have them); the backward adds every slab's adjoint into
    one coarse gradient, slab after slab on the stream, and passes the fine
    gradient through where the fill did not write."""

    @staticmethod
    def forward(ctx, crse, fine, P, crse_box, fine_box, interior):
        from .workloads import shell_slabs
        crse = crse.contiguous()
        out = torch.empty_like(fine, memory_format=torch.contiguous_format)
        out.copy_(fine)
        ncomp = crse.shape[0]
        slabs = shell_slabs(fine_box, interior)
        cf, of = _fab(crse, crse_box), _fab(out, fine_box)
        branches = [None] * len(slabs)
        for i, slab in enumerate(slabs):
            nb = P.branch_bytes(slab, ncomp)
            if nb > 0:
                branches[i] = torch.empty(nb, dtype=torch.uint8, device=crse.device)
                P.interp_record(cf, 0, of, 0, ncomp, slab, branches[i])
        # operators without branches: every slab in one call
        linear = [s for s, br in zip(slabs, branches) if br is None]
        if linear:
            P.interp_regions(cf, 0, of, 0, ncomp, linear)
        ctx.P, ctx.crse_box, ctx.fine_box = P, crse_box, fine_box
        ctx.slabs, ctx.branches, ctx.cshape = slabs, branches, crse.shape
        ctx.interior = interior
        return out

    @staticmethod
    def backward(ctx, gout):
        gout = gout.contiguous()
        ncomp = ctx.cshape[0]
        gcrse = torch.zeros(ctx.cshape, dtype=torch.float64, device=gout.device)
        gc, gf = _fab(gcrse, ctx.crse_box), _fab(gout, ctx.fine_box)
        if all(br is None for br in ctx.branches):
            # slabs whose coarse windows are apart share a launch
            ctx.P.interp_adjoint_regions(gc, 0, gf, 0, ncomp, ctx.slabs)
        else:
            for slab, br in zip(ctx.slabs, ctx.branches):
                ctx.P.interp_adjoint_recorded(br, gc, 0, gf, 0, ncomp, slab)
        # the fine gradient passes where the fill did not write: one select
        # against the interior mask instead of a copy and six slab fills
        gfine = torch.where(_interior_mask(ctx.fine_box, ctx.interior, gout.device),
                            gout, 0.0)
        return gcrse, gfine, None, None, None, None


_MASKS = {}


def _interior_mask(box, interior, device):
    """Bool tensor (1, nz, ny, nx) over box: True on interior (the points a
    ghost fill leaves alone), made once per (box, interior, device)."""
    key = (box.lo, box.hi, interior.lo, interior.hi, str(device))
    m = _MASKS.get(key)
    if m is None:
        m = torch.zeros((1,) + tuple(reversed(box.size)), dtype=torch.bool, device=device)
        m[(slice(None),) + _slices(interior, box)] = True
        _MASKS[key] = m
    return m


def _slices(region, box):
    return tuple(slice(region.lo[d] - box.lo[d], region.hi[d] - box.lo[d] + 1)
                 for d in (2, 1, 0))


def fill_ghosts(P, crse, crse_box, fine, fine_box, interior):
    """Differentiable ghost-shell fill: returns ``fine`` with every point of
    ``fine_box`` outside ``interior`` replaced by the prolongation of
    ``crse``.  The shell is split into the six slabs of
    cxp_worklist_append_shell."""
    return _FillGhosts.apply(crse, fine, P, crse_box, fine_box, interior)


__all__ = ["prolongate", "restrict", "fill_ghosts", "region_view", "Box"]
