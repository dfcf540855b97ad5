"""cxp_prolongate_adjoint_noalloc: the recomputing adjoint with its branch
codes in the library's static scratch array instead of an allocation.
Checked through the C ABI against cxp_prolongate_adjoint, bit for bit, on
two calls in a row (the second smaller), inside a captured graph, and for
the size limit before any device work."""
import ctypes

import numpy as np
import pytest
import torch

from carpetx_amd import _lib
from carpetx_amd._lib import i3
from carpetx_amd.batched import Graph
from carpetx_amd.interp import Box, FArrayBox, get_interpolator

SCRATCH_BYTES = 16 << 20   # CXP_NOALLOC_SCRATCH_BYTES
ERR_SCRATCH = -12


def _fake_fab(lo, hi, ncomp=1):
    f = _lib.CxpFab()
    f.data = ctypes.c_void_p(0x1000)
    f.lo, f.hi = i3(lo), i3(hi)
    f.comp_stride = (hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1)
    f.ncomp = ncomp
    return f


def test_noalloc_rejects_calls_beyond_the_scratch_before_device_work():
    """Branch codes of more than CXP_NOALLOC_SCRATCH_BYTES: CXP_ERR_SCRATCH,
    returned before any HIP call (fake device pointers, no GPU needed)."""
    L = _lib.lib()
    h = get_interpolator("eno", 3, (1, 1, 1)).handle
    # per-var, 3 components of 192^3: 3 x 7.08 MB > 16 MiB
    rlo, rhi = (0, 0, 0), (191, 191, 191)
    crse = _fake_fab((-3, -3, -3), (99, 99, 99), 3)
    fine = _fake_fab(rlo, rhi, 3)
    st = L.cxp_prolongate_adjoint_noalloc(h, ctypes.byref(crse), ctypes.byref(crse), 0,
                                          ctypes.byref(fine), 0, 3, i3(rlo), i3(rhi), 0, None)
    assert st == ERR_SCRATCH
    assert 3 * 192 ** 3 > SCRATCH_BYTES >= 2 * 192 ** 3
    # still the ordinary argument errors first
    st = L.cxp_prolongate_adjoint_noalloc(h, None, ctypes.byref(crse), 0,
                                          ctypes.byref(fine), 0, 3, i3(rlo), i3(rhi), 0, None)
    assert st == -3


def _case(fam, order, it, region, ncomp, seed):
    P = get_interpolator(fam, order, it)
    cb = P.CoarseBox(region)
    rng = np.random.default_rng(seed)
    crse = FArrayBox(cb, ncomp, tensor=torch.from_numpy(
        rng.uniform(-1, 1, (ncomp,) + tuple(reversed(cb.size)))).cuda())
    fbar = FArrayBox(region, ncomp, tensor=torch.from_numpy(
        rng.uniform(-1, 1, (ncomp,) + tuple(reversed(region.size)))).cuda())
    return P, crse, fbar


def _adjoint(fn, P, crse, fbar, region, ncomp, per_group):
    cbar = FArrayBox(crse.box, ncomp)
    ci, cb, fb = crse.c(), cbar.c(), fbar.c()
    st = fn(P.handle, ctypes.byref(ci), ctypes.byref(cb), 0, ctypes.byref(fb), 0, ncomp,
            i3(region.lo), i3(region.hi), int(per_group),
            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert st == 0, st
    return cbar


@pytest.mark.gpu
@pytest.mark.parametrize("fam,order,it", [
    ("eno", 3, (1, 1, 1)), ("poly-cons3lfb", 5, (1, 0, 1)), ("poly-eno3lfb", 3, (0, 1, 1)),
    ("natural", 5, (0, 0, 0)), ("minmod", 3, (1, 1, 1)), ("minmod", 5, (0, 1, 1)),
])
def test_noalloc_equals_the_allocating_adjoint(fam, order, it):
    L = _lib.lib()
    calls = [(Box((-3, 1, -2), (40, 36, 30)), 2, False, 5),    # first call
             (Box((2, 0, 1), (21, 17, 16)), 1, True, 6)]       # second, smaller
    for region, ncomp, pg, seed in calls:
        P, crse, fbar = _case(fam, order, it, region, ncomp, seed)
        want = _adjoint(L.cxp_prolongate_adjoint, P, crse, fbar, region, ncomp, pg)
        got = _adjoint(L.cxp_prolongate_adjoint_noalloc, P, crse, fbar, region, ncomp, pg)
        torch.cuda.synchronize()
        assert torch.equal(got.t, want.t), (fam, region, pg)
        assert want.t.abs().max() > 0


@pytest.mark.gpu
def test_noalloc_adjoint_replays_in_a_graph():
    """Captured once, replayed three times with crse_bar cleared eagerly in
    between: every replay equals the eager allocating adjoint."""
    L = _lib.lib()
    region = Box((0, -2, 1), (29, 25, 20))
    P, crse, fbar = _case("eno", 3, (1, 1, 1), region, 1, 9)
    want = _adjoint(L.cxp_prolongate_adjoint, P, crse, fbar, region, 1, False)
    cbar = FArrayBox(crse.box, 1)
    ci, cb, fb = crse.c(), cbar.c(), fbar.c()
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        def body():
            assert L.cxp_prolongate_adjoint_noalloc(
                P.handle, ctypes.byref(ci), ctypes.byref(cb), 0, ctypes.byref(fb), 0, 1,
                i3(region.lo), i3(region.hi), 0, ctypes.c_void_p(s.cuda_stream)) == 0
        g = Graph(body, s)
    for rep in range(3):
        cbar.t.zero_()
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            g.launch(s)
        torch.cuda.synchronize()
        assert torch.equal(cbar.t, want.t), rep
    g.close()
