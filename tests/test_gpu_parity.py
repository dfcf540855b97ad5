"""GPU parity: libcxp.so through the C ABI against the oracle on the same
seeded inputs.

Contract: relative 1e-13 against the CPU restatement (differences from FMA
reordering only).  The library and the strict oracle are both built without
FMA contraction and evaluate the reference's summation forms in the same
order, so the tests demand more: bit-identical results
(``np.array_equal``), with the 1e-13 bound reported alongside.
"""
import zlib

import numpy as np
import pytest
import torch

import oracle
from carpetx_amd.interp import Box, FArrayBox, get_interpolator

pytestmark = pytest.mark.gpu

RTOL = 1e-13  # the contract


def all_ops():
    for fam in oracle.FAMILIES:
        for order in oracle.FAMILY_ORDERS[fam]:
            for c in range(8):
                yield fam, order, ((c >> 2) & 1, (c >> 1) & 1, c & 1)


def run_case(fam, order, it, rlo, rhi, ncomp=1, per_group=False, seed=0,
             pad=(1, 2, 0), crse_comp=0, fine_comp=0, extra_comp=0):
    rng = np.random.default_rng(seed)
    op_o = oracle.lookup(fam, order, it)
    P = get_interpolator(fam, order, it)
    clo, chi = oracle.coarse_box(op_o, rlo, rhi)
    cbox = Box([clo[d] - pad[d] for d in range(3)], [chi[d] + pad[(d + 1) % 3] for d in range(3)])
    fbox = Box([rlo[d] - pad[(d + 2) % 3] for d in range(3)], [rhi[d] + pad[d] for d in range(3)])
    nc_c = ncomp + crse_comp + extra_comp
    nc_f = ncomp + fine_comp + extra_comp
    cdata = rng.uniform(-1, 1, size=(nc_c,) + tuple(reversed(cbox.size)))
    crse_h = oracle.Fab(cbox.lo, cbox.hi, nc_c, cdata)
    sentinel = 1234.5
    fine_h = oracle.Fab(fbox.lo, fbox.hi, nc_f, np.full((nc_f,) + tuple(reversed(fbox.size)), sentinel))
    oracle.prolongate(op_o, crse_h, fine_h, rlo, rhi, ncomp=ncomp, per_group=per_group,
                      crse_comp=crse_comp, fine_comp=fine_comp)
    crse = FArrayBox(cbox, nc_c, tensor=torch.from_numpy(cdata).cuda())
    fine = FArrayBox(fbox, nc_f, tensor=torch.full((nc_f,) + tuple(reversed(fbox.size)), sentinel,
                                                    dtype=torch.float64, device="cuda"))
    P.interp(crse, crse_comp, fine, fine_comp, ncomp, Box(rlo, rhi), per_group=per_group)
    torch.cuda.synchronize()
    got = fine.t.cpu().numpy()
    return got, fine_h.arr


def assert_parity(got, ref, what):
    scale = max(1.0, float(np.max(np.abs(ref[np.isfinite(ref)]))))
    err = float(np.max(np.abs(got - ref)))
    assert err <= RTOL * scale, "%s: max err %g" % (what, err)
    assert np.array_equal(got, ref), "%s: not bit-identical (max err %g)" % (what, err)


@pytest.mark.parametrize("fam", oracle.FAMILIES)
def test_every_operator_small_box(fam):
    """All registered operators of a family, all 8 centrings, on a region with
    odd and negative bounds; the sentinel outside the region must survive."""
    for f, order, it in all_ops():
        if f != fam:
            continue
        pgs = (False, True)
        for pg in pgs:
            got, ref = run_case(fam, order, it, (-3, 1, -2), (8, 9, 6), ncomp=2,
                                per_group=pg, seed=zlib.crc32(repr((fam, order, it)).encode()) & 0xffff)
            assert_parity(got, ref, "%s o%d %s pg=%d" % (fam, order, it, pg))


@pytest.mark.parametrize("fam", oracle.FAMILIES)
def test_every_operator_thin_z(fam):
    """Regions 1..4 fine planes thick in z (the whole-block kernel for thin
    patches) for every operator of a family: both z parities, several 16 x 16
    tiles with odd x/y bounds, per-var and per-group."""
    for f, order, it in all_ops():
        if f != fam:
            continue
        h = zlib.crc32(repr((fam, order, it, "thin")).encode()) & 0xffff
        nz = 1 + h % 4
        z0 = -3 + (h // 4) % 2
        for pg in (False, True):
            got, ref = run_case(fam, order, it, (-5, -3, z0), (37, 20, z0 + nz - 1),
                                ncomp=2, per_group=pg, seed=h)
            assert_parity(got, ref, "%s o%d %s nz=%d z0=%d pg=%d"
                          % (fam, order, it, nz, z0, pg))


@pytest.mark.parametrize("fam,order,it", [
    ("interpolate", 3, (0, 0, 0)), ("natural", 5, (0, 0, 0)),
    ("natural", 5, (1, 1, 1)), ("conservative", 1, (1, 1, 1)),
    ("natural", 7, (1, 0, 1)), ("minmod", 3, (1, 1, 1)),
    ("ddf", 7, (0, 1, 1)), ("natural", 1, (1, 0, 0)),
])
def test_multi_tile_regions(fam, order, it):
    """Regions spanning many tiles and z-chunks, odd starts, all three
    directions longer than a tile."""
    got, ref = run_case(fam, order, it, (-7, 3, -9), (70, 40, 36), ncomp=1, seed=7)
    assert_parity(got, ref, "%s o%d %s" % (fam, order, it))


@pytest.mark.parametrize("fam,order,it", [
    ("eno", 3, (1, 1, 1)), ("poly-cons3lfb", 5, (1, 0, 1)), ("poly-eno3lfb", 3, (0, 1, 1)),
])
def test_point_kernel_ops_larger(fam, order, it):
    for pg in (False, True):
        got, ref = run_case(fam, order, it, (-5, 0, 1), (30, 21, 17), ncomp=3,
                            per_group=pg, seed=11)
        assert_parity(got, ref, "%s o%d %s pg=%d" % (fam, order, it, pg))


@pytest.mark.parametrize("fam,order,it", [
    ("eno", 5, (1, 1, 1)), ("poly-cons3lfb", 5, (1, 0, 1)), ("poly-eno3lfb", 5, (0, 1, 1)),
])
def test_point_kernel_per_group_many_components(fam, order, it):
    """Per-group mode with 12 components: the coarse block of a tile no
    longer fits in LDS, so the one-thread-per-point kernel runs."""
    got, ref = run_case(fam, order, it, (-5, 0, 1), (20, 13, 11), ncomp=12,
                        per_group=True, seed=13)
    assert_parity(got, ref, "%s o%d %s pg=1 ncomp=12" % (fam, order, it))


def test_component_offsets():
    """crse_comp / fine_comp select components inside multi-component FABs."""
    got, ref = run_case("natural", 3, (0, 1, 0), (0, 0, 0), (17, 12, 9), ncomp=2,
                        crse_comp=1, fine_comp=2, extra_comp=1, seed=3)
    assert_parity(got, ref, "component offsets")


def test_thin_slabs_every_orientation():
    """Ghost-shell-like slabs 1..3 points thick in each direction."""
    for it in [(0, 0, 0), (1, 1, 1), (0, 1, 0)]:
        for d in range(3):
            for th in (1, 2, 3):
                lo = [-3, -3, -3]
                hi = [40, 33, 29]
                hi[d] = lo[d] + th - 1
                got, ref = run_case("natural", 5, it, tuple(lo), tuple(hi), ncomp=1,
                                    seed=th + 10 * d)
                assert_parity(got, ref, "slab it=%s d=%d th=%d" % (it, d, th))


@pytest.mark.parametrize("fam,order,it,pg", [("natural", 5, (0, 0, 0), 0),
                                              ("conservative", 1, (1, 1, 1), 0),
                                              ("eno", 3, (1, 0, 1), 1),
                                              ("poly-eno3lfb", 3, (1, 1, 1), 0),
                                              ("minmod", 3, (1, 1, 0), 1)])
def test_same_call_through_the_device_and_oracle_libraries(fam, order, it, pg):
    """One argument list, two libraries: cxp_prolongate on device FABs and
    cxp_ref_prolongate (the oracle behind the same signature) on host
    copies give identical fine values and status; then cxp_restrict and
    cxp_ref_restrict on the result."""
    import ctypes
    from carpetx_amd import _lib as L_
    from carpetx_amd._lib import i3
    D, R = L_.lib(), oracle.lib()

    def fab(lo, hi, ncomp, ptr):
        f = L_.CxpFab()
        f.data = ctypes.c_void_p(ptr)
        f.lo, f.hi = i3(lo), i3(hi)
        f.comp_stride = (hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1)
        f.ncomp = ncomp
        return f
    hd, hr = ctypes.c_void_p(), ctypes.c_void_p()
    assert D.cxp_op_lookup(fam.encode(), order, i3(it), ctypes.byref(hd)) == 0
    assert R.cxp_ref_op_lookup(fam.encode(), order, i3(it), ctypes.byref(hr)) == 0
    rlo, rhi = (-2, 0, 3), (29, 21, 16)
    clo, chi = i3((0, 0, 0)), i3((0, 0, 0))
    assert D.cxp_coarse_box(hd, i3(rlo), i3(rhi), clo, chi) == 0
    clo, chi = tuple(clo), tuple(chi)
    rng = np.random.default_rng(order + 7 * pg)
    cshape = (3,) + tuple(chi[d] - clo[d] + 1 for d in (2, 1, 0))
    fshape = (3,) + tuple(rhi[d] - rlo[d] + 1 for d in (2, 1, 0))
    c_h = rng.uniform(-1, 1, cshape)
    f_h = np.full(fshape, 5.0)
    c_d = torch.from_numpy(c_h).cuda()
    f_d = torch.from_numpy(f_h).cuda()
    args = (1, None, 1, 2, i3(rlo), i3(rhi), pg, None)
    cd, fd = fab(clo, chi, 3, c_d.data_ptr()), fab(rlo, rhi, 3, f_d.data_ptr())
    ch, fh = fab(clo, chi, 3, c_h.ctypes.data), fab(rlo, rhi, 3, f_h.ctypes.data)
    sd = D.cxp_prolongate(hd, ctypes.byref(cd), 1, ctypes.byref(fd), 1, 2, i3(rlo), i3(rhi),
                          pg, None)
    sr = R.cxp_ref_prolongate(hr, ctypes.byref(ch), 1, ctypes.byref(fh), 1, 2, i3(rlo),
                              i3(rhi), pg, None)
    torch.cuda.synchronize()
    assert sd == sr == 0
    assert np.array_equal(f_d.cpu().numpy(), f_h), (fam, order, it, pg)
    # restriction of the prolongated values back onto a coarse region
    qlo = tuple((rlo[d] + 1) // 2 for d in range(3))
    qhi = tuple((rhi[d] - it[d]) // 2 for d in range(3))
    q_h = np.zeros(cshape)
    q_d = torch.zeros(cshape, dtype=torch.float64, device="cuda")
    sd = D.cxp_restrict(i3(it), ctypes.byref(fd), 1, ctypes.byref(fab(clo, chi, 3, q_d.data_ptr())),
                        0, 2, i3(qlo), i3(qhi), None)
    sr = R.cxp_ref_restrict(i3(it), ctypes.byref(fh), 1, ctypes.byref(fab(clo, chi, 3, q_h.ctypes.data)),
                            0, 2, i3(qlo), i3(qhi), None)
    torch.cuda.synchronize()
    assert sd == sr == 0
    np.testing.assert_allclose(q_d.cpu().numpy(), q_h, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("fam,order,it,fb", [
    ("natural", 3, (1, 1, 1), 1), ("natural", 5, (1, 0, 1), 1), ("eno", 3, (1, 1, 0), 1),
    ("poly-cons3lfb", 3, (1, 1, 1), 0), ("poly-eno3lfb", 5, (0, 1, 1), 0),
])
def test_fallback_override_matches_the_oracle(fam, order, it, fb):
    """cxp_op_lookup_fallback: the family's stencils with the linear fallback
    switched on or off, against the oracle's operator with the same
    fallback, per-var and per-group, bit for bit."""
    rng = np.random.default_rng(31)
    rlo, rhi = (-3, 1, -2), (17, 12, 9)
    op_o = oracle.lookup(fam, order, it)
    op_o.fb = fb
    P = get_interpolator(fam, order, it, fallback=fb)
    clo, chi = oracle.coarse_box(op_o, rlo, rhi)
    cbox = Box(clo, chi)
    # a step in x so that the fallback fires somewhere
    x = np.arange(clo[0], chi[0] + 1)
    base = rng.uniform(-1, 1, (2,) + tuple(reversed(cbox.size)))
    cdata = base + 3.0 * (x >= (clo[0] + chi[0]) // 2)[None, None, None, :]
    for pg in (False, True):
        crse_h = oracle.Fab(cbox.lo, cbox.hi, 2, cdata.copy())
        fine_h = oracle.Fab(rlo, rhi, 2)
        oracle.prolongate(op_o, crse_h, fine_h, rlo, rhi, ncomp=2, per_group=pg)
        crse = FArrayBox(cbox, 2, tensor=torch.from_numpy(cdata.copy()).cuda())
        fine = FArrayBox(Box(rlo, rhi), 2)
        P.interp(crse, 0, fine, 0, 2, Box(rlo, rhi), per_group=pg)
        torch.cuda.synchronize()
        assert_parity(fine.t.cpu().numpy(), fine_h.arr, "%s o%d %s fb=%d pg=%d" % (fam, order, it, fb, pg))
    # the override changes the result where the family's fallback differs
    keep = get_interpolator(fam, order, it)
    fine2 = FArrayBox(Box(rlo, rhi), 2)
    keep.interp(crse, 0, fine2, 0, 2, Box(rlo, rhi))
    f1 = FArrayBox(Box(rlo, rhi), 2)
    P.interp(crse, 0, f1, 0, 2, Box(rlo, rhi))
    torch.cuda.synchronize()
    assert not torch.equal(fine2.t, f1.t)
