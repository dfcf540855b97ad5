"""The CCC ENO kernels (carpetx_amd/csrc/eno_ccc.hip: per-parent shift
search from the coarse Mall field, eight children per thread; adjoint
gathered from the branch codes).  Against the oracle bit for bit on ragged
boxes; at 96^3 -> 192^3 against the generic per-point kernels (the
per-group path with one component, whose search is the same), which the
small-box tests pin to the oracle."""
import numpy as np
import pytest
import torch

import oracle
from carpetx_amd.interp import Box, FArrayBox, get_interpolator

pytestmark = pytest.mark.gpu
IT = (1, 1, 1)


def _case(region, ncomp, seed):
    P = get_interpolator("eno", 3, IT)
    cb = P.CoarseBox(region)
    rng = np.random.default_rng(seed)
    u = rng.uniform(-1, 1, (ncomp,) + tuple(reversed(cb.size)))
    u[:, ::4] = np.abs(u[:, ::4])       # kinks and ties
    u[:, :, ::3] = 0.5
    crse = FArrayBox(cb, ncomp, tensor=torch.from_numpy(u).cuda())
    return P, cb, u, crse


@pytest.mark.parametrize("lo,hi", [((-5, 2, -3), (40, 37, 12)), ((3, -7, 1), (4, -6, 2)),
                                   ((0, 0, 0), (31, 15, 7))])
def test_forward_and_record_equal_the_oracle(lo, hi):
    region = Box(lo, hi)
    P, cb, u, crse = _case(region, 2, 3)
    oop = oracle.lookup("eno", 3, IT)
    fh = oracle.Fab(region.lo, region.hi, 2)
    oracle.prolongate(oop, oracle.Fab(cb.lo, cb.hi, 2, u), fh, region.lo, region.hi, ncomp=2)
    fine = FArrayBox(region, 2)
    P.interp(crse, 0, fine, 0, 2, region)
    nb = P.branch_bytes(region, 2)
    br = torch.full((nb,), 255, dtype=torch.uint8, device="cuda")
    f2 = FArrayBox(region, 2)
    P.interp_record(crse, 0, f2, 0, 2, region, br)
    torch.cuda.synchronize()
    assert np.array_equal(fine.t.cpu().numpy(), fh.arr)
    assert torch.equal(fine.t, f2.t)
    assert int(br.max()) <= 26


def test_at_size_against_the_generic_kernels():
    """96^3 -> 192^3, one component: values, codes and both adjoints of the
    CCC kernels against the generic per-point kernels (per_group = 1)."""
    region = Box((0, 0, 0), (191, 191, 191))
    P, cb, u, crse = _case(region, 1, 9)
    f_fast, f_gen = FArrayBox(region, 1), FArrayBox(region, 1)
    nb = P.branch_bytes(region, 1)
    b_fast = torch.full((nb,), 255, dtype=torch.uint8, device="cuda")
    b_gen = torch.full((nb,), 255, dtype=torch.uint8, device="cuda")
    P.interp_record(crse, 0, f_fast, 0, 1, region, b_fast)
    P.interp_record(crse, 0, f_gen, 0, 1, region, b_gen, per_group=1)
    torch.cuda.synchronize()
    assert torch.equal(f_fast.t, f_gen.t)
    assert torch.equal(b_fast, b_gen)
    g = torch.Generator(device="cuda").manual_seed(5)
    vb = FArrayBox(region, 1, tensor=torch.rand((1,) + tuple(reversed(region.size)),
                                                dtype=torch.float64, device="cuda", generator=g))
    a_re, a_rec, a_gen = FArrayBox(cb, 1), FArrayBox(cb, 1), FArrayBox(cb, 1)
    P.interp_adjoint(a_re, 0, vb, 0, 1, region, crse_in=crse)
    P.interp_adjoint_recorded(b_fast, a_rec, 0, vb, 0, 1, region)
    P.interp_adjoint(a_gen, 0, vb, 0, 1, region, crse_in=crse, per_group=1)
    torch.cuda.synchronize()
    assert torch.equal(a_re.t, a_rec.t)
    err = (a_re.t - a_gen.t).abs().max().item()
    assert err <= 1e-13 * a_gen.t.abs().max().item()
    # <P u, v> == <u, P^T v>: P is linear in u for its fixed shifts, and
    # P(u) = J(u) u since the search is invariant under u -> a u (a > 0)
    lhs = torch.sum(f_fast.t * vb.t).item()
    rhs = torch.sum(crse.t * a_re.t).item()
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
