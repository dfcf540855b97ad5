"""Reflux on the device (carpetx_amd.reflux over cxp_reflux_*) against the
oracle (oracle/reflux.py, pinned by conservation and a per-cell brute force
in test_oracle_reflux.py), bit for bit: both evaluate the register sum and
the correction in the same order."""
import numpy as np
import pytest
import torch

from carpetx_amd import workloads
from carpetx_amd.batched import Level
from carpetx_amd.interp import Box, FArrayBox
from carpetx_amd.reflux import FluxRegister
from oracle import reflux as rf
from test_oracle_reflux import coarse_fabs, make_fluxes

pytestmark = pytest.mark.gpu


def to_device(domain, crse, cflux, fine_boxes, fflux):
    cfabs = [FArrayBox(Box(lo, hi), data.shape[0], tensor=torch.from_numpy(data.copy()).cuda())
             for (lo, hi, data, vlo, vhi) in crse]
    clevel = Level(cfabs, [Box(vlo, vhi) for (_, _, _, vlo, vhi) in crse], Box(*domain))
    cfl = [[FArrayBox(Box(flo, fhi), fd.shape[0], tensor=torch.from_numpy(fd.copy()).cuda())
            for (flo, fhi, fd, _, _) in cflux[d]] for d in range(3)]
    # the fine level's FABs carry no data the register reads: one-point
    # placeholders with the valid boxes
    ffabs = [FArrayBox(Box(lo, hi), 1) for (lo, hi) in fine_boxes]
    flevel = Level(ffabs, [Box(lo, hi) for (lo, hi) in fine_boxes],
                   Box(tuple(2 * v for v in domain[0]), tuple(2 * v + 1 for v in domain[1])))
    ffl = [[FArrayBox(Box(lo, hi), fd.shape[0], tensor=torch.from_numpy(fd.copy()).cuda())
            for (lo, hi, fd) in fflux[d]] for d in range(3)]
    return clevel, cfl, flevel, ffl, cfabs


@pytest.mark.parametrize("periodic,mults", [
    ((0, 0, 0), (-1.0, 1.0, 1.0)),          # the reference's multipliers
    ((1, 0, 1), (-1.0, 1.0, 1.0)),
    ((0, 1, 0), (-0.01 / 64, 0.01 / 256, 1.0)),   # AMReX's dt * area weights
])
def test_reflux_matches_the_oracle(periodic, mults):
    rng = np.random.default_rng(20 + sum(periodic))
    ncomp = 3
    domain = ((0, 0, 0), (7, 7, 7))
    crse = coarse_fabs(domain, 4, ncomp, rng)
    # two fine boxes sharing a face, one on the low x / high z domain faces,
    # one across the coarse FABs' split at x = 4
    fine_boxes = [((4, 4, 6), (7, 11, 9)), ((8, 4, 6), (11, 11, 9)),
                  ((0, 0, 12), (3, 5, 15)), ((6, 12, 0), (9, 15, 3))]
    cflux, fflux, _ = make_fluxes(domain, crse, fine_boxes, ncomp, rng, zero_boundary=False)
    clevel, cfl, flevel, ffl, cfabs = to_device(domain, crse, cflux, fine_boxes, fflux)
    dx = (0.125, 0.25, 0.5)
    cm, fm, sc = mults
    # components 1..2 of 3
    want = [(lo, hi, data[1:3].copy(), vlo, vhi) for (lo, hi, data, vlo, vhi) in crse]
    cfl_o = [[(a, b, fd[1:3].copy(), c, e) for (a, b, fd, c, e) in cflux[d]] for d in range(3)]
    ffl_o = [[(a, b, fd[1:3].copy()) for (a, b, fd) in fflux[d]] for d in range(3)]
    rf.reflux(want, cfl_o, fine_boxes, ffl_o, dx, domain, periodic=periodic,
              crse_mult=cm, fine_mult=fm, scale=sc)
    fr = FluxRegister(clevel, cfl, flevel, ffl, dx, comp=1, ncomp=2, periodic=periodic,
                      crse_mult=cm, fine_mult=fm, scale=sc)
    fr.reflux()
    torch.cuda.synchronize()
    for fab, w, (lo, hi, data, vlo, vhi) in zip(cfabs, want, crse):
        got = fab.t.cpu().numpy()
        assert np.array_equal(got[0], data[0])          # component 0 untouched
        assert np.array_equal(got[1:3], w[2])
    st = fr.stats()
    assert st["faces"] > 0 and st["bytes"] == st["faces"] * 2 * 8 * 7
    fr.close()


def test_reflux_level64_matches_the_oracle_and_replays():
    """The 64-box level (128^3 coarse cells in 8 FABs, 64 fine boxes of
    32^3): one plan run against the oracle, and three graph replays, each
    adding the same correction again."""
    lv = workloads.level_fill_64(ncomp=1)["cons2_ccc"]
    crse_l, fine_l = lv["crse"], lv["fine"]
    rng = np.random.default_rng(5)
    dom = crse_l.domain
    domain = (dom.lo, dom.hi)
    crse = []
    for fab, v in zip(crse_l.fabs, crse_l.valid):
        crse.append((fab.box.lo, fab.box.hi, fab.t.cpu().numpy().copy(), v.lo, v.hi))
    fine_boxes = [(v.lo, v.hi) for v in fine_l.valid]
    cflux, fflux, _ = make_fluxes(domain, crse, fine_boxes, 1, rng, zero_boundary=False)
    cfl = [[FArrayBox(Box(flo, fhi), 1, tensor=torch.from_numpy(fd.copy()).cuda())
            for (flo, fhi, fd, _, _) in cflux[d]] for d in range(3)]
    ffl = [[FArrayBox(Box(lo, hi), 1, tensor=torch.from_numpy(fd.copy()).cuda())
            for (lo, hi, fd) in fflux[d]] for d in range(3)]
    dx = (1.0 / 128,) * 3
    want = [(lo, hi, data.copy(), vlo, vhi) for (lo, hi, data, vlo, vhi) in crse]
    rf.reflux(want, cflux, fine_boxes, fflux, dx, domain)
    fr = FluxRegister(crse_l, cfl, fine_l, ffl, dx)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        fr.reflux(s)
    torch.cuda.synchronize()
    for fab, w in zip(crse_l.fabs, want):
        assert np.array_equal(fab.t.cpu().numpy(), w[2])
    before = [fab.t.clone() for fab in crse_l.fabs]
    from carpetx_amd.batched import Graph
    with torch.cuda.stream(s):
        g = Graph(lambda: fr.reflux(s), s)
    corr = [w[2] - c[2] for w, c in zip(want, crse)]
    for rep in range(3):
        for fab, b in zip(crse_l.fabs, before):
            fab.t.copy_(b)
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            g.launch(s)
        torch.cuda.synchronize()
        for fab, b, c in zip(crse_l.fabs, before, corr):
            got = fab.t.cpu().numpy() - b.cpu().numpy()
            assert np.max(np.abs(got - c)) <= 1e-12 * max(1.0, np.max(np.abs(c))), rep
    g.close()
    fr.close()


def test_reflux_cells_two_orientations_reach_match_the_oracle():
    """Fine boxes one coarse cell apart in x and in y: the gap cells take an
    x-hi and an x-lo correction (or a y-hi and a y-lo one), which the plan
    applies in the reference's orientation order after the other faces, so
    the result equals the oracle's orientation-by-orientation sums bit for
    bit."""
    rng = np.random.default_rng(31)
    ncomp = 2
    domain = ((0, 0, 0), (7, 7, 7))
    crse = coarse_fabs(domain, 4, ncomp, rng)
    # coarse footprints [0,1]^3, x [3,4] (gap x = 2), y [3,4] (gap y = 2)
    fine_boxes = [((0, 0, 0), (3, 3, 3)), ((6, 0, 0), (9, 3, 3)),
                  ((0, 6, 0), (3, 9, 3))]
    cflux, fflux, _ = make_fluxes(domain, crse, fine_boxes, ncomp, rng, zero_boundary=False)
    clevel, cfl, flevel, ffl, cfabs = to_device(domain, crse, cflux, fine_boxes, fflux)
    dx = (0.125, 0.125, 0.125)
    want = [(lo, hi, data.copy(), vlo, vhi) for (lo, hi, data, vlo, vhi) in crse]
    rf.reflux(want, cflux, fine_boxes, fflux, dx, domain)
    fr = FluxRegister(clevel, cfl, flevel, ffl, dx, ncomp=ncomp)
    fr.reflux()
    torch.cuda.synchronize()
    for fab, w in zip(cfabs, want):
        assert np.array_equal(fab.t.cpu().numpy(), w[2])
    fr.close()
