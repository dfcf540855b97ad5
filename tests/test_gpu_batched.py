"""GPU batched path: the persistent kernel over a device-resident work list,
ghost-shell decomposition, hipGraph replay, autotuner -- all checked against
the per-patch path and the oracle."""
import numpy as np
import pytest
import torch

import oracle
from carpetx_amd import workloads
from carpetx_amd.batched import POISON_BITS, Graph, WorkList
from carpetx_amd.interp import Box, FArrayBox

pytestmark = pytest.mark.gpu


def small_uneven(nboxes=24, seed=5):
    return workloads.uneven_512(seed=seed, nboxes=nboxes, ncomp=3)


def _shell_mask(b):
    fab, ib = b["fine"], b["interior"]
    m = torch.ones_like(fab.t, dtype=torch.bool)
    sl = tuple(slice(ib.lo[d] - fab.box.lo[d], ib.hi[d] - fab.box.lo[d] + 1) for d in (2, 1, 0))
    m[(slice(None),) + sl] = False
    return m


def fill_oracle(w):
    """Expected shells from the oracle, on host copies."""
    out = []
    for b in w["boxes"]:
        op = b["op"]
        oop = oracle.lookup(op.prolongation_type, op.prolongation_order, op.indextype)
        crse = b["crse"]
        ch = oracle.Fab(crse.box.lo, crse.box.hi, crse.ncomp, crse.t.cpu().numpy())
        fab = b["fine"]
        fh = oracle.Fab(fab.box.lo, fab.box.hi, fab.ncomp, np.full(tuple(fab.t.shape), 7.0))
        for s in workloads.shell_slabs(fab.box, b["interior"]):
            oracle.prolongate(oop, ch, fh, s.lo, s.hi)
        out.append(fh.arr)
    return out


def test_batched_shell_fill_matches_oracle_and_per_patch():
    w = small_uneven()
    expect = fill_oracle(w)
    wl = WorkList()
    for b in w["boxes"]:
        b["fine"].t.fill_(7.0)   # interior must stay untouched
        wl.append_shell(b["op"], b["crse"], 0, b["fine"], 0, w["ncomp"], b["interior"])
    wl.upload()
    wl.run()
    torch.cuda.synchronize()
    for b, e in zip(w["boxes"], expect):
        assert np.array_equal(b["fine"].t.cpu().numpy(), e)
    # the per-patch path gives the same bits
    for b in w["boxes"]:
        b["fine"].t.fill_(7.0)
    wl.run_per_patch()
    torch.cuda.synchronize()
    for b, e in zip(w["boxes"], expect):
        assert np.array_equal(b["fine"].t.cpu().numpy(), e)


def test_poison_writes_exactly_the_shell():
    """cxp_worklist_poison writes CarpetX's poison bit pattern into every
    shell point and nothing else; a fill afterwards overwrites all of it."""
    w = small_uneven(nboxes=12, seed=9)
    wl = WorkList()
    for b in w["boxes"]:
        b["fine"].t.fill_(7.0)
        wl.append_shell(b["op"], b["crse"], 0, b["fine"], 0, w["ncomp"], b["interior"])
    wl.upload()
    wl.poison()
    torch.cuda.synchronize()
    for b in w["boxes"]:
        bits = b["fine"].t.cpu().numpy().view(np.uint64)
        m = _shell_mask(b).cpu().numpy()
        assert (bits[m] == np.uint64(POISON_BITS)).all()
        assert (b["fine"].t.cpu().numpy()[~m] == 7.0).all()
    wl.run()
    torch.cuda.synchronize()
    for b in w["boxes"]:
        assert torch.isfinite(b["fine"].t).all()


def test_graph_replays_match_eager():
    """A captured fill replayed several times, with eager work and syncs in
    between, reproduces the eager result every time."""
    w = small_uneven(nboxes=12, seed=6)
    wl = WorkList()
    for b in w["boxes"]:
        wl.append_shell(b["op"], b["crse"], 0, b["fine"], 0, w["ncomp"], b["interior"])
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        wl.upload(s)
        wl.run(s)
    torch.cuda.synchronize()
    eager = [b["fine"].t.clone() for b in w["boxes"]]
    with torch.cuda.stream(s):
        g = Graph(lambda: wl.run(s), s)
    for rep in range(4):
        for b in w["boxes"]:
            b["fine"].t.fill_(-1.0 - rep)
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            g.launch(s)
        torch.cuda.synchronize()
        # interior keeps the fill value, shell equals eager
        for b, e in zip(w["boxes"], eager):
            got = b["fine"].t
            assert torch.all(b["fine"].view(b["interior"]) == -1.0 - rep)
            shell = _shell_mask(b)
            assert torch.equal(got[shell], e[shell]), rep
    g.close()


def test_dynamic_tail_covers_every_tile_on_every_launch():
    """A list with at least four tiles per workgroup of the persistent grid,
    so the fill kernel hands out its smallest tiles through per-XCD claim
    counters.  The shell is poisoned before every launch, eager and
    replayed: each launch must reproduce the per-patch fill bit for bit, so
    no queued tile is skipped and the counters are back at zero for the
    next launch."""
    w = workloads.uneven_512(seed=11, nboxes=160, ncomp=10)
    wl = WorkList()
    for b in w["boxes"]:
        wl.append_shell(b["op"], b["crse"], 0, b["fine"], 0, w["ncomp"], b["interior"])
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        wl.upload(s)
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    assert wl.stats()["tiles"] >= 4 * 8 * cus
    for b in w["boxes"]:
        b["fine"].t.fill_(7.0)
    with torch.cuda.stream(s):
        wl.run_per_patch(s)
    torch.cuda.synchronize()
    ref = [b["fine"].t.clone() for b in w["boxes"]]
    with torch.cuda.stream(s):
        g = Graph(lambda: wl.run(s), s)
    for launch in range(5):
        with torch.cuda.stream(s):
            wl.poison(stream=s)
            if launch % 2:
                g.launch(s)
            else:
                wl.run(s)
        torch.cuda.synchronize()
        for b, e in zip(w["boxes"], ref):
            assert torch.equal(b["fine"].t, e), launch
    g.close()
    wl.close()


def test_autotune_keeps_results_identical():
    w = small_uneven(nboxes=16, seed=7)
    wl = WorkList()
    for b in w["boxes"]:
        wl.append_shell(b["op"], b["crse"], 0, b["fine"], 0, w["ncomp"], b["interior"])
    wl.upload()
    wl.run()
    torch.cuda.synchronize()
    before = [b["fine"].t.clone() for b in w["boxes"]]
    wl.autotune(reps=1)
    for b in w["boxes"]:
        b["fine"].t.zero_()
    wl.run()
    torch.cuda.synchronize()
    for b, e in zip(w["boxes"], before):
        m = _shell_mask(b)
        assert torch.equal(b["fine"].t[m], e[m])


def test_level64_batched_equals_oracle_sample():
    lv = workloads.level_64(nboxes=8)
    for name, d in lv.items():
        wl = WorkList()
        for fab, reg in zip(d["fines"], d["regions"]):
            wl.append(d["op"], d["crse"], 0, fab, 0, d["ncomp"], reg)
        wl.upload()
        wl.run()
        torch.cuda.synchronize()
        op = d["op"]
        oop = oracle.lookup(op.prolongation_type, op.prolongation_order, op.indextype)
        ch = oracle.Fab(d["crse"].box.lo, d["crse"].box.hi, d["ncomp"], d["crse"].t.cpu().numpy())
        for fab, reg in list(zip(d["fines"], d["regions"]))[:3]:
            fh = oracle.Fab(reg.lo, reg.hi, d["ncomp"])
            oracle.prolongate(oop, ch, fh, reg.lo, reg.hi)
            assert np.array_equal(fab.t.cpu().numpy(), fh.arr), name


def test_worklist_rejects_bad_patches():
    from carpetx_amd._lib import CxpError
    from carpetx_amd.interp import get_interpolator
    P = get_interpolator("natural", 5, (0, 0, 0))
    crse = FArrayBox(Box((0, 0, 0), (9, 9, 9)), 1)
    fine = FArrayBox(Box((0, 0, 0), (15, 15, 15)), 1)
    wl = WorkList()
    with pytest.raises(CxpError) as e:
        wl.append(P, crse, 0, fine, 0, 1, Box((0, 0, 0), (15, 15, 15)))
    assert e.value.status == -5
    with pytest.raises(CxpError):
        wl.run()   # not uploaded


def test_full_uneven_level_properties():
    """The whole 512-box, 10-field uneven level (the headline workload), by
    size-independent properties: after poisoning the shells, one batched
    fill leaves no poison in any shell and every interior bit as it was;
    the per-patch path and a graph replay give the same bits."""
    w = workloads.uneven_512()
    assert len(w["boxes"]) == 512 and w["ncomp"] == 10
    wl = WorkList()
    for b in w["boxes"]:
        wl.append_shell(b["op"], b["crse"], 0, b["fine"], 0, w["ncomp"], b["interior"])
    wl.upload()
    before = [b["fine"].t.clone() for b in w["boxes"]]
    wl.poison()
    wl.run()
    torch.cuda.synchronize()
    bits = torch.tensor(POISON_BITS - (1 << 64), dtype=torch.int64)
    filled = []
    for b, e in zip(w["boxes"], before):
        t = b["fine"].t
        m = _shell_mask(b)
        assert not torch.any(t.view(torch.int64)[m] == bits), "poison left in a shell"
        assert torch.equal(t.view(torch.int64)[~m], e.view(torch.int64)[~m])
        filled.append(t.clone())
    for b in w["boxes"]:
        b["fine"].t.fill_(7.0)
    wl.run_per_patch()
    torch.cuda.synchronize()
    for b, f in zip(w["boxes"], filled):
        m = _shell_mask(b)
        assert torch.equal(b["fine"].t.view(torch.int64)[m], f.view(torch.int64)[m])
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        g = Graph(lambda: wl.run(s), s)
        wl.poison(stream=s)
        g.launch(s)
    torch.cuda.synchronize()
    for b, f in zip(w["boxes"], filled):
        m = _shell_mask(b)
        assert torch.equal(b["fine"].t.view(torch.int64)[m], f.view(torch.int64)[m])
    g.close()
    wl.close()


def test_full_level_fill_with_boundary_conditions_batched_equals_per_patch():
    """The whole 64-box level ghost fill with boundary conditions on the
    coarse patches (reflection below, linear extrapolation above) and on
    the fine level.  Ghosts start as NaN except where another FAB's valid
    region covers them, which hold that FAB's values (the same-level sync
    the reference runs before its boundary conditions).  Then every point
    is written (no NaN left), valid and synced points keep their bits, and
    batched and per-patch runs give the same bits."""
    from carpetx_amd.batched import BoundaryConditions
    lv = workloads.level_fill_64()
    for name, d in lv.items():
        bc = BoundaryConditions([["reflection"] * 3, ["linear_extrapolation"] * 3],
                                parities=[(1, 1, 1)] * d["ncomp"])
        fine = d["fine"]
        valid_data = [torch.rand_like(f.t) for f in fine.fabs]

        def sl(f, lo, hi):
            return (slice(None),) + tuple(slice(lo[k] - f.box.lo[k], hi[k] - f.box.lo[k] + 1)
                                          for k in (2, 1, 0))

        def reset():
            for f in fine.fabs:
                f.t.fill_(float("nan"))
            for j, (fj, vj) in enumerate(zip(fine.fabs, fine.valid)):
                for f in fine.fabs:
                    lo = [max(vj.lo[k], f.box.lo[k]) for k in range(3)]
                    hi = [min(vj.hi[k], f.box.hi[k]) for k in range(3)]
                    if all(lo[k] <= hi[k] for k in range(3)):
                        f.t[sl(f, lo, hi)] = valid_data[j][sl(fj, lo, hi)]
            torch.cuda.synchronize()
        wl = WorkList()
        wl.append_level(d["op"], d["crse"], 0, fine, 0, d["ncomp"], bc=bc)
        wl.append_boundaries(fine, d["indextype"], 0, d["ncomp"], bc)
        wl.upload()
        reset()
        start = [f.t.clone() for f in fine.fabs]
        wl.run()
        torch.cuda.synchronize()
        got = [f.t.clone() for f in fine.fabs]
        for g_, s0 in zip(got, start):
            assert not torch.isnan(g_).any(), name
            kept = ~torch.isnan(s0)
            assert torch.equal(g_[kept], s0[kept]), name
        reset()
        wl.run_per_patch()
        torch.cuda.synchronize()
        for f, g_ in zip(fine.fabs, got):
            assert torch.equal(f.t.view(torch.int64), g_.view(torch.int64)), name
        wl.close()


def test_full_regrid_fills_batched_equal_per_patch():
    """The whole 64-box level's regrid-time fills (new level, and remake
    from the old level with half its boxes kept), with boundary conditions
    on the coarse patches and the new level: from the same starting
    values, batched and per-patch runs give the same bits, and every point
    of the new FABs is written."""
    from carpetx_amd.batched import BoundaryConditions
    lv = workloads.level_fill_64(regrid=True)
    for name, d in lv.items():
        bc = BoundaryConditions([["reflection"] * 3, ["linear_extrapolation"] * 3],
                                parities=[(1, 1, 1)] * d["ncomp"])
        for f in d["fine"].fabs:
            f.t.copy_(torch.rand_like(f.t))
        new = d["new"]
        for kind in ("new", "remake"):
            wl = WorkList()
            if kind == "new":
                wl.append_new_level(d["op"], d["crse"], 0, new, 0, d["ncomp"], bc=bc)
            else:
                wl.append_remake_level(d["op"], d["crse"], 0, d["fine"], new, 0, d["ncomp"], bc=bc)
            wl.append_boundaries(new, d["indextype"], 0, d["ncomp"], bc)
            wl.upload()
            for f in new.fabs:
                f.t.fill_(float("nan"))
            wl.run()
            torch.cuda.synchronize()
            got = [f.t.clone() for f in new.fabs]
            for g_ in got:
                assert not torch.isnan(g_).any(), (name, kind)
            for f in new.fabs:
                f.t.fill_(float("nan"))
            wl.run_per_patch()
            torch.cuda.synchronize()
            for f, g_ in zip(new.fabs, got):
                assert torch.equal(f.t.view(torch.int64), g_.view(torch.int64)), (name, kind)
            wl.close()
