#!/usr/bin/env python3
"""Benchmark of the rf=2 AMR inter-level transfer path on one MI355X.

Step = one ghost-shell fill (3 points) of the whole uneven level: 512 fine
boxes with extents from {8,...,64} per dimension, mixed VC/CC centring,
natural-family orders 1/3/5, 10 FP64 grid functions, replayed as one
captured hipGraph around the persistent batched kernel.  Inputs are resident
in HBM before the timed region.

Prints ONE JSON line.  `value` = fine points written per second (all boxes,
all grid functions).  `roofline.achieved` = algorithmic bytes per launch
(8 B per fine point written + 8 B per distinct coarse point of each box's
shell CoarseBoxes, per grid function) / average launch time from HIP events
on the launch stream.  `extra` carries the plain per-box path on the same
level and the single-level (64 x 32^3) operators.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default tile shapes (skip the in-context autotuner)")
    ap.add_argument("--cpu-boxes", type=int, default=512)
    ap.add_argument("--cpu-boxes-one-thread", type=int, default=48,
                    help="boxes of the one-thread CPU sample")
    ap.add_argument("--no-wave", action="store_true")
    ap.add_argument("--no-per-box", action="store_true", help="skip timing the per-box path")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="finish with three 512 MiB device copies (PMC byte-counter calibration)")
    return ap.parse_args()


def timed_events(fn, steps, stream):
    import torch
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    for a, b in ev:
        a.record(stream)
        fn()
        b.record(stream)
    torch.cuda.synchronize()
    return [a.elapsed_time(b) for a, b in ev]


def back_to_back(fn, reps, stream):
    """Average time of `reps` calls issued back to back between one pair of
    events (as a time-step loop issues fills): launch gaps overlap the
    previous call instead of adding to it."""
    import torch
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(reps):
        fn()
    b.record(stream)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def bench_uneven(args, stream, log):
    import torch
    from carpetx_amd import workloads
    from carpetx_amd.batched import Graph, WorkList

    t0 = time.time()
    wlz = workloads.uneven_512(seed=0)
    wl = WorkList()
    for b in wlz["boxes"]:
        wl.append_shell(b["op"], b["crse"], 0, b["fine"], 0, wlz["ncomp"], b["interior"])
    with torch.cuda.stream(stream):
        wl.upload(stream)
    tune_s = None
    if not args.no_tune:
        # tile shapes and z-chunks from the autotuner, timed in context
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        with torch.cuda.stream(stream):
            wl.autotune(reps=3, stream=stream)
        torch.cuda.synchronize()
        tune_s = time.perf_counter() - t1
        log("autotuned in %.1f s" % tune_s)
    st = wl.stats()
    log("uneven_512: %d patches, %d tiles, %.3e fine pts, %.3f GB algorithmic (setup %.1fs)"
        % (st["patches"], st["tiles"], st["fine_points"], st["algo_bytes"] / 1e9,
           time.time() - t0))

    with torch.cuda.stream(stream):
        graph = Graph(lambda: wl.run(stream), stream)
        for _ in range(args.warmup):
            graph.launch(stream)
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        times = timed_events(lambda: graph.launch(stream), args.steps, stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t_start
    ms = sum(times) / len(times)
    res = dict(ms=ms, wall_ms=1e3 * wall / args.steps, stats=st, tune_s=tune_s)

    # the list's write pattern alone: CarpetX's poisoning of the same points
    # (one store per point, no reads)
    with torch.cuda.stream(stream):
        wl.poison(stream=stream)
        tpz = timed_events(lambda: wl.poison(stream=stream), 10, stream)
        graph.launch(stream)  # the shells hold the fill again
        torch.cuda.synchronize()
    res["store_floor_ms"] = sum(tpz) / len(tpz)

    if args.no_per_box:
        res["per_box_ms"] = None
        return wlz, wl, res
    # plain per-box path on the same level (one launch per slab)
    nper = max(2, min(10, args.steps // 5))
    with torch.cuda.stream(stream):
        wl.run_per_patch(stream)
        torch.cuda.synchronize()
        tp = timed_events(lambda: wl.run_per_patch(stream), nper, stream)
    res["per_box_ms"] = sum(tp) / len(tp)
    return wlz, wl, res


def bench_level64(args, stream, log):
    import torch
    from carpetx_amd import workloads
    from carpetx_amd.batched import Graph, WorkList
    lv = workloads.level_64()
    out = {}
    for name, d in lv.items():
        wl = WorkList()
        for fab, reg in zip(d["fines"], d["regions"]):
            wl.append(d["op"], d["crse"], 0, fab, 0, d["ncomp"], reg)
        with torch.cuda.stream(stream):
            wl.upload(stream)
            if not args.no_tune:
                wl.autotune(reps=5, stream=stream)
            g = Graph(lambda: wl.run(stream), stream)
            for _ in range(3):
                g.launch(stream)
                wl.run(stream)
            # the list launched directly, back to back, as a time-step loop
            # issues it; and replays of a graph holding it alone, which add
            # the graph launch's own few microseconds to each
            mb = back_to_back(lambda: wl.run(stream), 50, stream)
            mg = back_to_back(lambda: g.launch(stream), 50, stream)
            tb = timed_events(lambda: g.launch(stream), 20, stream)
            wl.run_per_patch(stream)
            mp = back_to_back(lambda: wl.run_per_patch(stream), 10, stream)
        st = wl.stats()
        out[name] = dict(
            fine_pts=st["fine_points"], algo_MB=round(st["algo_bytes"] / 1e6, 2),
            batched_ms=round(mb, 4), batched_graph_ms=round(mg, 4),
            batched_single_launch_ms=round(sum(tb) / len(tb), 4),
            per_box_ms=round(mp, 4),
            batched_pts_per_s=st["fine_points"] / (mb * 1e-3),
            per_box_pts_per_s=st["fine_points"] / (mp * 1e-3),
            batched_hbm_frac=round(st["algo_bytes"] / (mb * 1e-3) / 1e9 / HBM_PEAK_GBS, 4))
        log("level_64 %s: batched %.4f ms, per-box %.4f ms" % (name, mb, mp))
    return out


def cpu_level_fill(d, log, whole=False):
    """The oracle's OpenMP build on a level fill's work: every fine FAB's
    ghost shell (its six slabs around its valid box, points in a
    neighbour's valid box included), or with whole every new FAB's box
    within the domain, prolongated from host copies of the same coarse
    data; median of a few runs."""
    import numpy as np
    import oracle
    from carpetx_amd.workloads import shell_slabs
    crse, ncomp = d["crse"], d["ncomp"]
    fine = d["new"] if whole else d["fine"]
    glo = [min(f.box.lo[k] for f in crse.fabs) for k in range(3)]
    ghi = [max(f.box.hi[k] for f in crse.fabs) for k in range(3)]
    g = np.zeros((ncomp,) + tuple(ghi[k] - glo[k] + 1 for k in (2, 1, 0)))
    for f in crse.fabs:
        sl = tuple(slice(f.box.lo[k] - glo[k], f.box.hi[k] - glo[k] + 1) for k in (2, 1, 0))
        g[(slice(None),) + sl] = f.t[:ncomp].cpu().numpy()
    op_d = d["op"]
    op = oracle.lookup(op_d.prolongation_type, op_d.prolongation_order, op_d.indextype,
                       kind="fast")
    ch = oracle.Fab(glo, ghi, ncomp, g)
    items, npts = [], 0
    dom = fine.domain
    for f, v in zip(fine.fabs, fine.valid):
        fh = oracle.Fab(f.box.lo, f.box.hi, ncomp)
        if whole:
            lo = [max(f.box.lo[k], dom.lo[k]) for k in range(3)]
            hi = [min(f.box.hi[k], dom.hi[k]) for k in range(3)]
            items.append((op, ch, fh, lo, hi, ncomp))
            npts += int(np.prod([hi[k] - lo[k] + 1 for k in range(3)])) * ncomp
            continue
        for sb in shell_slabs(f.box, v):
            items.append((op, ch, fh, sb.lo, sb.hi, ncomp))
            npts += sb.numPts * ncomp
    oracle.prolongate_many(items, kind="fast")  # warm-up
    reps = []
    for _ in range(5):
        t0 = time.perf_counter()
        oracle.prolongate_many(items, kind="fast")
        reps.append(time.perf_counter() - t0)
    med = sorted(reps)[len(reps) // 2]
    threads = oracle.num_threads("fast")
    return dict(cpu_ms=round(1e3 * med, 3), cpu_pts=npts, cpu_pts_per_s=npts / med,
                cpu_threads=threads)


def bench_level_fill(args, stream, log):
    """Ghost fill of a whole fine level (64 boxes of 32^3, 3 ghosts) from a
    coarse level of 2^3 FABs: the fine ghost points outside every fine valid
    box, coarse patches gathered where a stencil spans coarse FABs.  The
    "_bc" rows fill the same level with the coarse boundary points computed
    by the boundary conditions (reflection on the lower faces, linear
    extrapolation on the upper ones) instead of read from the coarse FABs."""
    import torch
    from carpetx_amd import workloads
    from carpetx_amd.batched import BoundaryConditions, Graph, WorkList
    lv = workloads.level_fill_64()
    cases = []
    for name, d in lv.items():
        cases.append((name, d, None))
        cases.append((name + "_bc", d, BoundaryConditions(
            [["reflection"] * 3, ["linear_extrapolation"] * 3],
            parities=[(1, 1, 1)] * d["ncomp"])))
    out = {}
    for name, d, bc in cases:
        wl = WorkList()
        wl.append_level(d["op"], d["crse"], 0, d["fine"], 0, d["ncomp"], bc=bc)
        with torch.cuda.stream(stream):
            wl.upload(stream)
            if not args.no_tune:
                wl.autotune(reps=5, stream=stream)
            g = Graph(lambda: wl.run(stream), stream)
            for _ in range(3):
                g.launch(stream)
            mb = back_to_back(lambda: g.launch(stream), 50, stream)
            # the chain launched directly (gather, fill: no graph)
            md = back_to_back(lambda: wl.run(stream), 50, stream)
            wl.run_per_patch(stream)
            mp = back_to_back(lambda: wl.run_per_patch(stream), 10, stream)
            wl.poison(stream=stream)
            mf = back_to_back(lambda: wl.poison(stream=stream), 20, stream)
        st = wl.stats()
        out[name] = dict(
            patches=st["patches"], fine_pts=st["fine_points"],
            algo_MB=round(st["algo_bytes"] / 1e6, 2),
            batched_ms=round(mb, 4), batched_direct_ms=round(md, 4),
            per_patch_ms=round(mp, 4), store_floor_ms=round(mf, 4),
            batched_pts_per_s=st["fine_points"] / (mb * 1e-3),
            batched_hbm_frac=round(st["algo_bytes"] / (mb * 1e-3) / 1e9 / HBM_PEAK_GBS, 4))
        if bc is None and not args.no_cpu:
            out[name].update(cpu_level_fill(d, log))
        log("level fill %s: %s" % (name, out[name]))
        g.close()
        wl.close()
    return out


def bench_regrid(args, stream, log):
    """Regrid-time fills of the 64-box level: a new level of 64 boxes (half
    at old boxes' slots) filled whole from the coarse level (new_level), or
    from the old fine level where it has valid data and the coarse level
    elsewhere (remake_level); both with reflection on the lower faces and
    linear extrapolation on the upper ones, on the coarse patches and then
    on the new level's own boundary points."""
    import torch
    from carpetx_amd import workloads
    from carpetx_amd.batched import BoundaryConditions, Graph, WorkList
    lv = workloads.level_fill_64(regrid=True)
    out = {}
    for name, d in lv.items():
        bc = BoundaryConditions([["reflection"] * 3, ["linear_extrapolation"] * 3],
                                parities=[(1, 1, 1)] * d["ncomp"])
        for kind in ("new_level", "remake_level"):
            wl = WorkList()
            if kind == "new_level":
                wl.append_new_level(d["op"], d["crse"], 0, d["new"], 0, d["ncomp"], bc=bc)
            else:
                wl.append_remake_level(d["op"], d["crse"], 0, d["fine"], d["new"], 0,
                                       d["ncomp"], bc=bc)
            wl.append_boundaries(d["new"], d["indextype"], 0, d["ncomp"], bc)
            with torch.cuda.stream(stream):
                wl.upload(stream)
                if not args.no_tune:
                    wl.autotune(reps=5, stream=stream)
                g = Graph(lambda: wl.run(stream), stream)
                for _ in range(3):
                    g.launch(stream)
                mb = back_to_back(lambda: g.launch(stream), 30, stream)
                md = back_to_back(lambda: wl.run(stream), 30, stream)
                wl.run_per_patch(stream)
                mp = back_to_back(lambda: wl.run_per_patch(stream), 5, stream)
            st = wl.stats()
            key = "%s_%s" % (name, kind)
            out[key] = dict(
                patches=st["patches"], fine_pts=st["fine_points"],
                algo_MB=round(st["algo_bytes"] / 1e6, 2),
                batched_ms=round(mb, 4), batched_direct_ms=round(md, 4),
                per_patch_ms=round(mp, 4),
                batched_pts_per_s=st["fine_points"] / (mb * 1e-3),
                batched_hbm_frac=round(st["algo_bytes"] / (mb * 1e-3) / 1e9 / HBM_PEAK_GBS, 4))
            if kind == "new_level" and not args.no_cpu:
                out[key].update(cpu_level_fill(d, log, whole=True))
            log("regrid %s: %s" % (key, out[key]))
            g.close()
            wl.close()
    return out


def bench_adjoint(stream, log):
    """96^3 -> 192^3 VC 5th-order prolongation: forward vs adjoint (per-box
    launches on one stream), plus the CCC minmod and CCC ENO forwards and
    adjoints (linearised at the forward input: recomputing the branches, and
    from the forward's branch record)."""
    import torch
    from carpetx_amd.interp import Box, FArrayBox, get_interpolator
    out = {}
    region = Box((0, 0, 0), (191, 191, 191))
    for name, typ, order, it in [("poly5_vvv", "interpolate", 5, (0, 0, 0)),
                                 ("minmod_ccc", "minmod", 1, (1, 1, 1)),
                                 ("eno_ccc", "eno", 3, (1, 1, 1))]:
        P = get_interpolator(typ, order, it)
        cb = P.CoarseBox(region)
        g = torch.Generator(device="cuda").manual_seed(4)
        crse = FArrayBox(cb, 1, tensor=torch.rand((1,) + tuple(reversed(cb.size)),
                                                   dtype=torch.float64, device="cuda",
                                                   generator=g) * 2 - 1)
        fine = FArrayBox(region, 1)
        cbar = FArrayBox(cb, 1)
        with torch.cuda.stream(stream):
            for _ in range(3):
                P.interp(crse, 0, fine, 0, 1, region, stream=stream)
            ms_f = back_to_back(
                lambda: P.interp(crse, 0, fine, 0, 1, region, stream=stream), 20, stream)
        algo = 8 * (region.numPts + cb.numPts)
        rec = dict(fine_pts=region.numPts, algo_MB=round(algo / 1e6, 2),
                   fwd_ms=round(ms_f, 4), fwd_hbm_frac=round(algo / (ms_f * 1e-3) / 1e9 / HBM_PEAK_GBS, 4))
        if name == "poly5_vvv":
            # the same forward over 4 grid functions in one launch: enough
            # workgroups that their z-window warm-ups overlap other tiles'
            # stores (one grid function is the adjoint check's size)
            crse4 = FArrayBox(cb, 4, tensor=torch.rand((4,) + tuple(reversed(cb.size)),
                                                        dtype=torch.float64, device="cuda",
                                                        generator=g) * 2 - 1)
            fine4 = FArrayBox(region, 4)
            with torch.cuda.stream(stream):
                for _ in range(3):
                    P.interp(crse4, 0, fine4, 0, 4, region, stream=stream)
                ms_f4 = back_to_back(
                    lambda: P.interp(crse4, 0, fine4, 0, 4, region, stream=stream), 20, stream)
            rec.update(fwd_4gf_ms=round(ms_f4, 4),
                       fwd_4gf_hbm_frac=round(4 * algo / (ms_f4 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4))
            del crse4, fine4
        try:
            with torch.cuda.stream(stream):
                P.interp_adjoint(cbar, 0, fine, 0, 1, region, crse_in=crse, stream=stream)
                ms_a = back_to_back(lambda: P.interp_adjoint(cbar, 0, fine, 0, 1, region,
                                                             crse_in=crse, stream=stream),
                                    20, stream)
            rec.update(adj_ms=round(ms_a, 4), adj_over_fwd=round(ms_a / ms_f, 3),
                       adj_hbm_frac=round(algo / (ms_a * 1e-3) / 1e9 / HBM_PEAK_GBS, 4))
        except Exception as e:  # operators without a device adjoint
            rec.update(adj_ms=None, adj_note=str(e)[:120])
        nb = P.branch_bytes(region, 1)
        if nb > 0 and rec.get("adj_ms") is not None:
            # the forward records each fine point's branch, the adjoint is
            # the gather alone (what the autograd wrapper runs)
            br = torch.empty(nb, dtype=torch.uint8, device="cuda")
            with torch.cuda.stream(stream):
                P.interp_record(crse, 0, fine, 0, 1, region, br, stream=stream)
                ms_fr = back_to_back(lambda: P.interp_record(
                    crse, 0, fine, 0, 1, region, br, stream=stream), 20, stream)
                P.interp_adjoint_recorded(br, cbar, 0, fine, 0, 1, region, stream=stream)
                ms_ar = back_to_back(lambda: P.interp_adjoint_recorded(
                    br, cbar, 0, fine, 0, 1, region, stream=stream), 20, stream)
            rec.update(fwd_record_ms=round(ms_fr, 4), adj_recorded_ms=round(ms_ar, 4),
                       adj_recorded_over_fwd_record=round(ms_ar / ms_fr, 3))
        out[name] = rec
        log("adjoint %s: %s" % (name, rec))
    return out


def bench_wave(log, n=128, steps=16, ck=4):
    """Two-level subcycled wave evolution (coarse n^3, fine covering the
    central half), RK4, ghost shells prolongated at every stage; forward and
    backward (gradient w.r.t. the initial data, checkpoint every ck steps)."""
    import torch
    from carpetx_amd.wave import TwoLevelWave, loss
    w = TwoLevelWave(n=n)
    sc0, sf0 = w.gaussian(width=0.05)
    # warm-up: one step forward and backward (library and allocator set-up)
    yc = sc0.clone().requires_grad_(True)
    yf = sf0.clone().requires_grad_(True)
    loss(*w.evolve(yc, yf, 1, checkpoint_every=1)).backward()
    del yc, yf
    xc = sc0.clone().requires_grad_(True)
    xf = sf0.clone().requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    a, b = w.evolve(xc, xf, steps, checkpoint_every=ck)
    L = loss(a, b)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    L.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    # forward only, one CUDA graph per coarse step (every RK4 stage's ghost
    # fills and the restriction replayed together)
    advance, _, _ = w.graphed_step(sc0, sf0)
    advance()
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    for _ in range(steps):
        advance()
    torch.cuda.synchronize()
    t4 = time.perf_counter()
    with torch.no_grad():
        c, f = sc0.clone(), sf0.clone()
        w.step(c, f)
        torch.cuda.synchronize()
        t5 = time.perf_counter()
        for _ in range(steps):
            c, f = w.step(c, f)
        torch.cuda.synchronize()
        t6 = time.perf_counter()
    # forward only with the native RK4 stage (both levels) and batched
    # fills, eagerly and as one CUDA graph per coarse step
    from carpetx_amd.wave import NativeStep
    with torch.no_grad():
        ns = NativeStep(w, sc0, sf0)
        ns.run()
        torch.cuda.synchronize()
        t7 = time.perf_counter()
        for _ in range(steps):
            ns.run()
        torch.cuda.synchronize()
        t8 = time.perf_counter()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            ns.run()
        torch.cuda.current_stream().wait_stream(side)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            ns.run()
        graph.replay()
        torch.cuda.synchronize()
        t9 = time.perf_counter()
        for _ in range(steps):
            graph.replay()
        torch.cuda.synchronize()
        t10 = time.perf_counter()
        ns.close()
    # the same evolution with the native stage and its adjoint under
    # autograd (TwoLevelWave(native=True))
    wn = TwoLevelWave(n=n, native=True)
    yc = sc0.clone().requires_grad_(True)
    yf = sf0.clone().requires_grad_(True)
    loss(*wn.evolve(yc, yf, 1, checkpoint_every=1)).backward()
    del yc, yf
    nc = sc0.clone().requires_grad_(True)
    nf = sf0.clone().requires_grad_(True)
    torch.cuda.synchronize()
    t11 = time.perf_counter()
    an, bn = wn.evolve(nc, nf, steps, checkpoint_every=ck)
    Ln = loss(an, bn)
    torch.cuda.synchronize()
    t12 = time.perf_counter()
    Ln.backward()
    torch.cuda.synchronize()
    t13 = time.perf_counter()
    rec = dict(n=n, steps=steps, checkpoint_every=ck,
               fwd_native_autograd_ms_per_step=round(1e3 * (t12 - t11) / steps, 3),
               bwd_native_ms_per_step=round(1e3 * (t13 - t12) / steps, 3),
               native_grad_rel_diff=float((nc.grad - xc.grad).abs().max() /
                                          xc.grad.abs().max()),
               fwd_native_ms_per_step=round(1e3 * (t8 - t7) / steps, 3),
               fwd_native_graph_ms_per_step=round(1e3 * (t10 - t9) / steps, 3),
               fwd_ms_per_step=round(1e3 * (t1 - t0) / steps, 3),
               bwd_ms_per_step=round(1e3 * (t2 - t1) / steps, 3),
               fwd_nograd_ms_per_step=round(1e3 * (t6 - t5) / steps, 3),
               fwd_graph_ms_per_step=round(1e3 * (t4 - t3) / steps, 3),
               grad_norm=float(xc.grad.norm()), loss=float(L.detach()))
    log("wave: %s" % rec)
    return rec


def bench_reflux(stream, log):
    """Reflux of the 64-box level (128^3 coarse cells in 8 FABs, 64 fine
    boxes of 32^3 cells, 4 grid functions): one plan run (6 launches, one
    per face orientation), back to back; bytes per run from the plan."""
    import torch
    from carpetx_amd import workloads
    from carpetx_amd.interp import Box, FArrayBox
    from carpetx_amd.reflux import FluxRegister
    lv = workloads.level_fill_64(ncomp=4)["cons2_ccc"]
    crse, fine = lv["crse"], lv["fine"]
    g = torch.Generator(device="cuda").manual_seed(13)

    def faces(level):
        out = []
        for d in range(3):
            fl = []
            for v in level.valid:
                hi = list(v.hi)
                hi[d] += 1
                b = Box(v.lo, hi)
                fl.append(FArrayBox(b, 4, tensor=torch.rand((4,) + tuple(reversed(b.size)),
                                                             dtype=torch.float64, device="cuda",
                                                             generator=g)))
            out.append(fl)
        return out
    cf, ff = faces(crse), faces(fine)
    fr = FluxRegister(crse, cf, fine, ff, (1.0 / 128,) * 3)
    st = fr.stats()
    with torch.cuda.stream(stream):
        for _ in range(3):
            fr.reflux(stream)
        ms = back_to_back(lambda: fr.reflux(stream), 20, stream)
    fr.close()
    rec = dict(register_faces=st["faces"], MB=round(st["bytes"] / 1e6, 2), ms=round(ms, 4),
               GB_per_s=round(st["bytes"] / (ms * 1e-3) / 1e9, 1),
               hbm_frac=round(st["bytes"] / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4))
    log("reflux: %s" % rec)
    return rec


def bench_tagging(stream, log, n=256):
    """Box-in-box tags and the threshold test on a level of n^3 cells (two
    regions, a cube and a sphere): one double written per cell, then one
    read (and one tag byte written) per cell."""
    import torch
    from carpetx_amd.interp import Box, FArrayBox
    from carpetx_amd.tagging import BoxInBoxRegion, boxinbox_setup, error_est
    h = 1.0 / n
    regions = [BoxInBoxRegion("cube", num_levels=2, radius=[1.0, 0.25]),
               BoxInBoxRegion("sphere", num_levels=2, position=(0.2, -0.1, 0.05),
                              radius=[1.0, 0.3])]
    with torch.cuda.stream(stream):
        err = FArrayBox(Box((0, 0, 0), (n - 1,) * 3), 1, device="cuda")
        tags = torch.empty(tuple(err.t.shape), dtype=torch.int8, device="cuda")
        box_out = torch.empty(7, dtype=torch.int32, device="cuda")

        def tag():
            boxinbox_setup(regions, 1, (-0.5 + h / 2,) * 3, (h,) * 3, err, stream=stream)

        def est():
            error_est(err, 1.0, box_out=box_out, stream=stream, read=False)

        def est_tags():
            error_est(err, 1.0, tags=tags, box_out=box_out, stream=stream, read=False)

        for f in (tag, est, est_tags):
            f()
        torch.cuda.synchronize()
        count = int(box_out[0])
        out = {"cells": n ** 3, "tagged": count}
        for name, f, nbytes in (("boxinbox_tag", tag, 8), ("error_est", est, 8),
                                ("error_est_with_tags", est_tags, 9)):
            ms = back_to_back(f, 20, stream)
            gbs = nbytes * n ** 3 / (ms * 1e-3) / 1e9
            out[name] = {"ms": round(ms, 4), "algo_MB": round(nbytes * n ** 3 / 1e6, 1),
                         "GB_per_s": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 3)}
    log("tagging %d^3: %s" % (n, out))
    return out


def load_traffic(name):
    """HBM traffic per launch from the committed rocprofv3 PMC summary."""
    path = os.path.join(ROOT, "profiles", "k_fill_batched_uneven_512_pmc.json")
    try:
        with open(path) as f:
            return json.load(f).get(name)
    except (OSError, ValueError):
        return None


def _cpu_items(wlz, nboxes):
    """Host copies of the first `nboxes` boxes' coarse data under their
    shells, and one oracle work item per shell slab."""
    import numpy as np
    import oracle
    from carpetx_amd.workloads import shell_slabs
    items = []
    npts = 0
    for b in wlz["boxes"][:nboxes]:
        op_d = b["op"]
        op = oracle.lookup(op_d.prolongation_type, op_d.prolongation_order, op_d.indextype,
                           kind="fast")
        fab = b["fine"]
        slabs = shell_slabs(fab.box, b["interior"])
        # coarse sub-box covering every slab's CoarseBox
        clo = [1 << 30] * 3
        chi = [-(1 << 30)] * 3
        for s in slabs:
            lo, hi = oracle.coarse_box(op, s.lo, s.hi)
            clo = [min(clo[d], lo[d]) for d in range(3)]
            chi = [max(chi[d], hi[d]) for d in range(3)]
        crse_dev = b["crse"]
        sub = crse_dev.t[:, clo[2] - crse_dev.box.lo[2]:chi[2] - crse_dev.box.lo[2] + 1,
                         clo[1] - crse_dev.box.lo[1]:chi[1] - crse_dev.box.lo[1] + 1,
                         clo[0] - crse_dev.box.lo[0]:chi[0] - crse_dev.box.lo[0] + 1]
        ch = oracle.Fab(clo, chi, fab.ncomp, np.ascontiguousarray(sub.cpu().numpy()))
        fh = oracle.Fab(fab.box.lo, fab.box.hi, fab.ncomp)
        for s in slabs:
            items.append((op, ch, fh, s.lo, s.hi, fab.ncomp))
            npts += s.numPts * fab.ncomp
    return items, npts


def _cpu_median(items, warmups=2, runs=10):
    """Median seconds of prolongate_many over the items: `warmups` untimed
    runs, then the median of `runs` (the protocol of the baseline plan)."""
    import oracle
    for _ in range(warmups):
        oracle.prolongate_many(items, kind="fast")
    reps = []
    for _ in range(runs):
        t0 = time.perf_counter()
        oracle.prolongate_many(items, kind="fast")
        reps.append(time.perf_counter() - t0)
    s = sorted(reps)
    return 0.5 * (s[(runs - 1) // 2] + s[runs // 2]), reps


HOST_CPU = None  # taken at start, before the OpenMP runtime binds this thread


def host_cpu():
    """CPU model, physical cores and the CPUs this process may run on."""
    model, cores = "?", set()
    try:
        phys = core = None
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name":
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
    except OSError:
        pass
    return {"model": model, "physical_cores": len(cores) or None,
            "allowed_cpus": len(os.sched_getaffinity(0)),
            "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")}}


def cpu_baseline(wlz, nboxes, log):
    """The oracle's OpenMP build (reference CI flags) on the shells of the
    first `nboxes` boxes of the same level, host copies of the same data,
    on every host thread OpenMP offers."""
    import oracle
    items, npts = _cpu_items(wlz, nboxes)
    med, reps = _cpu_median(items)
    threads = oracle.num_threads("fast")
    cpu = HOST_CPU or host_cpu()
    log("cpu baseline: %d boxes, %.3e pts, median %.3fs over %d runs, %d threads, %s"
        % (nboxes, npts, med, len(reps), threads, cpu))
    return dict(value=npts / med, unit="fine pts/s", cores=threads, kind="port",
                host=cpu,
                sample="ghost shells of %d of the 512 boxes, 10 GF, %d fine pts per run; "
                       "oracle/liboracle_fast.so (-O3 -ffp-contract=fast -march=x86-64-v3 "
                       "-fopenmp), %d threads, OMP_PROC_BIND=%s OMP_PLACES=%s, 2 warm-ups "
                       "then the median of %d runs, %.1f s of CPU work in total"
                       % (nboxes, npts, threads, cpu["omp"]["OMP_PROC_BIND"],
                          cpu["omp"]["OMP_PLACES"], len(reps), sum(reps)))


def cpu_one_thread(wlz, nboxes, log):
    """The same oracle build on one host thread, over the shells of the
    first `nboxes` boxes (a bounded sample: a few seconds)."""
    import oracle
    items, npts = _cpu_items(wlz, nboxes)
    threads = oracle.num_threads("fast")
    oracle.set_num_threads(1)
    try:
        med, reps = _cpu_median(items)
    finally:
        oracle.set_num_threads(threads)
    log("cpu one thread: %d boxes, %.3e pts, median %.3fs over %d runs"
        % (nboxes, npts, med, len(reps)))
    return dict(value=npts / med, unit="fine pts/s", cores=1,
                sample="ghost shells of the first %d of the 512 boxes, 10 GF, %d fine "
                       "pts per run, 2 warm-ups then the median of %d runs"
                       % (nboxes, npts, len(reps)))


def main():
    # the CPU baseline's OpenMP threads: one per core, packed (read by the
    # OpenMP runtime when it starts, so before anything loads it)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    global HOST_CPU
    HOST_CPU = host_cpu()
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: the transfer path has no multi-GPU "
              "launch (each GPU would fill its own share of the boxes); run with "
              "--gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)

    def log(msg):
        print("[bench] " + msg, file=sys.stderr, flush=True)

    stream = torch.cuda.Stream()
    wlz, wl, res = bench_uneven(args, stream, log)
    st = res["stats"]
    ms = res["ms"]
    value = st["fine_points"] / (ms * 1e-3)
    achieved = st["algo_bytes"] / (ms * 1e-3) / 1e9
    line = {
        "metric": "fine pts/s + % of HBM peak, FP64, batched vs per-box, fwd and adjoint",
        "value": value,
        "unit": "fine pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded sin-product fields + 1e-6 noise, resident in HBM)",
        "config": {
            "workload": "uneven_512: ghost-shell fill (3 pts) of 512 boxes, extents "
                        "{8..64}^3, mixed VC/CC, natural o1/3/5, 10 GF, one hipGraph",
            "tile_shapes": ("autotuned in context (%.1f s)" % res["tune_s"]
                            if res["tune_s"] is not None else "defaults"),
            "patches": st["patches"], "tiles": st["tiles"],
            "fine_points": st["fine_points"], "algo_bytes": st["algo_bytes"],
        },
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": (load_traffic("uneven_512") or {}).get("bytes_per_launch"),
            # the same algorithmic bytes over the time of the list's store
            # pattern alone (one store per fill point, no reads): the most
            # this layout of partial-line writes allows
            "frac_ceiling_store_floor": round(
                st["algo_bytes"] / (res["store_floor_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        },
        "extra": {
            "wall_ms_per_step": round(res["wall_ms"], 4),
            "store_floor_ms": round(res["store_floor_ms"], 4),
            "fill_over_store_floor": round(ms / res["store_floor_ms"], 3),
        },
    }
    if res["per_box_ms"] is not None:
        line["extra"].update(
            per_box_ms=round(res["per_box_ms"], 4),
            per_box_pts_per_s=st["fine_points"] / (res["per_box_ms"] * 1e-3),
            batched_over_per_box=round(res["per_box_ms"] / ms, 2))
    if not args.no_extra:
        line["extra"]["level_64"] = bench_level64(args, stream, log)
        line["extra"]["level_fill_64"] = bench_level_fill(args, stream, log)
        line["extra"]["regrid_64"] = bench_regrid(args, stream, log)
        line["extra"]["adjoint_96_192"] = bench_adjoint(stream, log)
        line["extra"]["boxinbox_tag_256"] = bench_tagging(stream, log)
        line["extra"]["reflux_64"] = bench_reflux(stream, log)
        if not args.no_wave:
            line["extra"]["wave_two_level"] = bench_wave(log)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(wlz, args.cpu_boxes, log)
        line["extra"]["cpu_one_thread"] = cpu_one_thread(wlz, args.cpu_boxes_one_thread, log)
    else:
        line["cpu_baseline"] = None
    if args.copy_calibration:
        src = torch.rand(64 << 20, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
