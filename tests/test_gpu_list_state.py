"""Work-list state rules enforced by the library: a graph refuses to replay a
list changed since its capture, launches of one list on two streams are
ordered, and the recomputing ENO adjoint refuses graph capture."""
import numpy as np
import pytest
import torch

from carpetx_amd import workloads
from carpetx_amd._lib import CxpError
from carpetx_amd.batched import Graph, WorkList
from carpetx_amd.interp import Box, FArrayBox, get_interpolator

pytestmark = pytest.mark.gpu

STATE = -10


def shell_list(nboxes, seed, ncomp):
    w = workloads.uneven_512(seed=seed, nboxes=nboxes, ncomp=ncomp)
    wl = WorkList()
    for b in w["boxes"]:
        wl.append_shell(b["op"], b["crse"], 0, b["fine"], 0, w["ncomp"], b["interior"])
    return w, wl


def test_graph_refuses_replay_after_the_list_changes():
    w, wl = shell_list(12, 6, 2)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        wl.upload(s)
        g = Graph(lambda: wl.run(s), s)
        g.launch(s)
        s.synchronize()
        # a re-upload may reallocate the list's buffers: the graph is stale
        wl.upload(s)
        with pytest.raises(CxpError) as e:
            g.launch(s)
        assert e.value.status == STATE
        g.close()
        g = Graph(lambda: wl.run(s), s)
        g.launch(s)
        s.synchronize()
        # an append changes the tiles: stale again
        b = w["boxes"][0]
        wl.append_shell(b["op"], b["crse"], 0, b["fine"], 0, w["ncomp"], b["interior"])
        with pytest.raises(CxpError) as e:
            g.launch(s)
        assert e.value.status == STATE
        g.close()
        wl.upload(s)
        g = Graph(lambda: wl.run(s), s)
        wl.close()   # destroyed: the graph refuses too
        with pytest.raises(CxpError) as e:
            g.launch(s)
        assert e.value.status == STATE
        g.close()


def test_launches_on_two_streams_are_ordered():
    """Poison and fill one list alternately on two streams with no host
    synchronisation; the library orders each launch after the list's
    previous one, so the claim counters stay consistent and a final launch
    reproduces the per-patch fill bit for bit."""
    w, wl = shell_list(160, 11, 10)
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    with torch.cuda.stream(streams[0]):
        wl.upload(streams[0])
        wl.run_per_patch(streams[0])
    torch.cuda.synchronize()
    ref = [b["fine"].t.clone() for b in w["boxes"]]
    for i in range(8):
        s = streams[i % 2]
        with torch.cuda.stream(s):
            wl.poison(stream=s)
            wl.run(s)
    torch.cuda.synchronize()
    for b, e in zip(w["boxes"], ref):
        assert torch.equal(b["fine"].t, e)
    s = streams[1]
    with torch.cuda.stream(s):
        wl.poison(stream=s)
        wl.run(s)
    torch.cuda.synchronize()
    for b, e in zip(w["boxes"], ref):
        assert torch.equal(b["fine"].t, e)
    wl.close()


def test_recomputing_eno_adjoint_refuses_capture():
    P = get_interpolator("eno", 3, (1, 1, 1))
    region = Box((0, 0, 0), (15, 15, 15))
    cb = P.CoarseBox(region)
    rng = np.random.default_rng(3)
    crse = FArrayBox(cb, 1, tensor=torch.from_numpy(
        rng.uniform(-1, 1, (1,) + tuple(reversed(cb.size)))).cuda())
    crse_bar = FArrayBox(cb, 1)
    fine_bar = FArrayBox(region, 1, tensor=torch.ones((1, 16, 16, 16), dtype=torch.float64,
                                                      device="cuda"))
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        # eager: works, twice (the second call reuses the stream's buffer)
        P.interp_adjoint(crse_bar, 0, fine_bar, 0, 1, region, crse_in=crse, stream=s)
        P.interp_adjoint(crse_bar, 0, fine_bar, 0, 1, region, crse_in=crse, stream=s)
        s.synchronize()
        once = crse_bar.t.clone() / 2
        err = []

        def body():
            try:
                P.interp_adjoint(crse_bar, 0, fine_bar, 0, 1, region, crse_in=crse, stream=s)
            except CxpError as e:
                err.append(e.status)
        g = Graph(body, s)
        g.close()
    assert err == [STATE]
    assert torch.isfinite(once).all() and once.abs().max() > 0
