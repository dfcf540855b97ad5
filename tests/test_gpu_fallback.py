"""The hand-derived linear-fallback fixtures (golden/fallback_kinks.json)
through the device library: every case along every axis, per-var and
per-group, must give the exact rational worked out by hand
(prolongate_3d_rf2_impl.hxx:1535-1667, :1990-2137)."""
from fractions import Fraction

import numpy as np
import pytest
import torch

from carpetx_amd.interp import Box, FArrayBox, get_interpolator
from test_oracle_fallback import CASES, kink_fabs

pytestmark = pytest.mark.gpu


def run_device(case, axis, profiles, per_group):
    _, crse_h, rlo, rhi = kink_fabs(case, axis, profiles)
    it = tuple(1 if d == axis else 0 for d in range(3))
    P = get_interpolator(case["family"], case["order"], it)
    crse = FArrayBox(Box(crse_h.lo, crse_h.hi), crse_h.ncomp,
                     tensor=torch.from_numpy(crse_h.arr.copy()).cuda())
    fbox = Box([x - 1 for x in rlo], [x + 1 for x in rhi])
    fine = FArrayBox(fbox, len(profiles),
                     tensor=torch.full((len(profiles),) + tuple(reversed(fbox.size)), 777.0,
                                       dtype=torch.float64, device="cuda"))
    P.interp(crse, 0, fine, 0, len(profiles), Box(rlo, rhi), per_group=per_group)
    torch.cuda.synchronize()
    arr = fine.t.cpu().numpy()
    inner = arr[:, 1:-1, 1:-1, 1:-1]
    outside = arr.copy()
    outside[:, 1:-1, 1:-1, 1:-1] = 777.0
    assert np.all(outside == 777.0)
    return inner


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_fallback_kink_on_device(case):
    for axis in range(3):
        for per_group, key in ((False, "per_var"), (True, "per_group")):
            got = run_device(case, axis, [case["profile"]], per_group)
            want = float(Fraction(case[key]))
            assert np.all(got == want), (case["name"], axis, key, got.ravel(), want)


def test_per_group_slope_test_spans_components_on_device():
    case = {"family": "poly-cons3lfb", "order": 3, "off": 0}
    comps = [["2", "0", "1", "2", "3"], ["0", "1", "0", "1", "2"]]
    for axis in range(3):
        v = run_device(case, axis, comps, per_group=False)
        assert np.all(v[0] == 39 / 64) and np.all(v[1] == 0.25)
        g = run_device(case, axis, comps, per_group=True)
        assert np.all(g[0] == 0.75) and np.all(g[1] == 0.25)
