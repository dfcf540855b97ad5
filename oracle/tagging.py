"""Box-in-box refinement tags and ErrorEst's threshold test, restated in
numpy (BoxInBox/src/boxinbox.cxx:20-45, 102-174; CarpetX/src/driver.cxx:
1025-1080).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
tag kernels, never by the product (carpetx_amd).

For each cell i of [lo, hi] (global indices of the cell-centred interior
loop, boxinbox.cxx:153) the cell centre is X = x0 + i dx
(Loop/src/loop.hxx:152-153, cell centring), and for each region that is
active with level < num_levels (:157-159), Rmax = radius - dx (:166) and
q = (X - position) / Rmax:
  * sphere: sum(pow2(q)) <= 1, the sum folded from zero in x, y, z order
    (:20-24; Arith/src/vect.hxx:776-779);
  * cube:   maximum(abs(q)) <= 1, folded from -inf with max1, which
    propagates NaN (:26-31; Arith/src/vect.hxx:737-741, defs.hxx:224-227).
regrid_error = 1 if any region holds the cell, else 0.  ErrorEst sets a tag
where regrid_error >= regrid_error_threshold (driver.cxx:1064-1066).

Every operation is one IEEE double operation in numpy (no contraction), as
in the device build.  The restatement is pinned by
tests/test_oracle_tagging.py against integer brute force on dyadic grids,
where every coordinate and quotient is exact.
"""
import numpy as np

SPHERE, CUBE = 0, 1


def _max1(x, y):
    return np.where(x != x, x, np.where(y != y, y, np.where(x < y, y, x)))


def boxinbox_tag(regions, level, x0, dx, lo, hi):
    """regrid_error over [lo, hi] as a float64 array (nz, ny, nx).

    regions: sequence of dicts with keys active, shape, num_levels,
    position (3), radius (3; the per-direction radius at this level)."""
    applying = [r for r in regions if r["active"] and level < r["num_levels"]]
    for r in applying:
        if any(float(v) <= 0.0 for v in r["radius"]):
            raise ValueError("non-positive radius")
    axes = [x0[d] + np.arange(lo[d], hi[d] + 1, dtype=np.float64) * dx[d]
            for d in range(3)]
    X = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")[::-1]  # x, y, z
    refine = np.zeros(X[0].shape, dtype=bool)
    with np.errstate(divide="ignore", invalid="ignore"):
        for r in applying:
            rmax = [float(r["radius"][d]) - float(dx[d]) for d in range(3)]
            q = [(X[d] - float(r["position"][d])) / rmax[d] for d in range(3)]
            if r["shape"] == SPHERE:
                s = np.zeros_like(q[0])
                for d in range(3):
                    s = s + q[d] * q[d]
                sr = s <= 1.0
            elif r["shape"] == CUBE:
                m = np.full_like(q[0], -np.inf)
                for d in range(3):
                    m = _max1(m, np.abs(q[d]))
                sr = m <= 1.0
            else:
                raise ValueError("unknown shape")
            refine |= sr
    return refine.astype(np.float64)


def error_est(err, lo, threshold):
    """Tags (int8: amrex::TagBox::SET = 2, CLEAR = 0) of an array over the box
    from lo, and the number and bounding box (lo, hi; None when empty) of the
    set cells (driver.cxx:1064-1066)."""
    tags = np.where(err >= threshold, 2, 0).astype(np.int8)
    idx = np.nonzero(tags)
    count = int(idx[0].size)
    if count == 0:
        return tags, 0, None
    blo = tuple(int(idx[2 - d].min()) + lo[d] for d in range(3))
    bhi = tuple(int(idx[2 - d].max()) + lo[d] for d in range(3))
    return tags, count, (blo, bhi)
