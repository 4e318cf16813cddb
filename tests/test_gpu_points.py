"""GPU tests of point output for the prebuilt families (arclengthReverse with
gcurvalReverseReturnPoints): points at the sampled parameters against the
curve evaluated on the host (ellipse formula, scipy's B-spline), FP64."""
import numpy as np
import pytest
from scipy.interpolate import BSpline

import gcurval_amd as G
from gcurval_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_ellipse_points(ctx):
    e = W.ellipse_batch(300, seed=91)
    counts = np.full(300, 37, np.int64)
    counts[:3] = (0, 1, 2)
    with G.ArclengthCalculateGpu(ctx, e, normalize=True) as eng:
        eng.createArclengthTable()
        u = eng.arclengthReverse(counts, None, True)
        pts = eng.arclengthReverse(counts, None, True, what=G.gcurvalReverseReturnPoints)
    a, b = np.repeat(e.a, counts), np.repeat(e.b, counts)
    ref = np.stack([a * np.cos(u), b * np.sin(u)], axis=1)
    assert pts.shape == (int(counts.sum()), 2)
    assert np.max(np.abs(pts - ref)) <= 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize("learned", [False, True])
def test_bspline_points_mixed(ctx, learned):
    bs = W.bspline_batch(60, seed=92, degrees=(1, 2, 3, 5), dims=(2, 3), n_ctrl_range=(8, 90), hairpin_frac=0.2)
    counts = 9 * bs.intervals_per_curve()
    with G.ArclengthCalculateGpu(ctx, bs, normalize=True) as eng:
        if learned:
            from gcurval_amd.seed import SeedMLP
            with SeedMLP.pretrained(ctx) as m:
                eng.createArclengthTableLearned(m)
        else:
            eng.createArclengthTable()
        u, seg = eng.arclengthReverse(counts, None, True, return_segments=True)
        pts = eng.arclengthReverse(counts, None, True, what=G.gcurvalReverseReturnPoints)
        L = eng.getAllCurveArclengths()
    assert pts.shape == (int(counts.sum()), bs.ctrl_stride)
    off = np.zeros(bs.n_curve + 1, np.int64)
    np.cumsum(counts, out=off[1:])
    for c in range(bs.n_curve):
        p, D = int(bs.degree[c]), int(bs.dim[c])
        t = bs.knots[bs.knot_off[c]:bs.knot_off[c + 1]]
        P = bs.ctrl[bs.ctrl_off[c]:bs.ctrl_off[c + 1], :D]
        uc = u[off[c]:off[c + 1]]
        ref = BSpline(t, P, p)(uc)
        got = pts[off[c]:off[c + 1]]
        scale = np.max(np.abs(P))
        assert np.max(np.abs(got[:, :D] - ref)) <= 1e-12 * scale, c
        if D < bs.ctrl_stride:
            assert np.all(got[:, D:] == 0.0)
        # arc-length-uniform samples: consecutive chords never exceed the spacing
        chords = np.linalg.norm(np.diff(got[:, :D], axis=0), axis=1)
        assert np.all(chords <= L[c] / (counts[c] - 1) * (1 + 1e-5) + 1e-12)
