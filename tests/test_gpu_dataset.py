"""Real-data parity: the reference's B-spline dataset (committed fixture: the
first 300 curves of Bundle_0, normalised by the loader as the reference's
driver does), plain and learned-seed builds against the long-double truth
and the reference GPU engine's restatement: lengths <= 1e-6 relative, sample
parameters <= 1e-6, per-interval sample counts identical outside the tie band."""
import os

import numpy as np
import pytest

import gcurval_amd as G
from gcurval_amd import dataset as D
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dataset_bundle0_head.txt")


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("learned", [False, True])
def test_dataset_parity(ctx, dim, learned):
    b = D.load_bspline_dataset([FIX], dim=dim)
    tr = O.Truth(b)
    Lt, cum = tr.lengths()
    counts = 32 * b.intervals_per_curve()
    pos = O.uniform_positions(counts)
    ut, segt = tr.reverse(counts, pos)
    ref = O.RefGpuPlain(b, O.opts(0, 3))
    ref.build()
    Lref = ref.tables()[0]
    with G.ArclengthCalculateGpu(ctx, b) as eng:
        if learned:
            from gcurval_amd.seed import SeedMLP
            with SeedMLP.pretrained(ctx) as m:
                eng.createArclengthTableLearned(m)
        else:
            eng.createArclengthTable()
        L = eng.getAllCurveArclengths()
        u, seg = eng.arclengthReverse(counts, None, True, return_segments=True)
    assert np.max(np.abs(L - Lt) / Lt) <= 1e-6
    assert np.max(np.abs(L - Lref) / Lref) <= 1e-6
    assert np.max(np.abs(u - ut)) <= 1e-6
    off = O.offsets(counts)
    ioff = O.offsets(b.intervals_per_curve())
    bad = 0
    for c in range(b.n_curve):
        s = pos[off[c]:off[c + 1]] * Lt[c]
        bounds = cum[ioff[c]:ioff[c + 1]]
        near = np.min(np.abs(s[:, None] - bounds[None, :]), axis=1) <= 2e-6 * Lt[c]
        bad += int(np.sum((seg[off[c]:off[c + 1]] != segt[off[c]:off[c + 1]]) & ~near))
    assert bad == 0
