"""GPU parity of uniform-k and seeded (per-interval k) table builds.

A seeded build starts knot interval i from 2^k_i roots and still runs the
adaptive error test on every root, so its lengths and samples must meet the
same bars as the plain build: <= 1e-6 relative length and <= 1e-6 sample
parameters against the long-double truth (FP32 mode, normalised curves),
identical per-interval sample counts outside the tie band.
"""
import numpy as np
import pytest

import gcurval_amd as G
from gcurval_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

def assert_quads(cost, st):
    """Per-interval quadrature counts add up to the build's: exactly without
    re-runs; a tile re-run with the full task stack also counted its first pass."""
    if st["reruns"] == 0:
        assert int(cost.sum()) == st["quadratures"]
    else:
        assert int(cost.sum()) < st["quadratures"]



@pytest.fixture(scope="module")
def cubic():
    b = W.bspline_batch(40, seed=51, degrees=(3,), dims=(3,), n_ctrl_range=(8, 150))
    tr = O.Truth(b)
    L, cum = tr.lengths()
    counts = 32 * b.intervals_per_curve()
    ut, segt = tr.reverse(counts, O.uniform_positions(counts))
    return b, L, cum, counts, ut, segt


def check_samples(eng, b, counts, ut, segt, L, cum):
    u, seg = eng.arclengthReverse(counts, None, True, return_segments=True)
    assert np.max(np.abs(u - ut)) <= 1e-6
    off = O.offsets(counts)
    ioff = O.offsets(b.intervals_per_curve())
    bad = 0
    pos = O.uniform_positions(counts)
    for c in range(b.n_curve):
        s = pos[off[c]:off[c + 1]] * L[c]
        bounds = cum[ioff[c]:ioff[c + 1]]
        near = np.min(np.abs(s[:, None] - bounds[None, :]), axis=1) <= 2e-6 * L[c]
        bad += int(np.sum((seg[off[c]:off[c + 1]] != segt[off[c]:off[c + 1]]) & ~near))
    assert bad == 0


@pytest.mark.parametrize("k", [0, 2, 4, 6])
def test_uniform_k_builds(ctx, cubic, k):
    b, L, cum, counts, ut, segt = cubic
    with G.ArclengthCalculateGpu(ctx, b, normalize=True) as eng:
        d_cost = ctx.empty(b.n_intervals, np.int32)
        eng.createArclengthTableUniform(k, d_cost)
        Lg = eng.getAllCurveArclengths()
        assert np.max(np.abs(Lg - L) / L) <= 1e-6
        cost = d_cost.download()
        st = eng.stats()
        assert_quads(cost, st)
        assert np.all(cost >= 3 * (1 << k))
        assert st["overflow"] == 0
        check_samples(eng, b, counts, ut, segt, L, cum)
        d_cost.free()


def test_seeded_random_k_matches_truth(ctx, cubic):
    b, L, cum, counts, ut, segt = cubic
    rng = np.random.default_rng(7)
    k = rng.integers(0, 7, b.n_intervals).astype(np.int8)
    with G.ArclengthCalculateGpu(ctx, b, normalize=True) as eng:
        d_k = ctx.array(k)
        d_cost = ctx.empty(b.n_intervals, np.int32)
        eng.createArclengthTableSeeded(d_k, d_cost)
        seeded = eng.root_layout()[0]
        assert seeded
        Lg = eng.getAllCurveArclengths()
        assert np.max(np.abs(Lg - L) / L) <= 1e-6
        cost = d_cost.download()
        keff = np.maximum(k, eng.seed_floor()).astype(np.int64)
        assert np.all(cost >= 3 * (1 << keff))
        assert_quads(cost, eng.stats())
        _, rp, _, _, _ = eng.tables()
        assert rp.size == int(np.sum(1 << keff))
        # dense roots: interval i's block starts at the sum of 2^k_j over j < i
        np.testing.assert_array_equal(eng.root_counts(), keff)
        np.testing.assert_array_equal(eng.root_offsets(), np.concatenate([[0], np.cumsum(1 << keff)[:-1]]))
        check_samples(eng, b, counts, ut, segt, L, cum)
        d_k.free()
        d_cost.free()


def test_seeded_all_six_equals_plain(ctx, cubic):
    b = cubic[0]
    with G.ArclengthCalculateGpu(ctx, b, normalize=True) as eng:
        eng.createArclengthTable()
        plain = eng.tables()
        d_k = ctx.array(np.full(b.n_intervals, 6, np.int8))
        eng.createArclengthTableSeeded(d_k)
        seeded = eng.tables()
        for x, y in zip(plain, seeded):
            np.testing.assert_array_equal(x, y)
        d_k.free()


def test_seeded_mixed_classes_and_ellipses(ctx):
    b = W.bspline_batch(40, seed=52, degrees=(2, 3, 5), dims=(2, 3), n_ctrl_range=(16, 80), hairpin_frac=0.3)
    tr = O.Truth(b)
    L, cum = tr.lengths()
    counts = 16 * b.intervals_per_curve()
    ut, segt = tr.reverse(counts, O.uniform_positions(counts))
    k = np.random.default_rng(8).integers(0, 7, b.n_intervals).astype(np.int8)
    with G.ArclengthCalculateGpu(ctx, b, normalize=True) as eng:
        d_k = ctx.array(k)
        eng.createArclengthTableSeeded(d_k)
        assert np.max(np.abs(eng.getAllCurveArclengths() - L) / L) <= 1e-6
        check_samples(eng, b, counts, ut, segt, L, cum)
        d_k.free()
    e = W.ellipse_batch(100, seed=53)
    Le, _ = O.Truth(e).lengths()
    ke = np.random.default_rng(9).integers(0, 7, 100).astype(np.int8)
    with G.ArclengthCalculateGpu(ctx, e, normalize=True) as eng:
        d_k = ctx.array(ke)
        eng.createArclengthTableSeeded(d_k)
        assert np.max(np.abs(eng.getAllCurveArclengths() - Le) / Le) <= 1e-6
        c = np.full(100, 50)
        u = eng.arclengthReverse(c, None, True)
        uet, _ = O.Truth(e).reverse(c, O.uniform_positions(c))
        bound = np.repeat(np.maximum(1e-6, 2e-7 * Le / e.b), c)
        assert np.all(np.abs(u - uet) <= bound)
        d_k.free()


def test_dip_guard_refines_wide_roots_like_plain(ctx):
    # near-cusps and near-collinear loops in degrees 2/3/5: with one root per
    # interval (raised to 4 where the speed dips), every segment wider than a
    # plain root whose node speeds dip is refined, so lengths and samples
    # stay with the plain build and the truth
    b = W.bspline_batch(40, seed=91, degrees=(2, 3, 5), dims=(2, 3), n_ctrl_range=(16, 80), cusp_frac=0.2,
                        hairpin_frac=0.5)
    counts = 8 * b.intervals_per_curve()
    tr = O.Truth(b)
    L, _ = tr.lengths()
    ut, _ = tr.reverse(counts, O.uniform_positions(counts))
    with G.ArclengthCalculateGpu(ctx, b, normalize=True) as eng:
        eng.createArclengthTable()
        Lp = eng.getAllCurveArclengths()
        up = eng.arclengthReverse(counts, None, True)
        d_k = ctx.array(np.zeros(b.n_intervals, np.int8))
        eng.createArclengthTableSeeded(d_k)
        d_k.free()
        floor = eng.seed_floor()
        Ls = eng.getAllCurveArclengths()
        us = eng.arclengthReverse(counts, None, True)
        deep = eng.tables()[2]
        st = eng.stats()
    assert st["overflow"] == 0
    assert np.any(floor == 2) and set(np.unique(floor)) <= {0, 2}
    # the dips show up as refined roots of the raised intervals
    assert np.count_nonzero(deep[floor == 2]) >= 0.5 * np.count_nonzero(floor == 2)
    assert np.max(np.abs(Ls - L) / L) <= 1e-6
    assert np.max(np.abs(Ls - Lp) / Lp) <= 2e-7
    assert np.max(np.abs(us - ut)) <= 1e-6
    assert np.max(np.abs(us - up)) <= 1e-6
