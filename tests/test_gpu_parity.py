"""GPU parity of the gfx950 table build and sampler against the oracle.

Bars (stated per test): arc length <= 1e-6 relative in FP32 mode and <= 1e-10
in FP64 mode against the long-double truth and against the serial
restatement of the reference's GPU path; per-interval sample counts equal to
the truth's except inside the tie band |s - S_i| <= 2 tol L_c; sample
parameters within the stated absolute bound of the truth. Every case runs
through the C ABI (gcurval_amd -> libgcurval_amd.so).
"""
import math

import numpy as np
import pytest

import gcurval_amd as G
from gcurval_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def build(ctx, batch, **kw):
    eng = G.ArclengthCalculateGpu(ctx, batch, **kw)
    eng.createArclengthTable()
    return eng


def count_mismatches(batch, counts, pos, seg_gpu, seg_truth, L_truth, cum_truth, tol):
    """Samples whose interval differs from the truth's, outside the tie band."""
    off = O.offsets(counts)
    bad = 0
    ties = 0
    ipc = np.ones(batch.n_curve, np.int64) if batch.kind == "ellipse" else batch.intervals_per_curve()
    ioff = np.zeros(batch.n_curve + 1, np.int64)
    np.cumsum(ipc, out=ioff[1:])
    for c in range(batch.n_curve):
        s = pos[off[c]:off[c + 1]] * L_truth[c]
        bounds = cum_truth[ioff[c]:ioff[c + 1]]
        near = np.min(np.abs(s[:, None] - bounds[None, :]), axis=1) <= 2 * tol * L_truth[c]
        diff = seg_gpu[off[c]:off[c + 1]] != seg_truth[off[c]:off[c + 1]]
        bad += int(np.sum(diff & ~near))
        ties += int(np.sum(diff & near))
    return bad, ties


# ---------------------------------------------------------------------------
# C0: one ellipse a=3, b=1, FP64, tolerance 1e-8 (as float), GL-5 and GL-3
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("gl", [5, 3])
def test_c0_ellipse_fp64(ctx, gl):
    tol = float(np.float32(1e-8))
    b = W.ellipse_c0()
    with build(ctx, b, precision=G.GCV_FP64, gl_order=gl, tolerance=tol) as eng:
        L = eng.getAllCurveArclengths()[0]
        assert abs(L - W.ELLIPSE_3_1_LENGTH) / W.ELLIPSE_3_1_LENGTH <= 1e-10
        ref = O.RefGpuPlain(b, O.opts(1, gl, tol=tol))
        ref.build()
        assert abs(L - ref.tables()[0][0]) / L <= 1e-10
        n = 1000
        for closed in (False, True):
            pos = O.uniform_positions([n], closed=closed)
            u = eng.arclengthReverse(np.array([n]), None, True, uniform_closed=closed)
            ut, _ = O.Truth(b).reverse([n], pos)
            assert np.max(np.abs(u - ut)) <= 1e-9
        assert eng.stats()["overflow"] == 0


def test_unit_circle_known_answer(ctx):
    b = G.EllipseBatch([1.0], [1.0])
    n = 1000
    for prec, atol in ((G.GCV_FP64, 1e-12), (G.GCV_FP32, 1e-6)):
        with build(ctx, b, precision=prec) as eng:
            assert abs(eng.getAllCurveArclengths()[0] - 2 * math.pi) <= (1e-12 if prec else 1e-6)
            u = eng.arclengthReverse(np.array([n]), None, True, uniform_closed=False)
            np.testing.assert_allclose(u, 2 * math.pi * np.arange(n) / n, rtol=0, atol=atol)


# ---------------------------------------------------------------------------
# C1 distribution (small): FP32 quadrature, FP64 scan, normalised curves
# ---------------------------------------------------------------------------
def test_ellipse_batch_fp32(ctx):
    b = W.ellipse_batch(400, seed=21)
    counts = W.ellipse_sample_counts(400, seed=22, lo=16, hi=256)
    tr = O.Truth(b)
    Lt, cumt = tr.lengths()
    with build(ctx, b, normalize=True) as eng:
        Lg = eng.getAllCurveArclengths()
        assert np.max(np.abs(Lg - Lt) / Lt) <= 1e-6
        ref = O.RefGpuPlain(b, O.opts(0, 3, normalize=True))
        ref.build()
        assert np.max(np.abs(Lg - ref.tables()[0]) / Lt) <= 1e-6
        pos = O.uniform_positions(counts)
        u = eng.arclengthReverse(counts, None, True)
        ut, _ = tr.reverse(counts, pos)
        # t-error bound: arc-length error (1e-6 L) over the slowest speed b
        off = O.offsets(counts)
        bound = np.repeat(np.maximum(1e-6, 2e-7 * Lt / b.b), counts)
        assert np.all(np.abs(u - ut) <= bound)
        assert eng.stats()["overflow"] == 0


@pytest.mark.parametrize("prec", [G.GCV_FP32, G.GCV_FP64])
def test_tall_ellipses_b_above_a(ctx, prec):
    """b > a: the speed's minimum moves to the quarter points (the squared
    cosine branch of the evaluator). Lengths equal those of the swapped
    ellipses (same curve, rotated); samples against the truth."""
    wide = W.ellipse_batch(120, seed=23)
    tall = G.EllipseBatch(wide.b, wide.a)
    counts = W.ellipse_sample_counts(120, seed=24, lo=16, hi=256)
    tr = O.Truth(tall)
    Lt, _ = tr.lengths()
    len_bar, u_floor = (1e-6, 1e-6) if prec == G.GCV_FP32 else (1e-10, 1e-9)
    with build(ctx, tall, precision=prec, normalize=True) as eng, build(ctx, wide, precision=prec,
                                                                            normalize=True) as ew:
        Lg = eng.getAllCurveArclengths()
        assert np.max(np.abs(Lg - Lt) / Lt) <= len_bar
        assert np.max(np.abs(Lg - ew.getAllCurveArclengths()) / Lt) <= 2 * len_bar
        pos = O.uniform_positions(counts)
        u = eng.arclengthReverse(counts, None, True)
        ut, _ = tr.reverse(counts, pos)
        bound = np.repeat(np.maximum(u_floor, 2 * len_bar / 10 * Lt / np.minimum(tall.a, tall.b)), counts)
        assert np.all(np.abs(u - ut) <= bound), float(np.max(np.abs(u - ut) - bound))
        assert eng.stats()["overflow"] == 0


@pytest.mark.parametrize("closed", [True, False])
def test_uniform_runs_agree_with_explicit_positions(ctx, closed):
    """Uniform positions take the persistent run sampler (per-curve data in
    LDS; on curves with axis ratio >= 0.2 a quintic start followed directly
    by the FP64 step); the same positions passed explicitly take the
    per-sample path (cubic start, FP32 Newton, FP64 step). Both solve the
    same FP64 residual on the same table, so they agree far below the
    table's own error: 1e-9 rad (the quintic start's worst case after the
    FP64 step is 6e-11 of a root's width, 6e-12 rad). Ragged counts (1 to
    300) make tiles that span several curves."""
    b = W.ellipse_batch(300, seed=61)
    counts = W.ellipse_sample_counts(300, seed=62, lo=1, hi=300)
    with build(ctx, b, normalize=True) as eng:
        u_run = eng.arclengthReverse(counts, None, True, uniform_closed=closed)
        pos = O.uniform_positions(counts, closed=closed)
        u_pos = eng.arclengthReverse(counts, pos, True)
        assert np.max(np.abs(u_run - u_pos)) <= 1e-9, float(np.max(np.abs(u_run - u_pos)))
        # and both against the truth, as the other ellipse tests
        ut, _ = O.Truth(b).reverse(counts, pos)
        Lt, _ = O.Truth(b).lengths()
        bound = np.repeat(np.maximum(1e-6, 2e-7 * Lt / np.minimum(b.a, b.b)), counts)
        assert np.all(np.abs(u_run - ut) <= bound)


def test_reference_sample1_dataset_lengths(ctx):
    b = W.reference_sample1_ellipses(1000)
    ref = O.RefGpuPlain(b, O.opts(0, 3))
    ref.build()
    with build(ctx, b) as eng:
        Lg = eng.getAllCurveArclengths()
        Lt, _ = O.Truth(b).lengths()
        assert np.max(np.abs(Lg - Lt) / Lt) <= 1e-6
        assert np.max(np.abs(Lg - ref.tables()[0]) / Lt) <= 1e-6


# ---------------------------------------------------------------------------
# B-splines
# ---------------------------------------------------------------------------
def bspline_case(ctx, batch, prec=G.GCV_FP32, gl=3, per_interval=32, len_bar=1e-6, u_bar=1e-6, tol=1e-6):
    tr = O.Truth(batch)
    Lt, cumt = tr.lengths()
    counts = per_interval * batch.intervals_per_curve()
    with build(ctx, batch, precision=prec, gl_order=gl, normalize=True) as eng:
        Lg = eng.getAllCurveArclengths()
        rel = np.max(np.abs(Lg - Lt) / Lt)
        assert rel <= len_bar, rel
        u, seg = eng.arclengthReverse(counts, None, True, return_segments=True)
        pos = O.uniform_positions(counts)
        ut, segt = tr.reverse(counts, pos)
        bad, ties = count_mismatches(batch, counts, pos, seg, segt, Lt, cumt, tol)
        assert bad == 0, (bad, ties)
        err = np.abs(u - ut)
        assert np.max(err) <= u_bar, np.max(err)
        assert eng.stats()["overflow"] == 0
        return rel, float(np.max(err)), ties


def test_bspline_cubic3d_fp32(ctx):
    b = W.bspline_batch(60, seed=31, degrees=(3,), dims=(3,), n_ctrl_range=(8, 200))
    bspline_case(ctx, b)


def test_bspline_mixed_degrees_fp32(ctx):
    b = W.bspline_batch(60, seed=32, degrees=(2, 3, 5), dims=(2, 3), n_ctrl_range=(16, 120), hairpin_frac=0.2)
    bspline_case(ctx, b)


def pick_curves(batch, ids):
    ko, co = batch.knot_off, batch.ctrl_off
    knots = np.concatenate([batch.knots[ko[c]:ko[c + 1]] for c in ids])
    ctrl = np.concatenate([batch.ctrl[co[c]:co[c + 1]] for c in ids])
    sko = np.zeros(len(ids) + 1, np.int64)
    sko[1:] = np.cumsum(ko[ids + 1] - ko[ids])
    sco = np.zeros(len(ids) + 1, np.int64)
    sco[1:] = np.cumsum(co[ids + 1] - co[ids])
    return G.BSplineBatch(batch.degree[ids], batch.dim[ids], sko, sco, knots, ctrl)


def test_sampler_converges_next_to_near_zero_speed(ctx):
    """Quadratic B-splines of the mixed-degree 50,000-curve batch whose leaves
    hold near-zero speeds: after the two fixed FP32 Newton steps some targets
    were still 2e-6 from the converged parameter, until unconverged lanes
    kept iterating. Bar: every sample within 1e-6 of the truth."""
    full = W.bspline_c4(50_000)
    b = pick_curves(full, np.array([23524, 14511, 48002, 21452, 49380, 18918, 33275]))
    tr = O.Truth(b)
    counts = np.full(b.n_curve, 4000, np.int64)
    with build(ctx, b, normalize=True) as eng:
        u = eng.arclengthReverse(counts, None, True)
    ut, _ = tr.reverse(counts, O.uniform_positions(counts))
    err = np.abs(u - ut)
    assert np.max(err) <= 1e-6, np.max(err)


@pytest.mark.parametrize("gl", [3, 5])
def test_bspline_fp64(ctx, gl):
    b = W.bspline_batch(20, seed=33, degrees=(2, 3, 5), dims=(2, 3), n_ctrl_range=(8, 60), hairpin_frac=0.3)
    bspline_case(ctx, b, prec=G.GCV_FP64, gl=gl, per_interval=8, len_bar=1e-10, u_bar=1e-9, tol=1e-14)


def test_bspline_vs_reference_gpu_plain(ctx):
    b = W.bspline_batch(40, seed=34, degrees=(3,), dims=(2,), n_ctrl_range=(8, 100))
    ref = O.RefGpuPlain(b, O.opts(0, 3, normalize=True))
    ref.build()
    Lr, cumr, _, _ = ref.tables()
    with build(ctx, b, normalize=True) as eng:
        int_cum, _, _, Lg, _ = eng.tables()
        assert np.max(np.abs(Lg - Lr) / Lr) <= 1e-6
        ioff = O.offsets(b.intervals_per_curve())
        scale = np.repeat(Lr, np.diff(ioff))
        assert np.max(np.abs(int_cum - cumr) / scale) <= 1e-6


# ---------------------------------------------------------------------------
# Options, edge cases and invariants
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("nsub", [1, 8, 64])
def test_n_subdivided_variants(ctx, nsub):
    b = W.ellipse_batch(50, seed=41)
    Lt, _ = O.Truth(b).lengths()
    with build(ctx, b, normalize=True, n_subdivided=nsub) as eng:
        assert np.max(np.abs(eng.getAllCurveArclengths() - Lt) / Lt) <= 1e-6
        counts = np.full(50, 64)
        u = eng.arclengthReverse(counts, None, True)
        assert np.all(np.diff(u.reshape(50, 64), axis=1) > 0)


def test_determinism_and_rebuild(ctx):
    b = W.bspline_batch(30, seed=42, degrees=(3,), dims=(3,), n_ctrl_range=(8, 80))
    with build(ctx, b, normalize=True) as eng:
        t1 = eng.tables()
        eng.createArclengthTable()
        t2 = eng.tables()
        for x, y in zip(t1, t2):
            np.testing.assert_array_equal(x, y)


def test_empty_batch_and_zero_samples(ctx):
    b = G.EllipseBatch(np.zeros(0), np.zeros(0))
    with build(ctx, b) as eng:
        assert eng.getAllCurveArclengths().size == 0
    b2 = W.ellipse_batch(5, seed=3)
    with build(ctx, b2) as eng:
        u = eng.arclengthReverse(np.array([0, 3, 0, 0, 2]), None, True)
        assert u.size == 5 and np.all(np.isfinite(u))


def test_empty_bspline_batch_seeded_and_learned_builds(ctx):
    """An empty B-spline batch has no evaluator classes: seeded, uniform and
    learned builds must launch nothing and leave an empty table."""
    from gcurval_amd.seed import SeedMLP
    z64, z32 = np.zeros(0, np.int64), np.zeros(0, np.int32)
    b = G.BSplineBatch(z32, z32, np.zeros(1, np.int64), np.zeros(1, np.int64), np.zeros(0), np.zeros((0, 3)))
    with G.ArclengthCalculateGpu(ctx, b, normalize=True) as eng:
        k = ctx.array(np.zeros(1, np.int8))
        eng.createArclengthTableSeeded(k)
        assert eng.getAllCurveArclengths().size == 0
        eng.createArclengthTableUniform(3)
        assert eng.getAllCurveArclengths().size == 0
        with SeedMLP.pretrained(ctx) as mlp:
            eng.createArclengthTableLearned(mlp)
            assert eng.getAllCurveArclengths().size == 0
        assert eng.arclengthReverse(z64, None, True).size == 0
        k.free()


@pytest.mark.parametrize("closed", [True, False])
def test_ragged_sample_counts_bspline(ctx, closed):
    """Per-curve counts 0, 1, 2, 3 and several hundred in one batch, so sample
    blocks of 256 start mid-curve, span many short curves and skip empty
    ones: the block interval bounds must leave every sample's interval and
    parameter equal to the unbounded search's (the truth, within 1e-6)."""
    b = W.bspline_batch(48, seed=34, degrees=(2, 3), dims=(2, 3), n_ctrl_range=(8, 300))
    rng = np.random.default_rng(35)
    counts = rng.choice([0, 1, 2, 3, 257, 700], size=b.n_curve).astype(np.int64)
    counts[:4] = [0, 1, 2, 700]
    tr = O.Truth(b)
    Lt, cumt = tr.lengths()
    with build(ctx, b, normalize=True) as eng:
        u, seg = eng.arclengthReverse(counts, None, True, return_segments=True, uniform_closed=closed)
    pos = O.uniform_positions(counts, closed=closed)
    ut, segt = tr.reverse(counts, pos)
    bad, ties = count_mismatches(b, counts, pos, seg, segt, Lt, cumt, 1e-6)
    assert bad == 0, (bad, ties)
    assert np.max(np.abs(u - ut)) <= 1e-6, np.max(np.abs(u - ut))


def test_explicit_positions_absolute_and_clamped(ctx):
    b = W.ellipse_batch(10, seed=5)
    Lt, _ = O.Truth(b).lengths()
    rng = np.random.default_rng(6)
    counts = np.full(10, 20)
    rel = rng.uniform(0, 1, 200)
    rel[::37] = 1.0
    rel[5] = 0.0
    with build(ctx, b, normalize=True) as eng:
        u_rel = eng.arclengthReverse(counts, rel, True)
        u_abs = eng.arclengthReverse(counts, rel * np.repeat(Lt, 20), False)
        ut, _ = O.Truth(b).reverse(counts, rel)
        bound = np.repeat(np.maximum(1e-6, 2e-7 * Lt / b.b), counts)
        assert np.all(np.abs(u_rel - ut) <= bound)
        assert np.all(np.abs(u_abs - ut) <= 2 * bound)
        # positions beyond the curve end clamp to the end of the last interval
        u_over = eng.arclengthReverse(np.array([1] + [0] * 9), np.array([1.5]), True)
        assert abs(u_over[0] - 2 * math.pi) <= 1e-6


def test_degenerate_intervals_have_zero_length(ctx):
    # repeated interior knot -> a zero-width knot interval
    knots = np.array([0, 0, 0, 0, 0.4, 0.4, 0.7, 1, 1, 1, 1.0])
    ctrl = np.random.default_rng(9).normal(size=(7, 3))
    b = G.BSplineBatch([3], [3], [0, knots.size], [0, 7], knots, ctrl)
    Lt, cumt = O.Truth(b).lengths()
    with build(ctx, b) as eng:
        int_cum, _, _, Lg, _ = eng.tables()
        assert abs(Lg[0] - Lt[0]) / Lt[0] <= 1e-6
        assert int_cum[1] == int_cum[0]  # interval 1 is [0.4, 0.4]
        u = eng.arclengthReverse(np.array([100]), None, True)
        assert np.all(np.isfinite(u)) and np.all(np.diff(u) >= 0)


def test_tolerance_below_float_floor_saturates_depth(ctx):
    b = W.ellipse_batch(20, seed=8)
    with build(ctx, b, tolerance=1e-13, max_depth=10) as eng:
        st = eng.stats()
        assert st["overflow"] == 0
        assert st["max_stack"] <= 64 * (10 + 1)
        # the first pass runs a small task stack (256 entries): a tile that
        # outgrows it is re-run with the full one, which alone records a
        # deeper stack
        assert (st["max_stack"] > 256) == (st["reruns"] > 0)
        Lt, _ = O.Truth(b).lengths()
        assert np.max(np.abs(eng.getAllCurveArclengths() - Lt) / Lt) <= 1e-6


@pytest.mark.parametrize("kind", ["ellipse", "bspline"])
def test_rerun_of_deep_tiles_with_the_full_stack(ctx, kind):
    # FP64 below its rounding floor: every root refines to the depth cap, so
    # each tile's task stack grows to 64 (depth + 1) entries, past the first
    # pass's 256; the tiles are re-run with the full stack (one re-run list
    # per class for the mixed B-spline batch)
    if kind == "ellipse":
        b = W.ellipse_batch(6, seed=82)
    else:
        b = W.bspline_batch(4, seed=83, degrees=(2, 3, 5), dims=(2, 3), n_ctrl_range=(5, 8))
    with build(ctx, b, precision=G.GCV_FP64, tolerance=1e-300, max_depth=6, normalize=True) as eng:
        st = eng.stats()
        assert st["overflow"] == 0
        assert 1 <= st["reruns"] <= b.n_intervals
        assert 256 < st["max_stack"] <= 64 * (6 + 1)
        Lt, _ = O.Truth(b).lengths()
        assert np.max(np.abs(eng.getAllCurveArclengths() - Lt) / Lt) <= 1e-9
        counts = 16 * np.asarray(b.intervals_per_curve()) if kind == "bspline" else np.full(b.n_curve, 64)
        u = eng.arclengthReverse(counts, None, True)
        ut, _ = O.Truth(b).reverse(counts, O.uniform_positions(counts))
        assert np.max(np.abs(u - ut)) <= 1e-9


def test_max_depth_one_refines_once(ctx):
    # depth 1: a failing root is split once and its halves are leaves
    b = W.ellipse_batch(20, seed=8)
    with build(ctx, b, max_depth=1, tolerance=1e-9) as eng:
        st = eng.stats()
        assert st["deep_roots"] > 0
        assert st["tasks"] == 2 * st["deep_roots"]
        assert st["quadratures"] == 3 * 64 * 20 + 2 * st["tasks"]
        assert st["max_stack"] <= 128


def test_pool_context_grows_and_reuses():
    # the default context is a pool: chunks of >= 256 MiB are added on demand,
    # freed blocks are reused, neighbours coalesce inside their chunk only
    MiB = 1 << 20
    with G.Context(0) as c:
        used0, cap0 = c.arena_stats()
        assert used0 == 0 and cap0 == 0
        a = c.empty(100 * MiB // 8, np.float64)
        used, cap = c.arena_stats()
        assert used == 100 * MiB and cap == 256 * MiB
        big = c.empty(300 * MiB // 8, np.float64)  # does not fit the first chunk: a second one
        used, cap = c.arena_stats()
        assert used == 400 * MiB and cap == 256 * MiB + 300 * MiB
        x = np.arange(1000, dtype=np.float64)
        small = c.array(x)
        np.testing.assert_array_equal(small.download(), x)
        a.free()
        small.free()
        b2 = c.empty(200 * MiB // 8, np.float64)  # the first chunk, coalesced again, takes it
        used, cap = c.arena_stats()
        assert used == 500 * MiB and cap == 256 * MiB + 300 * MiB
        for buf in (big, b2):
            buf.free()
        assert c.arena_stats()[0] == 0
        b = W.bspline_batch(10, seed=4, n_ctrl_range=(8, 40))
        with build(c, b, normalize=True) as eng:
            Lt, _ = O.Truth(b).lengths()
            assert np.max(np.abs(eng.getAllCurveArclengths() - Lt) / Lt) <= 1e-6


def test_arena_context(ctx):
    with G.Context(0, arena_bytes=64 << 20) as c2:
        b = W.bspline_batch(10, seed=4, n_ctrl_range=(8, 40))
        with build(c2, b, normalize=True) as eng:
            Lt, _ = O.Truth(b).lengths()
            assert np.max(np.abs(eng.getAllCurveArclengths() - Lt) / Lt) <= 1e-6
        used, cap = c2.arena_stats()
        assert cap == 64 << 20
    with G.Context(0, arena_bytes=1 << 20) as c3:
        big = W.ellipse_batch(100000, seed=1)
        with pytest.raises(G.GcvError) as ei:
            G.ArclengthCalculateGpu(c3, big)
        assert "OOM" in str(ei.value) or ei.value.code == 2
