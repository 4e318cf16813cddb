"""Full-size properties of the benchmark workloads (the batches the bench times,
at their real sizes), where a per-sample truth for everything would take the
oracle too long:
  * 1,000,000 ellipses: every length against the closed form 4a E(1 - b^2/a^2)
    (scipy.special.ellipe) to 1e-6 relative, plain and learned builds; 16
    uniform samples per curve non-decreasing, starting at 0 and ending at
    2 pi; a seeded subset of curves against the oracle truth;
  * 20,000 cubic 3-D B-splines (7.3M knot intervals): learned and plain
    lengths agree to 2e-6 on every curve, a seeded subset of 60 curves against
    the truth to 1e-6, samples non-decreasing within every curve and inside
    the curve's knot range;
  * 50,000 mixed-degree 2-D/3-D B-splines (12M knot intervals, 200M samples,
    4,000 per curve), plain and learned: the same properties, with the
    subset's samples also in the truth's knot intervals except inside the tie
    band (targets within 2 tol L of an interval boundary, counted and
    printed);
  * the first 100,000 of the 1M ellipses with their real 64..4,096 sample
    counts (97M samples), plain and learned: endpoints, monotonicity and a
    seeded subset against the truth.
"""
import math

import numpy as np
import pytest
import scipy.special as sp

import gcurval_amd as G
from gcurval_amd import workloads as W
from gcurval_amd.seed import SeedMLP
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _monotone_per_curve(u, counts):
    off = O.offsets(counts)
    d = np.diff(u)
    starts = off[1:-1] - 1  # differences across a curve boundary are not compared
    ok = d >= -1e-12
    ok[starts[(starts >= 0) & (starts < d.size)]] = True
    return bool(np.all(ok))


def test_c1_million_ellipses_lengths_and_samples(ctx):
    b = W.ellipse_batch(1_000_000, seed=413)
    L_closed = 4.0 * b.a * sp.ellipe(1.0 - (b.b / b.a) ** 2)
    counts = np.full(b.n_curve, 16, np.int64)
    mlp = SeedMLP.pretrained(ctx)
    try:
        with G.ArclengthCalculateGpu(ctx, b, normalize=True) as eng:
            for learned in (False, True):
                if learned:
                    eng.createArclengthTableLearned(mlp)
                else:
                    eng.createArclengthTable()
                eng._lengths = None
                L = eng.getAllCurveArclengths()
                assert np.max(np.abs(L - L_closed) / L_closed) <= 1e-6
                u = eng.arclengthReverse(counts, None, True)
                assert _monotone_per_curve(u, counts)
                u2 = u.reshape(-1, 16)
                assert np.all(u2[:, 0] == 0.0)
                assert np.max(np.abs(u2[:, -1] - 2 * math.pi)) <= 1e-6
                # a seeded subset against the truth (bound as for the small batch:
                # arc-length error over the slowest speed b)
                pick = np.random.default_rng(7).choice(b.n_curve, 300, replace=False)
                sub = G.EllipseBatch(b.a[pick], b.b[pick])
                cs = np.full(pick.size, 16, np.int64)
                ut, _ = O.Truth(sub).reverse(cs, O.uniform_positions(cs))
                bound = np.repeat(np.maximum(1e-6, 2e-7 * L_closed[pick] / b.b[pick]), 16)
                assert np.all(np.abs(u2[pick].ravel() - ut) <= bound)
    finally:
        mlp.close()


def test_c2_cubic_bsplines_full_size(ctx):
    b = W.bspline_c2(20_000)
    ipc = b.intervals_per_curve()
    counts = 4 * ipc
    mlp = SeedMLP.pretrained(ctx)
    try:
        with G.ArclengthCalculateGpu(ctx, b, normalize=True) as eng:
            eng.createArclengthTable()
            eng._lengths = None
            Lp = eng.getAllCurveArclengths().copy()
            up = eng.arclengthReverse(counts, None, True)
            eng.createArclengthTableLearned(mlp)
            eng._lengths = None
            Ll = eng.getAllCurveArclengths().copy()
            ul = eng.arclengthReverse(counts, None, True)
        assert np.max(np.abs(Ll - Lp) / Lp) <= 2e-6
        for u in (up, ul):
            assert _monotone_per_curve(u, counts)
        # parameters inside each curve's parameter range [t_p, t_{n_ctrl}]
        off = O.offsets(counts)
        lo = np.array([b.knots[b.knot_off[c] + b.degree[c]] for c in range(b.n_curve)])
        hi = np.array([b.knots[b.knot_off[c + 1] - b.degree[c] - 1] for c in range(b.n_curve)])
        for u in (up, ul):
            assert np.all(u >= np.repeat(lo, counts) - 1e-12) and np.all(u <= np.repeat(hi, counts) + 1e-12)
        # seeded subset against the truth
        pick = np.sort(np.random.default_rng(8).choice(b.n_curve, 60, replace=False))
        ko, co = b.knot_off, b.ctrl_off
        sko = np.zeros(pick.size + 1, np.int64)
        sko[1:] = np.cumsum(ko[pick + 1] - ko[pick])
        sco = np.zeros(pick.size + 1, np.int64)
        sco[1:] = np.cumsum(co[pick + 1] - co[pick])
        sub = G.BSplineBatch(b.degree[pick], b.dim[pick], sko, sco,
                             np.concatenate([b.knots[ko[c]:ko[c + 1]] for c in pick]),
                             np.concatenate([b.ctrl[co[c]:co[c + 1]] for c in pick]))
        tr = O.Truth(sub)
        Lt, _ = tr.lengths()
        assert np.max(np.abs(Lp[pick] - Lt) / Lt) <= 1e-6
        assert np.max(np.abs(Ll[pick] - Lt) / Lt) <= 1e-6
        ut, _ = tr.reverse(counts[pick], O.uniform_positions(counts[pick]))
        for u in (up, ul):
            us = np.concatenate([u[off[c]:off[c + 1]] for c in pick])
            assert np.max(np.abs(us - ut)) <= 1e-6
    finally:
        mlp.close()


def _pick(b, ids):
    ko, co = b.knot_off, b.ctrl_off
    sko = np.zeros(ids.size + 1, np.int64)
    sko[1:] = np.cumsum(ko[ids + 1] - ko[ids])
    sco = np.zeros(ids.size + 1, np.int64)
    sco[1:] = np.cumsum(co[ids + 1] - co[ids])
    return G.BSplineBatch(b.degree[ids], b.dim[ids], sko, sco, np.concatenate([b.knots[ko[c]:ko[c + 1]] for c in ids]),
                          np.concatenate([b.ctrl[co[c]:co[c + 1]] for c in ids]))


def _tie_band_split(sub, counts, pos, seg, segt, Lt, cumt, tol):
    """(interval mismatches outside the tie band, inside it) on a subset."""
    off = O.offsets(counts)
    ioff = O.offsets(sub.intervals_per_curve())
    bad = ties = 0
    for c in range(sub.n_curve):
        s = pos[off[c]:off[c + 1]] * Lt[c]
        near = np.min(np.abs(s[:, None] - cumt[ioff[c]:ioff[c + 1]][None, :]), axis=1) <= 2 * tol * Lt[c]
        diff = seg[off[c]:off[c + 1]] != segt[off[c]:off[c + 1]]
        bad += int(np.sum(diff & ~near))
        ties += int(np.sum(diff & near))
    return bad, ties


def test_c4_mixed_degree_bsplines_full_size(ctx):
    b = W.bspline_c4(50_000)
    counts = np.full(b.n_curve, 4000, np.int64)
    off = O.offsets(counts)
    lo = np.repeat(np.array([b.knots[b.knot_off[c] + b.degree[c]] for c in range(b.n_curve)]), counts)
    hi = np.repeat(np.array([b.knots[b.knot_off[c + 1] - b.degree[c] - 1] for c in range(b.n_curve)]), counts)
    pick = np.sort(np.random.default_rng(9).choice(b.n_curve, 60, replace=False))
    sub = _pick(b, pick)
    tr = O.Truth(sub)
    Lt, cumt = tr.lengths()
    pos = O.uniform_positions(counts[pick])
    ut, segt = tr.reverse(counts[pick], pos)
    rows = np.concatenate([np.arange(off[c], off[c + 1]) for c in pick])
    mlp = SeedMLP.pretrained(ctx)
    lengths = {}
    try:
        with G.ArclengthCalculateGpu(ctx, b, normalize=True) as eng:
            for learned in (False, True):
                print(f"[c4 full size] learned={learned}", flush=True)
                if learned:
                    eng.createArclengthTableLearned(mlp)
                else:
                    eng.createArclengthTable()
                eng._lengths = None
                L = eng.getAllCurveArclengths().copy()
                lengths[learned] = L
                assert eng.stats()["overflow"] == 0
                u, seg = eng.arclengthReverse(counts, None, True, return_segments=True)
                assert _monotone_per_curve(u, counts)
                assert np.all(u >= lo - 1e-12) and np.all(u <= hi + 1e-12)
                assert np.max(np.abs(L[pick] - Lt) / Lt) <= 1e-6
                us, ss = u[rows], seg[rows]
                del u, seg
                err = float(np.max(np.abs(us - ut)))
                assert err <= 1e-6, err
                bad, ties = _tie_band_split(sub, counts[pick], pos, ss, segt, Lt, cumt, 1e-6)
                print(f"[c4 full size] learned={learned}: subset max |u - truth| {err:.2e}, interval mismatches "
                      f"outside the tie band {bad}, inside it {ties} of {us.size}", flush=True)
                assert bad == 0, (bad, ties)
        assert np.max(np.abs(lengths[True] - lengths[False]) / lengths[False]) <= 2e-6
    finally:
        mlp.close()


def test_c1_ellipse_slice_with_real_sample_counts(ctx):
    n_all = 1_000_000
    b_all = W.ellipse_batch(n_all, seed=413)
    counts_all = W.ellipse_sample_counts(n_all, seed=414)
    n = 100_000
    b = G.EllipseBatch(b_all.a[:n], b_all.b[:n])
    counts = counts_all[:n]
    assert counts.min() >= 64 and counts.max() <= 4096
    off = O.offsets(counts)
    L_closed = 4.0 * b.a * sp.ellipe(1.0 - (b.b / b.a) ** 2)
    pick = np.sort(np.random.default_rng(10).choice(n, 200, replace=False))
    sub = G.EllipseBatch(b.a[pick], b.b[pick])
    ut, _ = O.Truth(sub).reverse(counts[pick], O.uniform_positions(counts[pick]))
    rows = np.concatenate([np.arange(off[c], off[c + 1]) for c in pick])
    bound = np.repeat(np.maximum(1e-6, 2e-7 * L_closed[pick] / b.b[pick]), counts[pick])
    mlp = SeedMLP.pretrained(ctx)
    try:
        with G.ArclengthCalculateGpu(ctx, b, normalize=True) as eng:
            for learned in (False, True):
                print(f"[c1 slice] learned={learned}, {int(off[-1])} samples", flush=True)
                if learned:
                    eng.createArclengthTableLearned(mlp)
                else:
                    eng.createArclengthTable()
                eng._lengths = None
                L = eng.getAllCurveArclengths()
                assert np.max(np.abs(L - L_closed) / L_closed) <= 1e-6
                u = eng.arclengthReverse(counts, None, True)
                assert _monotone_per_curve(u, counts)
                assert np.all(u[off[:-1]] == 0.0)
                assert np.max(np.abs(u[off[1:] - 1] - 2 * math.pi)) <= 1e-6
                assert np.all(np.abs(u[rows] - ut) <= bound)
    finally:
        mlp.close()
