"""The cumulative scan of interval lengths against closed forms and the
oracle, on one curve of more than a million intervals and on many short
ones, and its time against how the intervals split into curves. Batches of
curves of 64..8,192 intervals on average scan a curve per wave
(k_curve_scan); the others take one pass over all intervals with decoupled
look-back (k_seg_scan, curves as segments); the oracle test runs both on one
batch.

The per-curve scan it replaces (gcurval/detail/cuda/kernel/generic.inl.cuh:
117-137 scans each curve's intervals) costs one serial chain per curve; here
the time follows the interval count alone.
"""
import numpy as np
import pytest

import gcurval_amd as G
from gcurval_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    with G.Context(0) as c:
        yield c


def line_batch(n_curve, n_int, step=(1.0, 2.0, 0.5)):
    """Clamped uniform cubic curves whose control points lie evenly on a line:
    each curve is that line segment, of length |P_last - P_first|."""
    p = 3
    n_ctrl = n_int + p
    inner = np.arange(1, n_int) / n_int
    knots1 = np.concatenate([np.zeros(p + 1), inner, np.ones(p + 1)])
    t = np.arange(n_ctrl, dtype=np.float64)[:, None]
    ctrl1 = t * np.asarray(step)[None, :]
    return G.BSplineBatch(np.full(n_curve, p, np.int32), np.full(n_curve, 3, np.int32),
                          np.arange(n_curve + 1, dtype=np.int64) * knots1.size,
                          np.arange(n_curve + 1, dtype=np.int64) * n_ctrl,
                          np.tile(knots1, n_curve), np.tile(ctrl1, (n_curve, 1))), (n_ctrl - 1) * np.linalg.norm(step)


def test_one_curve_of_1_2m_intervals(ctx):
    n_int = 1_200_000
    b, exact = line_batch(1, n_int)
    print(f"[scan] one curve, {n_int} intervals", flush=True)
    with G.ArclengthCalculateGpu(ctx, b, precision=G.GCV_FP64, normalize=True) as eng:
        eng.createArclengthTable()
        L = eng.getAllCurveArclengths()
        cum, _, _, cl, _ = eng.tables()
    assert abs(L[0] - exact) <= 1e-9 * exact, (L[0], exact)
    assert cum[-1] == cl[0] == L[0]
    d = np.diff(cum)
    assert np.all(d > 0)
    # interior intervals of an evenly spaced clamped cubic all have the same length
    interior = d[10:-10]
    assert np.max(np.abs(interior - interior.mean())) <= 1e-9 * interior.mean()


@pytest.mark.parametrize("mode", [1, 2])
def test_many_curves_against_the_oracle(ctx, mode):
    # both scan kernels on the same batch (a wave per curve; the look-back pass)
    b = W.bspline_batch(300, seed=71, degrees=(2, 3, 5), dims=(2, 3), n_ctrl_range=(8, 3000), hairpin_frac=0.2)
    print(f"[scan] mode {mode}: 300 curves, {b.n_intervals} intervals", flush=True)
    ref = O.RefGpuPlain(b, O.opts(1, 3, normalize=True))
    ref.build()
    Lr, cumr, _, _ = ref.tables()
    with G.ArclengthCalculateGpu(ctx, b, precision=G.GCV_FP64, normalize=True) as eng:
        eng.set_scan(mode)
        eng.createArclengthTable()
        cum, _, _, Lg, io = eng.tables()
        eng.createArclengthTable()
        cum2 = eng.tables()[0]
    np.testing.assert_array_equal(cum, cum2)  # the same rounding from build to build
    assert np.max(np.abs(Lg - Lr) / Lr) <= 1e-10
    scale = np.repeat(Lr, np.diff(io))
    assert np.max(np.abs(cum - cumr) / scale) <= 1e-10
    ends = io[1:] - 1
    np.testing.assert_array_equal(cum[ends], Lg)


def test_time_follows_the_interval_count(ctx):
    # 1.2M intervals as 1 curve, 1,200 curves and 120,000 curves
    times = {}
    for n_curve, n_int in ((1, 1_200_000), (1200, 1000), (120_000, 10)):
        b, _ = line_batch(n_curve, n_int)
        with G.ArclengthCalculateGpu(ctx, b, normalize=True) as eng:
            eng.createArclengthTable()
            ctx.profile(True)
            best = np.inf
            for _ in range(3):
                ctx.profile_reset()
                eng.createArclengthTable()
                ctx.sync()
                t1, n1 = ctx.profile_query("k_seg_scan")
                t2, n2 = ctx.profile_query("k_curve_scan")
                t, n = t1 + t2, n1 + n2
                assert n == 1
                best = min(best, t)
            ctx.profile(False)
        times[n_curve] = best
        print(f"[scan] {n_curve} curves x {n_int} intervals: {best * 1e3:.1f} us", flush=True)
    assert max(times.values()) <= 3.0 * min(times.values()), times
