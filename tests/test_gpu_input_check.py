"""Bad batch data is rejected at table creation with GCV_ERR_DATA and a
message naming the first bad curve: non-finite knots, control points or
semi-axes, and knots that decrease within a curve. Well-formed batches with
repeated knots (zero-width spans) pass. The reference computes on such data
without checking (gcurval/detail/cuda/arclength_calculate.h:84-118 copies the
plug-in's arrays as given); here the check runs once per table, on the
device, before anything is built.
"""
import numpy as np
import pytest

import gcurval_amd as G
from gcurval_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    with G.Context(0) as c:
        yield c


def _batch(n=5, seed=3):
    return W.bspline_batch(n, seed=seed, degrees=(3,), dims=(3,), n_ctrl_range=(8, 40))


def _copy(b, knots=None, ctrl=None):
    return G.BSplineBatch(b.degree, b.dim, b.knot_off, b.ctrl_off, b.knots if knots is None else knots,
                          b.ctrl if ctrl is None else ctrl)


def _expect_data_error(ctx, batch, curve, what):
    print(f"[input_check] curve {curve}: {what}", flush=True)
    with pytest.raises(G.GcvError) as e:
        G.ArclengthCalculateGpu(ctx, batch, normalize=True)
    assert e.value.code == 6, str(e.value)
    assert f"curve {curve}:" in str(e.value) and what in str(e.value), str(e.value)


def test_good_batch_with_repeated_knots_builds(ctx):
    b = _batch()
    k = b.knots.copy()
    # a double interior knot in curve 2 (a zero-width span is legal)
    j = int(b.knot_off[2]) + 6
    k[j + 1] = k[j]
    with G.ArclengthCalculateGpu(ctx, _copy(b, knots=k), normalize=True) as eng:
        eng.createArclengthTable()
        L = eng.getAllCurveArclengths()
        assert np.all(np.isfinite(L)) and np.all(L > 0)


@pytest.mark.parametrize("val", [np.nan, np.inf, -np.inf])
def test_non_finite_control_point(ctx, val):
    b = _batch()
    ctrl = b.ctrl.copy()
    ctrl[int(b.ctrl_off[3]) + 2, 1] = val
    _expect_data_error(ctx, _copy(b, ctrl=ctrl), 3, "control point is not finite")


def test_non_finite_knot(ctx):
    b = _batch()
    k = b.knots.copy()
    k[int(b.knot_off[1]) + 5] = np.nan
    _expect_data_error(ctx, _copy(b, knots=k), 1, "knot is not finite")


def test_decreasing_knots(ctx):
    b = _batch()
    k = b.knots.copy()
    j = int(b.knot_off[4]) + 7
    k[j], k[j + 1] = k[j + 1] + 0.01, k[j]
    _expect_data_error(ctx, _copy(b, knots=k), 4, "knots decrease")


def test_first_bad_curve_is_reported(ctx):
    b = _batch(8)
    ctrl = b.ctrl.copy()
    for c in (6, 2, 5):
        ctrl[int(b.ctrl_off[c]), 0] = np.nan
    _expect_data_error(ctx, _copy(b, ctrl=ctrl), 2, "control point is not finite")


def test_unused_padding_coordinate_is_not_checked(ctx):
    # 2-D curves in a stride-3 batch: the third coordinate is padding
    b = W.bspline_batch(4, seed=9, degrees=(2,), dims=(2,), n_ctrl_range=(8, 30))
    ctrl3 = np.concatenate([b.ctrl[:, :2], np.full((b.ctrl.shape[0], 1), np.nan)], axis=1)
    with G.ArclengthCalculateGpu(ctx, _copy(b, ctrl=ctrl3), normalize=True) as eng:
        eng.createArclengthTable()
        assert np.all(np.isfinite(eng.getAllCurveArclengths()))


@pytest.mark.parametrize("val", [np.nan, np.inf])
def test_non_finite_semi_axis(ctx, val):
    a = np.array([1.0, 2.0, 3.0])
    bb = np.array([0.5, val, 1.0])
    print("[input_check] ellipse", val, flush=True)
    with pytest.raises(G.GcvError) as e:
        G.ArclengthCalculateGpu(ctx, G.EllipseBatch(a, bb), normalize=True)
    assert e.value.code == 6 and "curve 1:" in str(e.value) and "semi-axis" in str(e.value), str(e.value)
