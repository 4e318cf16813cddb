"""GPU tests of the header-only template surface (include/gcurval/gcurval.h)
through the C++ driver tests/cpp/template_surface, which instantiates the
engine on plug-ins written the way gcurval's samples write theirs: an ellipse
family (every length rule) and a cubic 3-D B-spline family with a curve
locator, a Cox-de Boor derivative and a per-thread scratch allocator.
The plug-ins are instantiated in FP32 (64 roots per interval, tolerance 1e-6
absolute on raw coordinates, depth 14: lengths within 1e-6 relative and
sample parameters within 2e-6 of the long-double truth for the
Gauss-Legendre rule) and in FP64 (tolerance 1e-10, depth 23: 1e-10 and 1e-9).
"""
import os
import subprocess

import numpy as np
import pytest

from gcurval_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "template_surface")


def _run(tmp_path, family, rule, payload, tol=None, prec="f32"):
    if not os.path.exists(EXE):
        pytest.fail("template-surface driver not built (python -m gcurval_amd.build)")
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    fin.write_bytes(payload)
    cmd = [EXE, prec, family, rule, str(fin), str(fout)] + ([repr(tol)] if tol else [])
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return np.fromfile(fout, dtype=np.float64)


def _ellipse_payload(a, b, counts, pos=None):
    parts = [np.int32(a.size).tobytes(), a.astype(np.float64).tobytes(), b.astype(np.float64).tobytes(),
             counts.astype(np.int32).tobytes(), np.int32(0 if pos is None else 1).tobytes()]
    if pos is not None:
        parts.append(pos.astype(np.float64).tobytes())
    return b"".join(parts)


@pytest.fixture(scope="module")
def ellipses():
    e = W.ellipse_batch(120, seed=71)
    a = e.a.astype(np.float32).astype(np.float64)
    b = e.b.astype(np.float32).astype(np.float64)
    counts = np.full(a.size, 40, np.int64)
    counts[0], counts[1] = 1, 0  # one sample (position 0) and none
    return e, a, b, counts


def _split(out, n, n_total, dim):
    L = out[:n]
    u = out[n:n + n_total]
    pts = out[n + n_total:].reshape(n_total, dim)
    return L, u, pts


def test_ellipse_gauss_legendre(tmp_path, ellipses):
    e, a, b, counts = ellipses
    n_total = int(counts.sum())
    out = _run(tmp_path, "ellipse", "gl", _ellipse_payload(a, b, counts))
    L, u, pts = _split(out, a.size, n_total, 2)
    tr = O.Truth(e)
    Lt, _ = tr.lengths()
    assert np.max(np.abs(L - Lt) / Lt) <= 1e-6
    ut, _ = tr.reverse(counts, O.uniform_positions(counts))
    assert np.max(np.abs(u - ut)) <= 2e-6
    # points are the curve at the returned parameters
    ar, br = np.repeat(a, counts), np.repeat(b, counts)
    ref = np.stack([ar * np.cos(u), br * np.sin(u)], axis=1)
    assert np.max(np.abs(pts - ref)) <= 2e-5 * np.max(np.abs(ref))
    # samples are non-decreasing within each curve and start at 0
    off = O.offsets(counts)
    for c in range(a.size):
        s = u[off[c]:off[c + 1]]
        if s.size:
            assert s[0] == 0.0 and np.all(np.diff(s) >= -1e-6)


def test_ellipse_relative_positions(tmp_path, ellipses):
    e, a, b, counts = ellipses
    rng = np.random.default_rng(5)
    pos = np.concatenate([np.sort(rng.uniform(0, 1, int(c))) for c in counts])
    out = _run(tmp_path, "ellipse", "gl", _ellipse_payload(a, b, counts, pos.astype(np.float32).astype(np.float64)))
    n_total = int(counts.sum())
    _, u, _ = _split(out, a.size, n_total, 2)
    ut, _ = O.Truth(e).reverse(counts, pos.astype(np.float32).astype(np.float64))
    assert np.max(np.abs(u - ut)) <= 2e-6


def test_ellipse_fp64(tmp_path, ellipses):
    e, a, b, counts = ellipses
    out = _run(tmp_path, "ellipse", "gl", _ellipse_payload(a, b, counts), tol=1e-10, prec="f64")
    n_total = int(counts.sum())
    L, u, pts = _split(out, a.size, n_total, 2)
    tr = O.Truth(e)
    Lt, _ = tr.lengths()
    assert np.max(np.abs(L - Lt) / Lt) <= 1e-10
    ut, _ = tr.reverse(counts, O.uniform_positions(counts))
    assert np.max(np.abs(u - ut)) <= 1e-9
    ar, br = np.repeat(a, counts), np.repeat(b, counts)
    assert np.max(np.abs(pts - np.stack([ar * np.cos(u), br * np.sin(u)], axis=1))) <= 1e-12 * np.max(a)


@pytest.mark.parametrize("prec,tol", [("f32", 1e-3), ("f64", 1e-8)])
@pytest.mark.parametrize("rule", ["simpson", "vincent", "polyline"])
def test_ellipse_point_rules(tmp_path, ellipses, rule, prec, tol):
    # Point-based rules refined by the same adaptive test converge to the arc
    # length within the per-interval tolerance. Their chords cancel: in FP32
    # the tolerance must stay above the rounding noise of point differences
    # (1e-3 absolute per interval on curves of length 3..55), or refinement
    # reaches the depth cap on noise and the noise biases every chord upward.
    e, a, b, counts = ellipses
    out = _run(tmp_path, "ellipse", rule, _ellipse_payload(a, b, counts), tol=tol, prec=prec)
    n_total = int(counts.sum())
    L, u, _ = _split(out, a.size, n_total, 2)
    Lt, _ = O.Truth(e).lengths()
    # one knot interval per curve: the adaptive test bounds its error by ~tol
    assert np.max(np.abs(L - Lt)) <= 2 * tol
    # a sample's arc-length error is below that of its curve, so its parameter
    # error is below 2 tol / (smallest speed b) over the curve
    ut, _ = O.Truth(e).reverse(counts, O.uniform_positions(counts))
    bound = np.repeat(2 * tol / np.minimum(a, b), counts)
    assert np.all(np.abs(u - ut) <= bound)


def test_ellipse_parameter_rule(tmp_path, ellipses):
    # the parameter-width rule measures 2 pi for every curve and samples uniformly in t
    e, a, b, counts = ellipses
    out = _run(tmp_path, "ellipse", "param", _ellipse_payload(a, b, counts))
    n_total = int(counts.sum())
    L, u, _ = _split(out, a.size, n_total, 2)
    assert np.max(np.abs(L - 2 * np.pi)) <= 2e-6 * 2 * np.pi
    ref = 2 * np.pi * O.uniform_positions(counts)
    assert np.max(np.abs(u - ref)) <= 4e-6


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_bspline_with_locator_and_scratch(tmp_path, prec):
    bs = W.bspline_batch(30, seed=72, degrees=(3,), dims=(3,), n_ctrl_range=(8, 80))
    # float32-representable data, knots offsets in the reference's layout
    n = bs.n_curve
    ko = bs.knot_off.astype(np.int32)
    knots = bs.knots.astype(np.float32).astype(np.float64)
    ctrl = bs.ctrl.reshape(-1, bs.ctrl_stride)[:, :3].astype(np.float32).astype(np.float64)
    counts = 16 * bs.intervals_per_curve()
    payload = b"".join([np.int32(n).tobytes(), ko.tobytes(), knots.tobytes(), ctrl.tobytes(),
                        counts.astype(np.int32).tobytes()])
    f64 = prec == "f64"
    out = _run(tmp_path, "bspline", "gl", payload, tol=1e-10 if f64 else None, prec=prec)
    n_total = int(counts.sum())
    L, u, pts = _split(out, n, n_total, 3)
    from gcurval_amd import BSplineBatch
    rb = BSplineBatch(bs.degree, np.full(n, 3, np.int32), bs.knot_off, bs.ctrl_off, knots, ctrl)
    tr = O.Truth(rb)
    Lt, _ = tr.lengths()
    assert np.max(np.abs(L - Lt) / Lt) <= (1e-10 if f64 else 1e-6)
    ut, _ = tr.reverse(counts, O.uniform_positions(counts))
    assert np.max(np.abs(u - ut)) <= (1e-9 if f64 else 2e-6)
    # points against scipy's B-spline at the returned parameters
    from scipy.interpolate import BSpline
    off = O.offsets(counts)
    worst = 0.0
    for c in range(n):
        t = knots[bs.knot_off[c]:bs.knot_off[c + 1]]
        P = ctrl[bs.ctrl_off[c]:bs.ctrl_off[c + 1]]
        ref = BSpline(t, P, 3)(u[off[c]:off[c + 1]])
        worst = max(worst, float(np.max(np.abs(pts[off[c]:off[c + 1]] - ref)) / np.max(np.abs(P))))
    assert worst <= (1e-12 if f64 else 1e-5)
