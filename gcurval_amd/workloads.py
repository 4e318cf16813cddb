"""Seeded synthetic curve batches for the benchmark configurations and tests.

Every generator is deterministic in its seed. Sizes are parameters so tests
can draw small batches from the same distributions the benchmark uses.
"""
from __future__ import annotations

import math

import numpy as np

from . import BSplineBatch, EllipseBatch

# Closed form of the 3-by-1 ellipse perimeter, 4 a E(1 - b^2/a^2) (mpmath, 30 digits).
ELLIPSE_3_1_LENGTH = 13.3648932205552582301


# ---------------------------------------------------------------------------
# Reference sample-1 dataset (samples/sample1_ellipse/main.cpp:25-42):
# std::default_random_engine(413) (libstdc++ minstd_rand0) feeding
# uniform_real_distribution<float>(0.1, 2.0); a = 1, b = U * a, both divided by
# max(a, b); curve 0 forced to the unit circle.
# ---------------------------------------------------------------------------
def _minstd_rand0(seed: int):
    x = seed % 2147483647
    if x == 0:
        x = 1
    while True:
        x = (x * 16807) % 2147483647
        yield x


def reference_sample1_ellipses(n_curve: int, seed: int = 413) -> EllipseBatch:
    f32 = np.float32
    gen = _minstd_rand0(seed)
    r = f32(2147483646.0)  # max - min + 1 of minstd_rand0
    lo, hi = f32(0.1), f32(2.0)
    a = np.empty(n_curve, np.float32)
    b = np.empty(n_curve, np.float32)
    for i in range(n_curve):
        # generate_canonical<float, 24>: one draw, (x - 1) / r in float
        v = f32(f32(next(gen) - 1) / r)
        if v >= f32(1.0):
            v = np.nextafter(f32(1.0), f32(0.0))
        u = f32(f32(f32(hi - lo) * v) + lo)
        aa = f32(1.0)
        bb = f32(u * aa)
        m = max(aa, bb)
        a[i] = f32(aa / m)
        b[i] = f32(bb / m)
    if n_curve:
        a[0] = b[0] = f32(1.0)
    return EllipseBatch(a.astype(np.float64), b.astype(np.float64))


# ---------------------------------------------------------------------------
# Benchmark configurations
# ---------------------------------------------------------------------------
def ellipse_c0() -> EllipseBatch:
    """One ellipse, a = 3, b = 1, t in [0, 2 pi] as one knot interval."""
    return EllipseBatch(np.array([3.0]), np.array([1.0]))


def ellipse_batch(n_curve: int, seed: int = 413, f32_inputs: bool = True) -> EllipseBatch:
    """a ~ U[0.5, 10], b = a * U[0.01, 1] (eccentricity up to 0.99995)."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.5, 10.0, n_curve)
    b = a * rng.uniform(0.01, 1.0, n_curve)
    if f32_inputs:  # values exactly representable in float, so both precisions see the same curve
        a = a.astype(np.float32).astype(np.float64)
        b = b.astype(np.float32).astype(np.float64)
    return EllipseBatch(a, b)


def ellipse_sample_counts(n_curve: int, seed: int = 414, lo: int = 64, hi: int = 4096) -> np.ndarray:
    """Per-curve sample counts floor(exp(U[ln lo, ln hi]))."""
    rng = np.random.default_rng(seed)
    return np.floor(np.exp(rng.uniform(math.log(lo), math.log(hi), n_curve))).astype(np.int64)


def _clamped_uniform_knots(n_ctrl: int, p: int) -> np.ndarray:
    n_inner = n_ctrl - p - 1
    inner = np.arange(1, n_inner + 1, dtype=np.float64) / (n_inner + 1)
    return np.concatenate([np.zeros(p + 1), inner, np.ones(p + 1)])


def _random_walk(rng, n: int, D: int, cusp_frac: float = 0.1) -> np.ndarray:
    steps = rng.normal(0.0, 1.0, (n, D))
    small = rng.random(n) < cusp_frac
    steps[small] *= 0.01
    steps[0] = 0.0
    return np.cumsum(steps, axis=0)


def _hairpin(rng, P: np.ndarray) -> None:
    """Insert a near-collinear high-curvature loop: six control points run
    out along a direction and straight back with a 1% lateral offset."""
    n, D = P.shape
    if n < 10:
        return
    j = int(rng.integers(1, n - 7))
    d = rng.normal(size=D)
    d /= np.linalg.norm(d)
    o = rng.normal(size=D)
    o -= o.dot(d) * d
    o /= max(np.linalg.norm(o), 1e-12)
    orig = P.copy()
    base = P[j - 1].copy()
    for k, (t, side) in enumerate([(1, 0), (2, 0), (3, 0), (3, 1), (2, 1), (1, 1)]):
        P[j + k] = base + d * t + o * 0.01 * side
    if j + 6 < n:  # the rest of the walk continues from the loop's last point
        P[j + 6:] = orig[j + 6:] - orig[j + 5] + P[j + 5]


def bspline_batch(n_curve: int, seed: int = 413, degrees=(3,), dims=(3,), n_ctrl_range=(8, 2048),
                  cusp_frac: float = 0.1, hairpin_frac: float = 0.0, f32_inputs: bool = True) -> BSplineBatch:
    """Clamped uniform-knot B-splines with Gaussian random-walk control points
    (step sigma 1, `cusp_frac` of the steps scaled by 0.01), control-point
    counts log-uniform on n_ctrl_range, degree/dim drawn uniformly from the
    given tuples; `hairpin_frac` of the curves get a near-collinear loop."""
    rng = np.random.default_rng(seed)
    lo, hi = n_ctrl_range
    stride = max(dims)
    deg = rng.choice(np.asarray(degrees, np.int32), n_curve).astype(np.int32)
    dim = rng.choice(np.asarray(dims, np.int32), n_curve).astype(np.int32)
    n_ctrl = np.rint(np.exp(rng.uniform(math.log(lo), math.log(hi), n_curve))).astype(np.int64)
    n_ctrl = np.maximum(n_ctrl, deg + 1)
    knot_off = np.zeros(n_curve + 1, np.int64)
    ctrl_off = np.zeros(n_curve + 1, np.int64)
    np.cumsum(n_ctrl + deg + 1, out=knot_off[1:])
    np.cumsum(n_ctrl, out=ctrl_off[1:])
    knots = np.empty(knot_off[-1], np.float64)
    ctrl = np.zeros((ctrl_off[-1], stride), np.float64)
    hp = rng.random(n_curve) < hairpin_frac
    for c in range(n_curve):
        p, D, n = int(deg[c]), int(dim[c]), int(n_ctrl[c])
        knots[knot_off[c]:knot_off[c + 1]] = _clamped_uniform_knots(n, p)
        P = _random_walk(rng, n, D, cusp_frac)
        if hp[c]:
            _hairpin(rng, P)
        ctrl[ctrl_off[c]:ctrl_off[c + 1], :D] = P
    if f32_inputs:
        knots = knots.astype(np.float32).astype(np.float64)
        ctrl = ctrl.astype(np.float32).astype(np.float64)
    return BSplineBatch(deg, dim, knot_off, ctrl_off, knots, ctrl)


def bspline_scattered(n_curve: int, seed: int = 413, degrees=(3,), dims=(2,), n_ctrl_range=(4, 16),
                      f32_inputs: bool = True) -> BSplineBatch:
    """Clamped B-splines with few control points scattered uniformly in a box
    and random interior knots: the regime of the reference's B-spline dataset
    (dataset/BSpline: cubic, 4..12 control points at unrelated positions,
    non-uniform knots), where one knot interval spans a large, sharply turning
    part of its curve. Control-point counts uniform on n_ctrl_range."""
    rng = np.random.default_rng(seed)
    lo, hi = n_ctrl_range
    stride = max(dims)
    deg = rng.choice(np.asarray(degrees, np.int32), n_curve).astype(np.int32)
    dim = rng.choice(np.asarray(dims, np.int32), n_curve).astype(np.int32)
    n_ctrl = np.maximum(rng.integers(lo, hi + 1, n_curve), deg + 1).astype(np.int64)
    knot_off = np.zeros(n_curve + 1, np.int64)
    ctrl_off = np.zeros(n_curve + 1, np.int64)
    np.cumsum(n_ctrl + deg + 1, out=knot_off[1:])
    np.cumsum(n_ctrl, out=ctrl_off[1:])
    knots = np.empty(knot_off[-1], np.float64)
    ctrl = np.zeros((ctrl_off[-1], stride), np.float64)
    for c in range(n_curve):
        p, D, n = int(deg[c]), int(dim[c]), int(n_ctrl[c])
        inner = np.sort(rng.uniform(0.0, 1.0, n - p - 1))
        knots[knot_off[c]:knot_off[c + 1]] = np.concatenate([np.zeros(p + 1), inner, np.ones(p + 1)])
        ctrl[ctrl_off[c]:ctrl_off[c + 1], :D] = rng.uniform(0.0, 20.0, (n, D))
    if f32_inputs:
        knots = knots.astype(np.float32).astype(np.float64)
        ctrl = ctrl.astype(np.float32).astype(np.float64)
    return BSplineBatch(deg, dim, knot_off, ctrl_off, knots, ctrl)


def bspline_c2(n_curve: int = 20000, seed: int = 413) -> BSplineBatch:
    """20,000 cubic 3-D curves, 8..2048 control points, 10% near-cusp steps."""
    return bspline_batch(n_curve, seed, degrees=(3,), dims=(3,), n_ctrl_range=(8, 2048), cusp_frac=0.1)


def bspline_c4(n_curve: int = 50000, seed: int = 417) -> BSplineBatch:
    """50,000 curves, degree in {2,3,5}, dim in {2,3}, 16..1024 control
    points, 20% with near-collinear high-curvature loops."""
    return bspline_batch(n_curve, seed, degrees=(2, 3, 5), dims=(2, 3), n_ctrl_range=(16, 1024), cusp_frac=0.1,
                         hairpin_frac=0.2)
