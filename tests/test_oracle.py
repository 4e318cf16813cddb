"""CPU pins of the oracle against closed forms, library routines and the
reference's own known answer (unit circle, samples/sample1_ellipse/main.cpp:40-41).

The oracle is the checker for every GPU parity test, so each of its parts is
pinned here by something that a plausible mistake (wrong sign, dropped term,
wrong index or tolerance schedule) would fail.
"""
import json
import math
import os

import numpy as np
import pytest

from gcurval_amd import BSplineBatch, EllipseBatch
from gcurval_amd import workloads as W
from oracle import oracle as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))


def one_bspline(case):
    knots = np.array(case["knots"])
    ctrl = np.array(case["ctrl"])
    p, D = case["p"], case["D"]
    stride = 3
    c3 = np.zeros((ctrl.shape[0], stride))
    c3[:, :D] = ctrl
    return BSplineBatch([p], [D], [0, knots.size], [0, ctrl.shape[0]], knots, c3)


# ---------------------------------------------------------------------------
# Gauss-Legendre rules (gcurval/base/constants.inl.h:74-92)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_gl_table_matches_leggauss(n):
    x, w = O.gl_table(n)
    order = np.argsort(x)
    g = GOLD["gl"][str(n)]
    # the reference lists 16 significant digits: agreement to a few ulps
    np.testing.assert_allclose(x[order], g["x"], rtol=2e-15, atol=2e-16)
    np.testing.assert_allclose(w[order], g["w"], rtol=2e-15, atol=0)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_gl_rule_exact_for_degree_2n_minus_1(n):
    x, w = O.gl_table(n)
    rng = np.random.default_rng(n)
    c = rng.normal(size=2 * n)  # degree 2n-1
    quad = np.sum(w * np.polyval(c, x))
    # exact integral over [-1, 1]
    deg = np.arange(2 * n - 1, -1, -1)
    exact = np.sum(c * (1 - (-1.0) ** (deg + 1)) / (deg + 1))
    assert abs(quad - exact) < 1e-13
    # and not exact for degree 2n (the rule is exactly n points)
    assert abs(np.sum(w * x ** (2 * n)) - 2.0 / (2 * n + 1)) > 1e-6


def test_eps_schedules():
    # device table: tol / n_sub * 2^-l computed in the data type (constants.inl.h:8-34)
    assert O.eps_device(0, 1e-6, 64, 0) == float(np.float32(np.float32(1e-6) / np.float32(64)))
    assert O.eps_device(0, 1e-6, 64, 3) == float(np.float32(1e-6) / 64) / 8
    assert O.eps_device(1, 1e-14, 64, 0) == 1e-14 / 64
    assert O.eps_device(1, 1e-14, 1, 5) == 1e-14 / 32


# ---------------------------------------------------------------------------
# Ellipse: closed forms
# ---------------------------------------------------------------------------
def test_truth_ellipse_closed_form():
    cases = GOLD["ellipse"]
    b = EllipseBatch([c["a"] for c in cases], [c["b"] for c in cases])
    L, _ = O.Truth(b).lengths()
    exact = np.array([float(c["L"]) for c in cases])
    np.testing.assert_allclose(L, exact, rtol=2e-15)


def test_truth_ellipse_partial_arcs():
    for case in GOLD["ellipse_partial"]:
        t = O.Truth(EllipseBatch([case["a"]], [case["b"]]))
        s = t.partial(0, 0, case["t"])
        assert abs(s - float(case["s"])) <= 1e-14 * float(case["s"])


def test_truth_inverse_round_trip():
    b = EllipseBatch([3.0, 1.0, 10.0], [1.0, 0.02, 9.0])
    tr = O.Truth(b)
    L, _ = tr.lengths()
    counts = np.array([50, 50, 50])
    pos = O.uniform_positions(counts)
    u, _ = tr.reverse(counts, pos, True)
    off = O.offsets(counts)
    for c in range(3):
        for k in range(off[c], off[c + 1]):
            s = tr.partial(c, 0, u[k])
            assert abs(s - pos[k] * L[c]) <= 1e-13 * L[c]


@pytest.mark.parametrize("prec,gl,tol,bound", [(1, 3, 1e-14, 1e-12), (1, 5, 1e-14, 1e-13),
                                                (1, 5, float(np.float32(1e-8)), 1e-11), (0, 3, 1e-6, 3e-7)])
def test_reference_engines_ellipse_lengths(prec, gl, tol, bound):
    cases = GOLD["ellipse"]
    a = np.array([c["a"] for c in cases])
    bb = np.array([c["b"] for c in cases])
    exact = np.array([float(c["L"]) for c in cases])
    b = EllipseBatch(a, bb)
    # reference drivers feed unit-size curves (normalize) in float mode
    o = O.opts(prec, gl, tol=tol, normalize=True)
    cpu = O.RefCpu(b, o)
    cpu.build()
    gpu = O.RefGpuPlain(b, o)
    gpu.build()
    assert np.max(np.abs(cpu.lengths() - exact) / exact) <= bound
    assert np.max(np.abs(gpu.tables()[0] - exact) / exact) <= bound


def test_unit_circle_known_answer():
    # samples/sample1_ellipse/main.cpp:40-41: curve 0 is the unit circle, so
    # relative positions j/N map to t = 2 pi j / N.
    b = EllipseBatch([1.0], [1.0])
    n = 1000
    pos = O.uniform_positions([n], closed=False)
    # the reference returns leaf-half midpoints: 2 pi / 64 / 2^23 / 4 = 2.9e-9 at depth 23
    # (float: a 1.5e-6 half-leaf plus float rounding of the cumulative table)
    for prec, width in ((1, 5e-9), (0, 3e-6)):
        o = O.opts(prec, 3)
        gpu = O.RefGpuPlain(b, o)
        gpu.build()
        assert abs(gpu.tables()[0][0] - 2 * math.pi) < (1e-13 if prec else 2e-6)
        u, _ = gpu.reverse([n], pos)
        np.testing.assert_allclose(u, 2 * math.pi * np.arange(n) / n, rtol=0, atol=width)
        cpu = O.RefCpu(b, o)
        cpu.build()
        u2, _ = cpu.reverse([n], pos)
        # the CPU engine has no pre-split: its bisection stops at depth 14 (float)
        # / 23 (double), so its half-leaf midpoint is within 2 pi 2^-(depth+2)
        cpu_width = 2 * math.pi * 2.0 ** -((23 if prec else 14) + 2) * 1.01
        np.testing.assert_allclose(u2, 2 * math.pi * np.arange(n) / n, rtol=0, atol=cpu_width)


def test_scale_equivariance_and_normalize():
    b1 = EllipseBatch([3.0], [1.0])
    b2 = EllipseBatch([30.0], [10.0])
    L1 = O.Truth(b1).lengths()[0][0]
    L2 = O.Truth(b2).lengths()[0][0]
    assert abs(L2 - 10 * L1) <= 1e-14 * L2
    o = O.opts(0, 3, normalize=True)
    g1, g2 = O.RefGpuPlain(b1, o), O.RefGpuPlain(b2, o)
    g1.build()
    g2.build()
    # normalised evaluation sees the same unit-size curve, so trees are identical
    assert g1.n_tasks() == g2.n_tasks()
    assert abs(g2.tables()[0][0] - 10 * g1.tables()[0][0]) <= 1e-6 * g2.tables()[0][0]


# ---------------------------------------------------------------------------
# B-splines: scipy derivatives, mpmath lengths, special cases
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("k", range(6))
def test_bspline_derivatives_match_scipy(k):
    case = GOLD["bspline"][k]
    b = one_bspline(case)
    ob = O.Batch(b)
    p = case["p"]
    knots = np.array(case["knots"])
    for u, d in zip(case["u"], case["deriv"]):
        span = int(np.searchsorted(knots, u, side="right") - 1)
        seg = span - p
        ref = O.deriv(ob, 0, seg, u)[:case["D"]]
        tru = O.deriv(ob, 0, seg, u, truth=True)[:case["D"]]
        np.testing.assert_allclose(ref, d, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(tru, d, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("k", range(6))
def test_bspline_lengths_match_mpmath(k):
    case = GOLD["bspline"][k]
    b = one_bspline(case)
    L, _ = O.Truth(b).lengths()
    assert abs(L[0] - float(case["length"])) <= 1e-14 * float(case["length"])
    o = O.opts(1, 5, tol=1e-12)
    g = O.RefGpuPlain(b, o)
    g.build()
    assert abs(g.tables()[0][0] - float(case["length"])) <= 1e-10 * float(case["length"])


def test_degree1_is_polyline_and_collinear_is_chord():
    rng = np.random.default_rng(5)
    P = rng.normal(size=(7, 3))
    knots = np.r_[0, 0, np.sort(rng.uniform(0, 1, 5)), 1, 1].astype(float)  # n_knots = n_ctrl + p + 1
    b = BSplineBatch([1], [3], [0, knots.size], [0, 7], knots, P)
    poly = np.sum(np.linalg.norm(np.diff(P, axis=0), axis=1))
    assert abs(O.Truth(b).lengths()[0][0] - poly) <= 1e-14 * poly
    c = O.RefCpu(b, O.opts(1, 3))
    c.build()
    assert abs(c.lengths()[0] - poly) <= 1e-12 * poly
    # cubic with collinear, monotone control points: length is the chord
    Q = np.zeros((8, 3))
    Q[:, 0] = np.cumsum(rng.uniform(0.1, 1.0, 8))
    kn = np.r_[np.zeros(4), [0.2, 0.4, 0.6, 0.8], np.ones(4)]
    b3 = BSplineBatch([3], [3], [0, kn.size], [0, 8], kn, Q)
    chord = Q[-1, 0] - Q[0, 0]
    assert abs(O.Truth(b3).lengths()[0][0] - chord) <= 1e-14 * chord


def test_truth_interval_cumulative_is_partial_at_interval_ends():
    b = W.bspline_batch(3, seed=3, n_ctrl_range=(8, 20))
    tr = O.Truth(b)
    L, cum = tr.lengths()
    ipc = b.intervals_per_curve()
    g = 0
    for c in range(3):
        for s in range(ipc[c]):
            k0 = b.knot_off[c]
            p = b.degree[c]
            u1 = b.knots[k0 + s + p + 1]
            assert abs(tr.partial(c, s, u1) - cum[g]) <= 1e-13 * L[c]
            g += 1
        assert cum[g - 1] == L[c]


# ---------------------------------------------------------------------------
# Tree logic of the GPU-plain restatement vs an independent brute-force
# recursion (float64, same GL-3 rule, same device epsilon schedule)
# ---------------------------------------------------------------------------
def _brute_gpuplain_length(a, b, n_sub, tol, max_depth):
    xs, ws = O.gl_table(3)

    def speed(t):
        return math.hypot(-a * math.sin(t), b * math.cos(t))

    def gl(x0, x1):
        k, m = (x1 - x0) / 2, (x1 + x0) / 2
        return k * sum(w * speed(k * x + m) for x, w in zip(xs, ws))

    def node(x0, x1, parent, level):
        mid = (x0 + x1) / 2
        L, R = gl(x0, mid), gl(mid, x1)
        if level < max_depth and abs(L + R - parent) > tol / n_sub * 2.0 ** -level:
            return node(x0, mid, L, level + 1) + node(mid, x1, R, level + 1)
        return L + R

    e = 2 * math.pi
    d = e / n_sub
    vals = [node(d * k, d * (k + 1), gl(d * k, d * (k + 1)), 0) for k in range(n_sub)]
    return vals


@pytest.mark.parametrize("max_depth", [0, 1, 2, 3])
def test_gpuplain_tree_vs_brute_force(max_depth):
    a, bb, n_sub, tol = 1.0, 0.05, 8, 1e-3
    brute = _brute_gpuplain_length(a, bb, n_sub, tol, max_depth)
    g = O.RefGpuPlain(EllipseBatch([a], [bb]), O.opts(1, 3, n_sub=n_sub, max_depth=max_depth, tol=tol))
    g.build()
    _, _, prefix, depth = g.tables()
    np.testing.assert_allclose(np.diff(np.r_[0, prefix]), brute, rtol=1e-13, atol=1e-15)
    assert depth.max() <= max_depth


def test_oracle_samples_agree_with_truth_fp64():
    b = W.bspline_batch(4, seed=9, n_ctrl_range=(8, 30))
    counts = 16 * b.intervals_per_curve()
    pos = O.uniform_positions(counts)
    ut, segt = O.Truth(b).reverse(counts, pos)
    o = O.opts(1, 5, normalize=True)
    g = O.RefGpuPlain(b, o)
    g.build()
    ug, segg = g.reverse(counts, pos)
    assert np.max(np.abs(ug - ut)) < 1e-9
    c = O.RefCpu(b, o)
    c.build()
    uc, _ = c.reverse(counts, pos)
    assert np.max(np.abs(uc - ut)) < 1e-9


# ---------------------------------------------------------------------------
# Reference sample-1 dataset generator vs libstdc++ itself
# ---------------------------------------------------------------------------
def test_sample1_dataset_matches_libstdcxx(tmp_path):
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src = tmp_path / "r.cpp"
    src.write_text("""
#include <random>
#include <cstdio>
int main(){ std::default_random_engine eg(413); std::uniform_real_distribution<float> dis(0.1, 2.0);
 for (int i=0;i<200;i++){ float a=1.0f; float b=dis(eg)*a; float m=std::max(a,b); std::printf("%.9g %.9g\\n", a/m, b/m);} }
""")
    exe = tmp_path / "r"
    subprocess.run(["g++", "-O0", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    vals = np.array(out, dtype=np.float64).reshape(-1, 2)
    b = W.reference_sample1_ellipses(200)
    np.testing.assert_array_equal(b.a[1:], vals[1:, 0].astype(np.float32).astype(np.float64))
    np.testing.assert_array_equal(b.b[1:], vals[1:, 1].astype(np.float32).astype(np.float64))
    assert b.a[0] == 1.0 and b.b[0] == 1.0
