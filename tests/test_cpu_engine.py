"""CPU engine of the template surface (include/gcurval/arclength_calculate_cpu.h,
the counterpart of gcurval/detail/cpu/arclength_calculate.h:10-440), driven
through plug-in headers whose GPU and CPU subclasses sit side by side as in
gcurval's samples (tests/cpp/plugin_ellipse.h, plugin_bspline.h).

The driver is built by g++ with AddressSanitizer and UndefinedBehaviorSanitizer
(no GPU code), so every run here is also a sanitizer run. The CPU engine runs
the reference CPU algorithm with the plug-in's functors, and the oracle's
RefCpu restates the same algorithm with its own evaluation of the same curves,
so the two agree bit for bit (float and double): lengths, table-driven
sample parameters, and point output. The unit circle of
samples/sample1_ellipse/main.cpp:40-41 pins both against the closed form.
"""
import math
import os

import numpy as np
import pytest

from gcurval_amd import EllipseBatch
from gcurval_amd import workloads as W
from oracle import oracle as O

import plugin_io as P


@pytest.fixture(scope="module")
def exe():
    return P.build_cpu()


def _ellipses(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.5, 3.0, n).astype(np.float32).astype(np.float64)
    b = (a * rng.uniform(0.05, 1.0, n)).astype(np.float32).astype(np.float64)
    return a, b


def _run(exe, tmp_path, prec, family, threads=1):
    inp, out = str(tmp_path / "in.bin"), str(tmp_path / f"out_{threads}.bin")
    if not os.path.exists(inp):
        raise AssertionError("input not written")
    r = P.run(exe, prec, family, "cpu", inp, out, threads, P.SAN_ENV)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr
    return out


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_ellipse_cpu_engine_matches_reference_algorithm(exe, tmp_path, prec):
    a, b = _ellipses(12, 5)
    rng = np.random.default_rng(6)
    counts = rng.integers(0, 40, a.size)
    counts[3] = 1
    total = int(counts.sum())
    pos = rng.random(total)
    pos[:2] = [0.0, 1.0]
    if prec == "f32":
        pos = pos.astype(np.float32).astype(np.float64)  # the engine receives Dt positions
    P.write_ellipse(tmp_path / "in.bin", a, b, counts, pos)
    out = _run(exe, tmp_path, prec, "ellipse")
    L, u, pts = P.read_out(out, a.size, total, 2)
    ref = O.RefCpu(EllipseBatch(a, b), O.opts(0 if prec == "f32" else 1, 3))
    ref.build()
    np.testing.assert_array_equal(L, ref.lengths())
    ur, _ = ref.reverse(counts, pos, True)
    np.testing.assert_array_equal(u, ur)
    # points: the plug-in's GetPoint at the returned parameters
    cid = np.repeat(np.arange(a.size), counts)
    f = np.float32 if prec == "f32" else np.float64
    uu = u.astype(f)
    exp = np.stack([(a[cid].astype(f) * np.cos(uu)).astype(f), (b[cid].astype(f) * np.sin(uu)).astype(f)], 1)
    np.testing.assert_allclose(pts, exp, rtol=0, atol=2e-6 if prec == "f32" else 1e-14)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_bspline_cpu_engine_matches_reference_algorithm(exe, tmp_path, prec):
    batch = W.bspline_batch(6, seed=21, degrees=(3,), dims=(3,), n_ctrl_range=(8, 40))
    counts = 4 * batch.intervals_per_curve()
    P.write_bspline(tmp_path / "in.bin", batch, counts)
    out = _run(exe, tmp_path, prec, "bspline")
    total = int(counts.sum())
    L, u, pts = P.read_out(out, batch.n_curve, total, 3)
    ref = O.RefCpu(batch, O.opts(0 if prec == "f32" else 1, 3))
    ref.build()
    np.testing.assert_array_equal(L, ref.lengths())
    # the driver's uniform positions j / (n - 1), divided as the engine divides them
    pos = np.concatenate([np.arange(n) / (n - 1) if n > 1 else np.zeros(n) for n in counts])
    if prec == "f32":
        pos = pos.astype(np.float32).astype(np.float64)
    ur, _ = ref.reverse(counts, pos, True)
    np.testing.assert_array_equal(u, ur)
    assert np.all(np.isfinite(pts))


def test_threads_do_not_change_results(exe, tmp_path):
    batch = W.bspline_batch(9, seed=22, degrees=(3,), dims=(3,), n_ctrl_range=(8, 30))
    counts = 3 * batch.intervals_per_curve()
    P.write_bspline(tmp_path / "in.bin", batch, counts)
    one = np.fromfile(_run(exe, tmp_path, "f32", "bspline", threads=1))
    four = np.fromfile(_run(exe, tmp_path, "f32", "bspline", threads=4))
    np.testing.assert_array_equal(one, four)


@pytest.mark.parametrize("prec,tol,depth", [("f32", 2e-6, 14), ("f64", 1e-12, 23)])
def test_unit_circle_known_answer(exe, tmp_path, prec, tol, depth):
    # samples/sample1_ellipse/main.cpp:40-41: curve 0 is the unit circle, L = 2 pi,
    # and uniform arc-length samples are t_j = 2 pi j / (N - 1). The speed is
    # constant, so the table is the single row [0, 2 pi] and the reference's
    # search descends from it until the depth cap (14 in float, 23 in double;
    # its stop width, the tolerance, is finer): every sample is the midpoint of
    # a half of width 2 pi 2^-(depth + 1), within 2 pi 2^-(depth + 2) of t_j,
    # and that bound is attained.
    n = 200
    P.write_ellipse(tmp_path / "in.bin", [1.0], [1.0], [n], np.arange(n) / (n - 1))
    L, u, _ = P.read_out(_run(exe, tmp_path, prec, "ellipse"), 1, n, 2)
    assert abs(L[0] - 2 * math.pi) <= tol * 2 * math.pi
    exact = 2 * math.pi * np.arange(n) / (n - 1)
    bound = 2 * math.pi * 2.0 ** -(depth + 2)
    err = np.max(np.abs(u - exact))
    assert bound * 0.9 <= err <= bound * (1 + 1e-3), (err, bound)


def test_empty_and_zero_sample_curves(exe, tmp_path):
    P.write_ellipse(tmp_path / "in.bin", [2.0, 1.0, 3.0], [1.0, 1.0, 0.5], [0, 5, 0], np.linspace(0, 1, 5))
    L, u, pts = P.read_out(_run(exe, tmp_path, "f64", "ellipse"), 3, 5, 2)
    assert np.all(L > 0) and u.size == 5 and np.all(np.diff(u) > 0)
