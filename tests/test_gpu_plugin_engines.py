"""GPU and CPU engine subclasses of one plug-in header, side by side as in
gcurval's samples (tests/cpp/plugin_ellipse.h, plugin_bspline.h), built by
hipcc into tests/cpp/plugin_engines and run on the same inputs.

* GPU lengths against the CPU subclass, the oracle's RefCpu and the
  long-double truth: 1e-6 relative in FP32, 1e-10 in FP64.
* GPU sample parameters against the truth: 1e-6 (FP32), 1e-9 (FP64).
* CPU subclass (the reference's leaf-midpoint search) against the GPU: within
  the midpoint resolution of its depth cap.
"""
import os

import numpy as np
import pytest

from gcurval_amd import EllipseBatch
from gcurval_amd import workloads as W
from oracle import oracle as O

import plugin_io as P

pytestmark = pytest.mark.gpu


def _exe():
    if not os.path.exists(P.GPU_EXE):
        pytest.fail("tests/cpp/plugin_engines is not built (python -m gcurval_amd.build)")
    return P.GPU_EXE


def _both(tmp_path, prec, family, n, total, dim):
    outs = {}
    for eng in ("gpu", "cpu"):
        out = str(tmp_path / f"{eng}.bin")
        print(f"[plugin_engines] {prec} {family} {eng}", flush=True)
        r = P.run(_exe(), prec, family, eng, str(tmp_path / "in.bin"), out, 4, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        outs[eng] = P.read_out(out, n, total, dim)
    return outs


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_ellipse_plugin_gpu_and_cpu_subclasses(tmp_path, prec):
    rng = np.random.default_rng(31)
    n = 64
    a = rng.uniform(0.5, 3.0, n).astype(np.float32).astype(np.float64)
    b = (a * rng.uniform(0.05, 1.0, n)).astype(np.float32).astype(np.float64)
    counts = rng.integers(0, 80, n)
    total = int(counts.sum())
    pos = np.sort(rng.random(total)).astype(np.float32).astype(np.float64)
    P.write_ellipse(tmp_path / "in.bin", a, b, counts, pos)
    o = _both(tmp_path, prec, "ellipse", n, total, 2)
    (Lg, ug, pg), (Lc, uc, pc) = o["gpu"], o["cpu"]
    batch = EllipseBatch(a, b)
    Lt, _ = O.Truth(batch).lengths()
    rtol = 1e-6 if prec == "f32" else 1e-10
    assert np.max(np.abs(Lg - Lt) / Lt) <= rtol
    assert np.max(np.abs(Lg - Lc) / Lc) <= rtol
    ref = O.RefCpu(batch, O.opts(0 if prec == "f32" else 1, 3))
    ref.build()
    assert np.max(np.abs(Lc - ref.lengths()) / Lc) <= (1e-6 if prec == "f32" else 1e-13)
    ut, _ = O.Truth(batch).reverse(counts, pos, True)
    assert np.max(np.abs(ug - ut)) <= (1e-6 if prec == "f32" else 1e-9)
    depth = 14 if prec == "f32" else 23
    assert np.max(np.abs(uc - ug)) <= 2 * np.pi * 2.0 ** -(depth + 2) + 2e-6
    cid = np.repeat(np.arange(n), counts)
    np.testing.assert_allclose(pg[:, 0], a[cid] * np.cos(ug), atol=2e-6 if prec == "f32" else 1e-12)
    np.testing.assert_allclose(pg[:, 1], b[cid] * np.sin(ug), atol=2e-6 if prec == "f32" else 1e-12)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_bspline_plugin_gpu_and_cpu_subclasses(tmp_path, prec):
    batch = W.bspline_batch(24, seed=32, degrees=(3,), dims=(3,), n_ctrl_range=(8, 120))
    counts = 8 * batch.intervals_per_curve()
    total = int(counts.sum())
    P.write_bspline(tmp_path / "in.bin", batch, counts)
    o = _both(tmp_path, prec, "bspline", batch.n_curve, total, 3)
    (Lg, ug, pg), (Lc, uc, pc) = o["gpu"], o["cpu"]
    Lt, _ = O.Truth(batch).lengths()
    rtol = 1e-6 if prec == "f32" else 1e-10
    assert np.max(np.abs(Lg - Lt) / Lt) <= rtol
    assert np.max(np.abs(Lg - Lc) / Lc) <= rtol
    pos = O.uniform_positions(counts, closed=True)
    ut, _ = O.Truth(batch).reverse(counts, pos, True)
    assert np.max(np.abs(ug - ut)) <= (1e-6 if prec == "f32" else 1e-9)
    # the CPU search stops at knot-span widths / 2^(depth + 1) or the tolerance
    assert np.max(np.abs(uc - ug)) <= 1e-4
    assert np.all(np.isfinite(pg)) and np.all(np.isfinite(pc))
