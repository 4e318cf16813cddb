"""gcv_sample_noalloc (include/gcurval_c.h): the sampling entry point that
allocates nothing, its block maps in the library's static device scratch.
Called through the C ABI and compared bit for bit with gcv_sample on the same
table and inputs (the same kernels run), on two calls in a row, the second
with a smaller input, so the scratch holds the first call's maps when the
second starts. B-splines (several intervals per curve: block -> interval map)
and ellipses (the run sampler), uniform and explicit positions."""
import ctypes as C

import numpy as np
import pytest

import gcurval_amd as G
from gcurval_amd import workloads as W

pytestmark = pytest.mark.gpu

MAX_SAMPLES = (1 << 20) * 256


def _noalloc(eng, d_off, total, d_pos, closed, d_out, d_seg):
    rc = G.lib().gcv_sample_noalloc(eng.h, d_off.ptr, total, d_pos.ptr if d_pos is not None else None,
                                    1 if closed else 0, 1, d_out.ptr, d_seg.ptr)
    return rc


def _both(ctx, eng, counts, pos=None, closed=True):
    off = np.zeros(counts.size + 1, np.int64)
    np.cumsum(counts, out=off[1:])
    total = int(off[-1])
    d_off = ctx.array(off)
    d_pos = ctx.array(pos) if pos is not None else None
    outs = []
    for which in ("plain", "noalloc"):
        d_out = ctx.empty(total, np.float64)
        d_seg = ctx.empty(total, np.int32)
        if which == "plain":
            eng.sample_device(d_off, total, d_pos, True, d_out, d_seg, uniform_closed=closed)
        else:
            ctx._check(_noalloc(eng, d_off, total, d_pos, closed, d_out, d_seg))
        ctx.sync()
        outs.append((d_out.download(), d_seg.download()))
        d_out.free()
        d_seg.free()
    d_off.free()
    if d_pos is not None:
        d_pos.free()
    return outs


def test_noalloc_sampling_matches_gcv_sample_bspline():
    with G.Context(0) as ctx:
        b = W.bspline_batch(300, seed=91, degrees=(2, 3, 5), dims=(2, 3), n_ctrl_range=(8, 400))
        with G.ArclengthCalculateGpu(ctx, b, normalize=True) as eng:
            eng.createArclengthTable()
            rng = np.random.default_rng(92)
            big = rng.integers(0, 3000, b.n_curve).astype(np.int64)
            small = rng.integers(0, 40, b.n_curve).astype(np.int64)
            print(f"[no_alloc] bspline uniform {int(big.sum())} then {int(small.sum())} samples", flush=True)
            for counts in (big, small):
                (u0, s0), (u1, s1) = _both(ctx, eng, counts)
                np.testing.assert_array_equal(u1, u0)
                np.testing.assert_array_equal(s1, s0)
            pos = np.concatenate([np.sort(rng.random(int(n))) for n in small])
            print(f"[no_alloc] bspline explicit positions {pos.size}", flush=True)
            (u0, s0), (u1, s1) = _both(ctx, eng, small, pos)
            np.testing.assert_array_equal(u1, u0)


def test_noalloc_sampling_matches_gcv_sample_ellipse():
    with G.Context(0) as ctx:
        b = W.ellipse_batch(5000, seed=93)
        with G.ArclengthCalculateGpu(ctx, b, normalize=True) as eng:
            eng.createArclengthTable()
            big = W.ellipse_sample_counts(b.n_curve, seed=94)
            small = W.ellipse_sample_counts(b.n_curve, seed=95, lo=1, hi=50)
            print(f"[no_alloc] ellipse {int(big.sum())} then {int(small.sum())} samples", flush=True)
            for counts, closed in ((big, True), (small, False)):
                (u0, _), (u1, _) = _both(ctx, eng, counts, closed=closed)
                np.testing.assert_array_equal(u1, u0)


def test_noalloc_rejects_more_samples_than_its_scratch():
    with G.Context(0) as ctx:
        b = W.ellipse_batch(4, seed=96)
        with G.ArclengthCalculateGpu(ctx, b) as eng:
            eng.createArclengthTable()
            d_off = ctx.array(np.array([0, 1, 2, 3, 4], np.int64))
            d_out = ctx.empty(4, np.float64)
            rc = G.lib().gcv_sample_noalloc(eng.h, d_off.ptr, MAX_SAMPLES + 1, None, 1, 1, d_out.ptr, None)
            assert rc == 1  # GCV_ERR_INVALID, before any launch
            assert b"static scratch" in G.lib().gcv_last_error(ctx.h)
            d_off.free()
            d_out.free()
