"""GPU tests of graph capture and the launch-shape autotuner.

A captured step (learned-seed or plain build, then uniform sampling) is
replayed several times with eager library work and host synchronisations in
between; every replay must reproduce the eager step's lengths and samples
bit for bit (the output buffer is cleared before each replay).
"""
import numpy as np
import pytest

import gcurval_amd as G
from gcurval_amd import seed as S
from gcurval_amd import workloads as W

pytestmark = pytest.mark.gpu


def _offsets(counts):
    off = np.zeros(counts.size + 1, np.int64)
    np.cumsum(counts, out=off[1:])
    return off


@pytest.mark.parametrize("learned", [False, True])
def test_graph_replays_match_eager(ctx, learned):
    b = W.bspline_batch(40, seed=81, degrees=(2, 3, 5), dims=(2, 3), n_ctrl_range=(16, 200), hairpin_frac=0.2)
    counts = 8 * b.intervals_per_curve()
    off = _offsets(counts)
    n_total = int(off[-1])
    mlp = S.SeedMLP.pretrained(ctx) if learned else None
    other = W.ellipse_batch(500, seed=82)
    with G.ArclengthCalculateGpu(ctx, b, normalize=True) as eng, \
            G.ArclengthCalculateGpu(ctx, other, normalize=True) as side:
        d_off = ctx.array(off)
        d_out = ctx.empty(n_total, np.float64)

        def step():
            if learned:
                eng.createArclengthTableLearned(mlp)
            else:
                eng.createArclengthTable()
            eng.sample_device(d_off, n_total, None, True, d_out, None, uniform_closed=True)

        step()
        ctx.sync()
        L_ref = eng.getAllCurveArclengths().copy()
        u_ref = d_out.download()
        ctx.graph_begin()
        step()
        g = ctx.graph_end()
        zeros = np.zeros(n_total, np.float64)
        for _ in range(3):
            side.createArclengthTable()          # eager work between replays
            side.getAllCurveArclengths()
            d_out.upload(zeros)
            g.launch()
            ctx.sync()
            np.testing.assert_array_equal(d_out.download(), u_ref)
            eng._lengths = None
            np.testing.assert_array_equal(eng.getAllCurveArclengths(), L_ref)
        g.close()
        d_off.free()
        d_out.free()
    if mlp is not None:
        mlp.close()


def test_capture_refuses_synchronising_calls(ctx):
    e = W.ellipse_batch(64, seed=83)
    with G.ArclengthCalculateGpu(ctx, e, normalize=True) as eng:
        eng.createArclengthTable()
        ctx.graph_begin()
        eng.createArclengthTable()
        with pytest.raises(G.GcvError):
            ctx.sync()
        with pytest.raises(G.GcvError):
            eng.stats()
        with pytest.raises(G.GcvError):
            eng.close()  # destroying a table synchronises: refused, the table stays usable
        g = ctx.graph_end()
        g.launch()
        ctx.sync()
        assert eng.stats()["quadratures"] >= 3 * 64 * 64
        g.close()


def test_autotune_keeps_results(ctx):
    b = W.bspline_batch(200, seed=84, degrees=(3,), dims=(3,), n_ctrl_range=(8, 400))
    with G.ArclengthCalculateGpu(ctx, b, normalize=True) as eng:
        eng.createArclengthTable()
        L0 = eng.getAllCurveArclengths().copy()
        rep = eng.autotune(reps=2)
        assert rep["n_tried"] == 9
        assert rep["sub_waves_per_block"] in (1, 2, 4) and rep["sub_grid_pct"] in (25, 50, 100)
        assert 0 < rep["best_build_ms"] <= rep["default_build_ms"] * 1.0001
        eng._lengths = None
        np.testing.assert_allclose(eng.getAllCurveArclengths(), L0, rtol=1e-13)
        for wpb, pct in ((1, 25), (2, 50)):
            eng.set_launch(wpb, pct)
            eng.createArclengthTable()
            np.testing.assert_allclose(eng.getAllCurveArclengths(), L0, rtol=1e-13)
        with pytest.raises(G.GcvError):
            eng.set_launch(3, 100)


def test_learned_bins_on_several_streams_match_one_stream(ctx):
    """The learned build's per-bin launches spread over the auxiliary streams
    (joined before the build returns) give the single-stream tables bit for
    bit; the autotuner tries the spread for learned builds."""
    b = W.bspline_batch(120, seed=85, degrees=(2, 3, 5), dims=(2, 3), n_ctrl_range=(16, 300), hairpin_frac=0.2)
    counts = 8 * b.intervals_per_curve()
    mlp = S.SeedMLP.pretrained(ctx)
    try:
        with G.ArclengthCalculateGpu(ctx, b, normalize=True) as eng:
            out = {}
            for streams in (1, 3):
                eng.set_launch(0, 100, streams)
                eng.createArclengthTableLearned(mlp)
                eng._lengths = None
                out[streams] = (eng.getAllCurveArclengths().copy(), eng.arclengthReverse(counts, None, True))
            np.testing.assert_array_equal(out[1][0], out[3][0])
            np.testing.assert_array_equal(out[1][1], out[3][1])
            rep = eng.autotune(mlp=mlp, reps=1)
            assert rep["n_tried"] == 10 and rep["bin_streams"] in (1, 3)
            eng._lengths = None
            np.testing.assert_array_equal(eng.getAllCurveArclengths(), out[1][0])
            with pytest.raises(G.GcvError):
                eng.set_launch(0, 100, 4)
    finally:
        mlp.close()
