"""GPU checks of the learned seed: the BF16 matrix-core network against the
numpy reference (forward outputs, loss, every parameter gradient), training
progress, the fused feature+predict kernel, and learned-seed builds against
the long-double truth.
"""
import os

import numpy as np
import pytest

import gcurval_amd as G
from gcurval_amd import seed as S
from gcurval_amd import workloads as W
from oracle import mlp_ref as R
from oracle import oracle as O

pytestmark = pytest.mark.gpu

def assert_quads(cost, st):
    """Per-interval quadrature counts add up to the build's: exactly without
    re-runs; a tile re-run with the full task stack also counted its first pass."""
    if st["reruns"] == 0:
        assert int(cost.sum()) == st["quadratures"]
    else:
        assert int(cost.sum()) < st["quadratures"]



def _params(seed):
    rng = np.random.default_rng(seed)
    p = S.he_init(seed)
    W1, b1, W2, b2, W3, b3 = R.unpack(p)
    b1[:] = rng.normal(0, 0.1, b1.size)
    b2[:] = rng.normal(0, 0.1, b2.size)
    b3[:] = rng.normal(0, 0.5, b3.size)
    return p


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, 16)).astype(np.float32)
    y = np.stack([rng.uniform(0, 6, n), rng.integers(0, 2, n), rng.uniform(-3, 3, n), rng.integers(0, 2, n)],
                 axis=1).astype(np.float32)
    y[::5, 2:] = np.nan  # rows without start labels leave those terms out
    return X, y


@pytest.mark.parametrize("n", [1, 16, 1000, 40007])
def test_forward_matches_reference(ctx, n):
    p = _params(1)
    X, _ = _data(n, 2)
    Xb = S.bf16_bits(X)
    with S.SeedMLP(ctx, p) as m:
        d_x = ctx.array(Xb)
        out = m.forward(d_x, n)
        d_x.free()
    ref = R.forward(p, R.bf16(X), quant=True)
    err = np.abs(out - ref)
    scale = np.max(np.abs(ref))
    # accumulation order differs from numpy, so an activation may round to the
    # neighbouring BF16 value; a layout or transpose mistake gives O(1) errors
    assert np.max(err) <= 0.02 * scale, (np.max(err), scale)
    assert np.mean(err) <= 2e-3 * np.mean(np.abs(ref))


def test_gradients_match_reference(ctx):
    n = 5003
    p = _params(5)
    X, y = _data(n, 6)
    with S.SeedMLP(ctx, p) as m:
        d_x = ctx.array(S.bf16_bits(X))
        d_y = ctx.array(y)
        loss, g = m.grad(d_x, d_y, n, 0.7)
        # gradients are reset every call: a second call gives the same values
        loss2, g2 = m.grad(d_x, d_y, n, 0.7)
        d_x.free()
        d_y.free()
    Lr, gr = R.loss_and_grad(p, R.bf16(X), y, 0.7, quant=True)
    assert abs(loss - Lr) <= 1e-3 * abs(Lr)
    assert abs(loss2 - loss) <= 1e-5 * abs(loss)
    np.testing.assert_allclose(g2, g, rtol=1e-4, atol=1e-7)
    blocks = {"W1": (0, 1024), "b1": (1024, 1088), "W2": (1088, 5184), "b2": (5184, 5248), "W3": (5248, 5504),
              "b3": (5504, 5508)}
    for name, (a, b) in blocks.items():
        rel = np.linalg.norm(g[a:b] - gr[a:b]) / np.linalg.norm(gr[a:b])
        assert rel <= 2e-2, (name, rel)


def test_training_reduces_loss(ctx):
    # learnable targets: smooth functions of the features and signs of features
    rng = np.random.default_rng(11)
    n = 32768
    X = rng.normal(size=(n, 16)).astype(np.float32)
    y = np.stack([3 + np.tanh(X[:, 0] - 0.5 * X[:, 3]) * 2, (X[:, 5] > 0), np.tanh(X[:, 7] + X[:, 1]),
                  (X[:, 9] - X[:, 2] > 0)], axis=1).astype(np.float32)
    with S.SeedMLP(ctx, S.he_init(2)) as m:
        d_x = ctx.array(S.bf16_bits(X))
        d_y = ctx.array(y)
        l0, _ = m.grad(d_x, d_y, n, 0.5)
        for _ in range(300):
            m.train_step(d_x, d_y, n, 3e-3, 0.5)
        l1, _ = m.grad(d_x, d_y, n, 0.5)
        p, _, _ = m.get_params()
        d_x.free()
        d_y.free()
    assert np.all(np.isfinite(p))
    assert l1 <= 0.1 * l0, (l0, l1)
    assert abs(l1 - R.loss(p, R.bf16(X), y, 0.5, quant=True)) <= 0.02 * l1 + 1e-4


@pytest.fixture(scope="module")
def mixed():
    return W.bspline_batch(60, seed=61, degrees=(2, 3, 5), dims=(2, 3), n_ctrl_range=(16, 120), hairpin_frac=0.2)


def test_predict_matches_features_and_forward(ctx, mixed):
    rng = np.random.default_rng(12)
    p = _params(7)
    mean = rng.normal(0, 1, 16).astype(np.float32)
    istd = rng.uniform(0.2, 2, 16).astype(np.float32)
    for batch in (mixed, W.ellipse_batch(333, seed=62)):
        with G.ArclengthCalculateGpu(ctx, batch, normalize=True) as eng, S.SeedMLP(ctx, p, mean, istd) as m:
            F = m.features(eng)
            assert F.shape == (eng.n_intervals, 16) and np.all(np.isfinite(F))
            assert np.all((F[:, 15] >= 0) & (F[:, 15] <= 1.0 + 1e-6))
            for hedge in (False, True):
                k, raw = m.predict(eng, hedge)
                ref = R.forward(p, R.bf16((F - mean) * istd), quant=True)
                assert np.max(np.abs(raw - ref)) <= 0.02 * np.max(np.abs(ref))
                kr = np.clip(np.rint(raw[:, 0]) + (hedge & (raw[:, 1] < 0)), 0, 6).astype(np.int8)
                np.testing.assert_array_equal(k, kr)


def _pretrained(ctx):
    if not os.path.exists(S.WEIGHTS):
        pytest.skip("no trained weights in the tree")
    return S.SeedMLP.pretrained(ctx)


def test_learned_build_matches_truth(ctx, mixed):
    tr = O.Truth(mixed)
    L, cum = tr.lengths()
    counts = 16 * mixed.intervals_per_curve()
    ut, _ = tr.reverse(counts, O.uniform_positions(counts))
    with _pretrained(ctx) as m, G.ArclengthCalculateGpu(ctx, mixed, normalize=True) as eng:
        d_cost = ctx.empty(mixed.n_intervals, np.int32)
        for hedge in (False, True):
            eng.createArclengthTableLearned(m, hedge, d_cost)
            assert np.max(np.abs(eng.getAllCurveArclengths() - L) / L) <= 1e-6
            u = eng.arclengthReverse(counts, None, True)
            assert np.max(np.abs(u - ut)) <= 1e-6
            k_used = eng.seed_k()
            k_pred, _ = m.predict(eng, hedge)
            np.testing.assert_array_equal(k_used, k_pred)
            keff = np.maximum(k_used, eng.seed_floor()).astype(np.int64)
            cost = d_cost.download()
            assert np.all(cost >= 3 * (1 << keff))
            assert_quads(cost, eng.stats())
        d_cost.free()


def test_learned_seed_halves_quadratures(ctx):
    b = W.bspline_batch(300, seed=63, degrees=(2, 3, 5), dims=(2, 3), n_ctrl_range=(16, 512), hairpin_frac=0.2)
    with _pretrained(ctx) as m, G.ArclengthCalculateGpu(ctx, b, normalize=True) as eng:
        eng.createArclengthTable()
        q_plain = eng.stats()["quadratures"]
        L_plain = eng.getAllCurveArclengths()
        eng.createArclengthTableLearned(m)
        q_seed = eng.stats()["quadratures"]
        L_seed = eng.getAllCurveArclengths()
    assert np.max(np.abs(L_seed - L_plain) / L_plain) <= 2e-6
    assert q_plain >= 2 * q_seed, (q_plain, q_seed)
