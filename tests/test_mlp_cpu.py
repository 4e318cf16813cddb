"""CPU checks of the learned seed's host logic and of its numpy reference network.

The reference network's gradients are pinned by central differences of the
exact FP64 network; the BF16 mode is pinned against the exact mode and the
rounding helper against hand-computed bit patterns.
"""
import numpy as np
import pytest

from gcurval_amd import seed as S
from oracle import mlp_ref as R


def _params(seed, bias=0.1):
    rng = np.random.default_rng(seed)
    p = R.init_params(seed).astype(np.float64)
    W1, b1, W2, b2, W3, b3 = R.unpack(p)
    b1[:] = rng.normal(0, bias, b1.size)
    b2[:] = rng.normal(0, bias, b2.size)
    b3[:] = rng.normal(0, bias, b3.size)
    return p


def test_sizes():
    assert R.N_PARAMS == S.N_PARAMS == 5508
    assert S.he_init(0).shape == (5508,)
    assert len(S.FEATURE_NAMES) == S.N_FEAT


def test_bf16_rounding_bits():
    x = np.array([1.0, -2.0, 0.0], np.float32)
    np.testing.assert_array_equal(S.bf16_bits(x), [0x3F80, 0xC000, 0x0000])
    # ties go to even, above-tie rounds up
    u = np.array([0x3F808000, 0x3F818000, 0x3F808001, 0x3F807FFF], np.uint32)
    np.testing.assert_array_equal(S.bf16_bits(u.view(np.float32)), [0x3F80, 0x3F82, 0x3F81, 0x3F80])
    r = np.random.default_rng(0).normal(size=1000).astype(np.float32)
    np.testing.assert_array_equal(S.bf16_bits(r), R.bf16_bits(r))
    # rounding error at most half an ulp of an 8-bit significand
    assert np.all(np.abs(R.bf16(r) - r) <= np.abs(r) * 2.0 ** -8 + 1e-38)


def _targets(rng, n):
    y = np.stack([rng.uniform(0, 6, n), rng.integers(0, 2, n), rng.uniform(-3, 3, n), rng.integers(0, 2, n)],
                 axis=1).astype(np.float64)
    y[::3, 2:] = np.nan  # rows without start labels (ellipses)
    return y


def test_exact_gradient_matches_central_differences():
    rng = np.random.default_rng(3)
    p = _params(3)
    X = rng.normal(size=(9, 16))
    y = _targets(rng, 9)
    L, g = R.loss_and_grad(p, X, y, 0.7, quant=False)
    assert abs(L - R.loss(p, X, y, 0.7, quant=False)) <= 1e-12 * abs(L)
    idx = np.concatenate([rng.choice(np.arange(0, 1024), 25, replace=False),   # W1
                          np.arange(1024, 1088, 7),                            # b1
                          rng.choice(np.arange(1088, 5184), 25, replace=False), # W2
                          np.arange(5184, 5248, 7),                            # b2
                          np.arange(5248, 5508)])                              # W3, b3
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (R.loss(p + e, X, y, 0.7, quant=False) - R.loss(p - e, X, y, 0.7, quant=False)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * (1 + abs(g[i])), i


def test_bf16_network_close_to_exact():
    rng = np.random.default_rng(4)
    p = _params(4)
    X = rng.normal(size=(200, 16))
    y = _targets(rng, 200)
    Oq = R.forward(p, X, quant=True)
    Oe = R.forward(p, X, quant=False)
    assert np.max(np.abs(Oq - Oe)) <= 0.05 * np.max(np.abs(Oe))
    _, gq = R.loss_and_grad(p, X, y, 0.5, quant=True)
    _, ge = R.loss_and_grad(p, X, y, 0.5, quant=False)
    assert np.linalg.norm(gq - ge) <= 0.03 * np.linalg.norm(ge)


def test_nan_targets_leave_their_terms_out():
    rng = np.random.default_rng(5)
    p = _params(5)
    X = rng.normal(size=(12, 16))
    y = _targets(rng, 12)
    L, g = R.loss_and_grad(p, X, y, 0.5, quant=False)
    # the same loss and gradient as the rows' seed terms plus the start terms of labelled rows only
    lab = ~np.isnan(y[:, 2])
    L_seed, g_seed = R.loss_and_grad(p, X, np.concatenate([y[:, :2], np.full((12, 2), np.nan)], 1), 0.5, quant=False)
    L_start, g_start = R.loss_and_grad(p, X[lab], np.concatenate([np.full((lab.sum(), 2), np.nan), y[lab, 2:]], 1),
                                       0.5, quant=False)
    n_lab = lab.sum()
    assert abs(L - (L_seed + L_start * n_lab / 12)) <= 1e-12 * abs(L)
    np.testing.assert_allclose(g, g_seed + g_start * n_lab / 12, rtol=1e-10, atol=1e-14)
    assert np.isfinite(L) and np.all(np.isfinite(g))


def test_start_targets():
    e = np.array([-30.0, -13.0, -12.9, 0.0, np.nan], np.float32)
    y = S.start_targets(e)
    np.testing.assert_allclose(y[:4, 0], (e[:4] + 20) / 8)
    np.testing.assert_array_equal(y[:4, 1], [1, 1, 0, 0])
    assert np.isnan(y[4]).all()


def test_labels_from_costs():
    # interval 0: k=1 passes cleanly; 1: floor 4 forces k>=4; 2: never clean, k=6 cheapest;
    # 3: k=1 refined but cheapest (9 < 12), k=2 the first clean seed; 4: k=4 the first clean
    # seed, k=2 with one refinement task (14 quadratures) cheapest even with refinement
    # quadratures weighted 6x (12 + 6 * 2 = 24 < 48)
    cost = np.array([[30, 3, 9000, 40, 63], [6, 6, 8000, 9, 26], [12, 12, 7000, 12, 14], [24, 24, 6000, 24, 26],
                     [48, 48, 5000, 48, 48], [96, 96, 4000, 96, 96], [192, 192, 1000, 192, 192]], np.int64)
    floor = np.array([0, 4, 0, 0, 0], np.int8)
    y = S.labels_from_costs(cost, floor, "clean")
    np.testing.assert_array_equal(y[:, 0], [1, 4, 6, 2, 4])
    np.testing.assert_array_equal(y[:, 1], [1, 1, 0, 1, 1])
    y = S.labels_from_costs(cost, floor, "cheapest")
    np.testing.assert_array_equal(y[:, 0], [1, 4, 6, 1, 2])
    np.testing.assert_array_equal(y[:, 1], [1, 1, 0, 0, 0])
    assert S.REFINE_WEIGHT == 6
    y = S.labels_from_costs(cost, floor, "weighted")
    # interval 3: k=1 weighs 6 + 6 * 3 = 24 > 12, so the clean k=2 wins
    np.testing.assert_array_equal(y[:, 0], [1, 4, 6, 2, 2])
    np.testing.assert_array_equal(y[:, 1], [1, 1, 0, 1, 0])
    with pytest.raises(ValueError):
        S.labels_from_costs(cost, floor, "nope")


def test_weights_file_shapes():
    import os
    if not os.path.exists(S.WEIGHTS):
        pytest.skip("no trained weights in the tree")
    with np.load(S.WEIGHTS, allow_pickle=False) as z:
        assert z["params"].shape == (5508,) and z["params"].dtype == np.float32
        assert z["mean"].shape == (16,) and z["istd"].shape == (16,)
        assert np.all(np.isfinite(z["params"])) and np.all(z["istd"] > 0)
