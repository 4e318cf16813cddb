"""CPU simulation of the wave-local refinement stack of k_subdivide.

The kernel pops the top min(size, 64) tasks, evaluates them, and pushes the
two children of every failing task on top (in lane order). Its LDS capacity
is 64 (max_depth + 2); the claimed bound is 64 (max_depth + 1). This test
drives the same discipline with adversarial and random failure patterns and
checks the bound and that every root's leaves tile it exactly.
"""
import numpy as np
import pytest


def simulate(fail, max_depth, n_roots=64):
    """fail(root, level, idx) -> bool. Returns (max size, leaves per root)."""
    stack = []
    leaves = [[] for _ in range(n_roots)]
    for r in range(n_roots):
        if max_depth > 0 and fail(r, 0, 0):
            stack += [(r, 1, 0), (r, 1, 1)]
        else:
            leaves[r].append((0, 0))
    max_size = len(stack)
    while stack:
        n = min(64, len(stack))
        popped = [stack[len(stack) - 1 - lane] for lane in range(n)]
        del stack[len(stack) - n:]
        for (r, l, idx) in popped:  # lane order
            if l < max_depth and fail(r, l, idx):
                stack += [(r, l + 1, 2 * idx), (r, l + 1, 2 * idx + 1)]
            else:
                leaves[r].append((l, idx))
        max_size = max(max_size, len(stack))
    return max_size, leaves


def check_tiling(leaves):
    for lv in leaves:
        total = sum(2.0 ** -l for l, _ in lv)
        assert abs(total - 1.0) < 1e-12
        spans = sorted((idx * 2.0 ** -l, (idx + 1) * 2.0 ** -l) for l, idx in lv)
        for (a0, a1), (b0, b1) in zip(spans, spans[1:]):
            assert a1 == b0


@pytest.mark.parametrize("max_depth", [1, 4, 10, 14])
def test_all_fail_is_the_worst_case(max_depth):
    size, leaves = simulate(lambda r, l, i: True, max_depth)
    assert size <= 64 * (max_depth + 1)
    check_tiling(leaves)


@pytest.mark.parametrize("seed", range(12))
def test_random_failures_stay_within_bound(seed):
    rng = np.random.default_rng(seed)
    max_depth = int(rng.integers(2, 11))
    p = rng.uniform(0.3, 0.95, max_depth + 1)
    roots_fail = rng.random(64) < rng.uniform(0.1, 1.0)
    cache = {}

    def fail(r, l, idx):
        if l == 0:
            return roots_fail[r]
        key = (r, l, idx)
        if key not in cache:
            cache[key] = rng.random() < p[l]
        return cache[key]

    size, leaves = simulate(fail, max_depth)
    assert size <= 64 * (max_depth + 1)
    check_tiling(leaves)


@pytest.mark.parametrize("max_depth", [6, 12])
def test_one_deep_path_per_root(max_depth):
    # only the leftmost descendant keeps failing: a thin deep tree per root
    size, leaves = simulate(lambda r, l, i: i == 0, max_depth)
    assert size <= 64 * (max_depth + 1)
    check_tiling(leaves)
