"""Learned seed of the table build (host side).

The plain build starts every knot interval from n_sub = 64 equal roots and
runs the halves-versus-whole error test on each (the reference's GPU engine,
gcurval/include/gcurval/arclengthCalculateGpu.h:266-297 and :360-380). Most
intervals of a smooth curve pass that test with far fewer roots. A seeded
build (gcv_table_build_seeded) starts interval i from 2^k_i roots, k_i in
0..6, still testing every root against the plain level-0 tolerance tol/n_sub,
so accuracy is that of the adaptive test whatever k_i is; the seed only moves
the work.

This module drives the device predictor of k_i: a 16-64-64-4 perceptron on
BF16 matrix cores (gcv_mlp_*). Training labels come from the library itself:
per-interval quadrature counts of uniform-k builds for k = 0..6; the label of
an interval is the smallest k at or above its safety floor whose roots all
pass the level-0 test (or, with the "cheapest" policy, the k of fewest
quadratures), and the second output learns whether that seed needs no
refinement.

The same network predicts the sampler's Newton start (outputs 2 and 3): on
the uniform-k table of the interval's label k, the library measures the
worst error of the quintic Hermite start of the inverse map over the
interval's roots against the converged Newton solution
(gcv_table_start_labels); output 2 learns its scaled log2 and output 3
whether it is at most 2^-13 of a root, in which case the sampler takes the
FP64 step straight from that start (gcv_table_set_newton_start).

``python -m gcurval_amd.seed --train`` regenerates ``seed_weights.npz``.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import time

import numpy as np

from . import (GCV_FP32, ArclengthCalculateGpu, Context, DeviceArray, _ptr, lib)

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "seed_weights.npz")

N_FEAT, N_HID, N_OUT = 16, 64, 4
N_PARAMS = N_FEAT * N_HID + N_HID + N_HID * N_HID + N_HID + N_HID * N_OUT + N_OUT
MAX_K = 6
# direct-start label: worst quintic-start error at most 2^-13 of a root (the
# sampler keeps an FP64 step that moves at most 2^-12)
START_LOG2_OK = -13.0

FEATURE_NAMES = [
    "log2 speed x=0", "log2 speed x=1/4", "log2 speed x=1/2", "log2 speed x=3/4", "log2 speed x=1",
    "log2 speed max/min", "cos(d0,d1)", "cos(d0,d1/2)", "cos(d1/2,d1)", "log2 tol0/length",
    "log2 lead coef/length", "log2 edge max/min", "min edge cos", "log2 span x intervals", "class code",
    "min/max speed (17)",
]


def bf16_bits(x: np.ndarray) -> np.ndarray:
    """float32 -> BF16 bit patterns, round to nearest even (finite inputs)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def he_init(seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    parts = [rng.normal(0, np.sqrt(2 / N_FEAT), N_FEAT * N_HID), np.zeros(N_HID),
             rng.normal(0, np.sqrt(2 / N_HID), N_HID * N_HID), np.zeros(N_HID),
             rng.normal(0, np.sqrt(1 / N_HID), N_HID * N_OUT), np.zeros(N_OUT)]
    return np.concatenate(parts).astype(np.float32)


class SeedMLP:
    """Device-resident predictor of per-interval log2 root counts."""

    def __init__(self, ctx: Context, params=None, mean=None, istd=None, seed: int = 0):
        self.ctx = ctx
        h = C.c_void_p()
        ctx._check(lib().gcv_mlp_create(ctx.h, C.byref(h)))
        self.h = h.value
        self.set_params(he_init(seed) if params is None else params,
                        np.zeros(N_FEAT, np.float32) if mean is None else mean,
                        np.ones(N_FEAT, np.float32) if istd is None else istd)

    @classmethod
    def pretrained(cls, ctx: Context, path: str = WEIGHTS) -> "SeedMLP":
        if not os.path.exists(path):
            raise FileNotFoundError(f"learned-seed weights not found: {path}")
        with np.load(path, allow_pickle=False) as z:
            return cls(ctx, z["params"], z["mean"], z["istd"])

    def set_params(self, params, mean, istd):
        p = np.ascontiguousarray(params, np.float32)
        m = np.ascontiguousarray(mean, np.float32)
        s = np.ascontiguousarray(istd, np.float32)
        if p.size != N_PARAMS or m.size != N_FEAT or s.size != N_FEAT:
            raise ValueError(f"parameter sizes must be {N_PARAMS}, 16, 16")
        self.ctx._check(lib().gcv_mlp_set_params(self.h, _ptr(p), _ptr(m), _ptr(s)))

    def get_params(self):
        p = np.empty(N_PARAMS, np.float32)
        m = np.empty(N_FEAT, np.float32)
        s = np.empty(N_FEAT, np.float32)
        self.ctx._check(lib().gcv_mlp_get_params(self.h, _ptr(p), _ptr(m), _ptr(s)))
        return p, m, s

    def save(self, path: str = WEIGHTS, **meta):
        p, m, s = self.get_params()
        np.savez(path, params=p, mean=m, istd=s, meta=np.array(json.dumps(meta)))

    def features(self, eng: ArclengthCalculateGpu) -> np.ndarray:
        n = eng.n_intervals
        d = self.ctx.empty(n * N_FEAT, np.float32)
        try:
            self.ctx._check(lib().gcv_mlp_features(eng.h, d.ptr))
            return d.download().reshape(n, N_FEAT)
        finally:
            d.free()

    def predict(self, eng: ArclengthCalculateGpu, hedge: bool = False):
        """(k int8 [n_intervals], raw outputs float [n_intervals][4])."""
        n = eng.n_intervals
        dk = self.ctx.empty(n, np.int8)
        dr = self.ctx.empty(n * N_OUT, np.float32)
        try:
            self.ctx._check(lib().gcv_mlp_predict_k(self.h, eng.h, 1 if hedge else 0, dk.ptr, dr.ptr))
            return dk.download(), dr.download().reshape(n, N_OUT)
        finally:
            dk.free()
            dr.free()

    def forward(self, d_x: DeviceArray | int, n: int) -> np.ndarray:
        d = self.ctx.empty(n * N_OUT, np.float32)
        try:
            self.ctx._check(lib().gcv_mlp_forward(self.h, _dptr(d_x), n, d.ptr))
            return d.download().reshape(n, N_OUT)
        finally:
            d.free()

    def grad(self, d_x, d_y, n: int, w_pass: float):
        g = np.empty(N_PARAMS, np.float32)
        loss = np.zeros(1, np.float32)
        self.ctx._check(lib().gcv_mlp_grad(self.h, _dptr(d_x), _dptr(d_y), n, w_pass, _ptr(g), _ptr(loss)))
        return float(loss[0]) / n, g

    def train_step(self, d_x, d_y, n: int, lr: float, w_pass: float, want_loss: bool = False):
        loss = np.zeros(1, np.float32)
        self.ctx._check(lib().gcv_mlp_train_step(self.h, _dptr(d_x), _dptr(d_y), n, lr, w_pass,
                                                 _ptr(loss) if want_loss else None))
        return float(loss[0]) / n if want_loss else None

    def close(self):
        if getattr(self, "h", None):
            lib().gcv_mlp_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def _dptr(x) -> int:
    return x.ptr if isinstance(x, DeviceArray) else int(x)


# ---------------------------------------------------------------------------
# Labels and training
# ---------------------------------------------------------------------------
def cost_matrix(ctx: Context, batch, precision=GCV_FP32, with_start: bool = False, **opts):
    """Per-interval quadrature counts of uniform-k builds, k = 0..6, and the
    safety floor: (cost int64 [7][n_intervals], floor int8 [n_intervals]);
    with_start adds the log2 quintic-start errors of those builds (float
    [7][n_intervals]; NaN for ellipses, whose sampler picks its start
    analytically)."""
    with ArclengthCalculateGpu(ctx, batch, precision=precision, normalize=True, **opts) as eng:
        n = eng.n_intervals
        d_cost = ctx.empty(n, np.int32)
        cost = np.empty((MAX_K + 1, n), np.int64)
        start = np.full((MAX_K + 1, n), np.nan, np.float32)
        for k in range(MAX_K + 1):
            eng.createArclengthTableUniform(k, d_cost)
            cost[k] = d_cost.download()
            if with_start and batch.kind != "ellipse":
                start[k] = eng.start_labels()
        d_k = ctx.array(np.zeros(n, np.int8))
        eng.createArclengthTableSeeded(d_k)
        floor = eng.seed_floor()
        d_k.free()
        d_cost.free()
    return (cost, floor, start) if with_start else (cost, floor)


REFINE_WEIGHT = int(os.environ.get("GCV_SEED_REFINE_WEIGHT", "6"))


def labels_from_costs(cost: np.ndarray, floor: np.ndarray, policy: str = "weighted") -> np.ndarray:
    """Per-interval targets from the uniform-k quadrature counts.

    y[:, 0]: with policy "clean", the smallest k at or above the floor whose
    2^k roots all pass the level-0 test (3 quadratures per root; k = 6 when
    none does), so the table has almost no refined roots and the sampler no
    tree re-walks; with "cheapest", the k of fewest quadratures; with
    "weighted", the k of least cost when a refinement quadrature counts
    REFINE_WEIGHT level-0 ones (refinement passes run few lanes and the
    sampler re-walks refined roots).
    y[:, 1] = 1 when the interval needs no refinement at that k.
    """
    ks = np.arange(MAX_K + 1)[:, None]
    ok = ks >= floor[None, :].astype(np.int64)
    clean = (cost == 3 * (1 << ks)) & ok
    if policy == "clean":
        k = np.where(clean.any(0), np.argmax(clean, axis=0), MAX_K)
    elif policy == "cheapest":
        k = np.argmin(np.where(ok, cost, np.iinfo(np.int64).max), axis=0)
    elif policy == "weighted":
        base = 3 * (1 << ks)
        w = base + REFINE_WEIGHT * np.maximum(cost - base, 0)
        k = np.argmin(np.where(ok, w, np.iinfo(np.int64).max), axis=0)
    else:
        raise ValueError(f"unknown label policy {policy}")
    c = cost[k, np.arange(cost.shape[1])]
    return np.stack([k.astype(np.float32), (c == 3 * (1 << k)).astype(np.float32)], axis=1)


def start_targets(log2_err: np.ndarray) -> np.ndarray:
    """Targets of outputs 2 and 3 from log2 quintic-start errors: the error
    mapped to about [-2.5, 2.5] and the direct-start flag (NaN stays NaN)."""
    e = np.asarray(log2_err, np.float32)
    y2 = (e + 20.0) / 8.0
    y3 = np.where(np.isnan(e), np.nan, (e <= START_LOG2_OK).astype(np.float32))
    return np.stack([y2, y3], axis=1).astype(np.float32)


def training_batches(seed: int = 100):
    """Synthetic batches spanning the classes the predictor serves (about 2.5M
    intervals): cubic 3-D random walks with near-cusps, mixed degrees 2/3/5 in
    2-D/3-D with near-collinear loops, degrees 1/4, few scattered control
    points with random knots (the reference dataset's regime, cubic 2-D and
    mixed), and ellipses. The seeds differ from every benchmark and test
    batch."""
    from . import workloads as W
    out = []
    for j in range(3):
        out.append(W.bspline_batch(2500, seed=seed + j, degrees=(3,), dims=(3,), n_ctrl_range=(8, 2048),
                                   cusp_frac=0.1))
    for j in range(4):
        out.append(W.bspline_batch(1500, seed=seed + 5 + j, degrees=(2, 3, 5), dims=(2, 3),
                                   n_ctrl_range=(16, 1024), hairpin_frac=0.2))
    out.append(W.bspline_batch(1000, seed=seed + 11, degrees=(1, 4), dims=(2, 3), n_ctrl_range=(8, 256),
                               hairpin_frac=0.1))
    out.append(W.bspline_scattered(30000, seed=seed + 30, degrees=(3,), dims=(2,)))
    out.append(W.bspline_scattered(30000, seed=seed + 31, degrees=(2, 3, 5), dims=(2, 3), n_ctrl_range=(4, 32)))
    out.append(W.ellipse_batch(200000, seed=seed + 20))
    return out


def make_dataset(ctx: Context, batches, mlp: SeedMLP, policy: str = "weighted"):
    """Feature rows and targets [n][4]: the seed labels, then the start targets
    measured on the uniform-k table of each interval's label k."""
    X, Y = [], []
    for b in batches:
        cost, floor, start = cost_matrix(ctx, b, with_start=True)
        with ArclengthCalculateGpu(ctx, b, normalize=True) as eng:
            X.append(mlp.features(eng))
        y = labels_from_costs(cost, floor, policy)
        k = y[:, 0].astype(np.int64)
        Y.append(np.concatenate([y, start_targets(start[k, np.arange(k.size)])], axis=1))
    return np.concatenate(X), np.concatenate(Y)


def standardisation(X: np.ndarray):
    mean = X.mean(0).astype(np.float32)
    std = X.std(0)
    return mean, (1.0 / np.maximum(std, 1e-3)).astype(np.float32)


def train(ctx: Context, X: np.ndarray, Y: np.ndarray, epochs: int = 30, batch: int = 16384, lr: float = 3e-3,
          w_pass: float = 0.5, seed: int = 0, log=print) -> SeedMLP:
    mean, istd = standardisation(X)
    mlp = SeedMLP(ctx, he_init(seed), mean, istd)
    Xs = bf16_bits((X - mean) * istd)
    rng = np.random.default_rng(seed)
    n = X.shape[0]
    d_x = ctx.empty(n * N_FEAT, np.uint16)
    d_y = ctx.empty(n * N_OUT, np.float32)
    try:
        for ep in range(epochs):
            perm = rng.permutation(n)
            d_x.upload(np.ascontiguousarray(Xs[perm]))
            d_y.upload(np.ascontiguousarray(Y[perm]))
            step_lr = lr * (0.5 * (1 + np.cos(np.pi * ep / epochs)))
            tot = 0.0
            for s in range(0, n - batch + 1, batch):
                last = s + 2 * batch > n
                l = mlp.train_step(d_x.ptr + s * N_FEAT * 2, d_y.ptr + s * N_OUT * 4, batch, step_lr, w_pass,
                                   want_loss=last)
                if last:
                    tot = l
            if log:
                log(f"epoch {ep:3d} lr {step_lr:.2e} last-batch loss {tot:.4f}")
    finally:
        d_x.free()
        d_y.free()
    return mlp


def _main():
    ap = argparse.ArgumentParser(description="train the learned seed")
    ap.add_argument("--train", action="store_true")
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--policy", default="weighted", choices=("clean", "cheapest", "weighted"))
    a = ap.parse_args()
    if not a.train:
        ap.print_help()
        return
    from . import workloads as W  # noqa: F401
    with Context(0) as ctx:
        t0 = time.time()
        probe = SeedMLP(ctx)
        X, Y = make_dataset(ctx, training_batches(), probe, a.policy)
        probe.close()
        print(f"dataset {X.shape[0]} intervals in {time.time() - t0:.1f}s; label histogram "
              f"{np.bincount(Y[:, 0].astype(int), minlength=7).tolist()}; direct-start share "
              f"{np.nanmean(Y[:, 3]):.3f} of {int(np.sum(~np.isnan(Y[:, 3])))} B-spline intervals", flush=True)
        mlp = train(ctx, X, Y, epochs=a.epochs)
        mlp.save(a.out, n_rows=int(X.shape[0]), epochs=a.epochs, policy=a.policy)
        print(f"saved {a.out}", flush=True)


if __name__ == "__main__":
    _main()
