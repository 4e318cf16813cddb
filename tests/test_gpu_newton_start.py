"""GPU checks of the learned Newton start of the B-spline sampler.

* The start labels (log2 of the worst quintic Hermite start error over each
  interval's roots) against the same definition restated in FP64 numpy, and,
  on intervals whose speed varies mildly, against the exact inverse
  arc-length map (exact end speeds and slopes from scipy's B-spline
  derivatives, a dense Simpson cumulative of the speed).
* The three Newton starts (plain, learned, quintic everywhere) against the
  long-double truth (<= 1e-6) and against each other, and the gain of the
  learned start: fewer FP32 Newton steps per sample than the plain start.
"""
import os

import numpy as np
import pytest
from scipy.interpolate import BSpline

import gcurval_amd as G
from gcurval_amd import seed as S
from gcurval_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


_GX = np.array([-np.sqrt(0.6), 0.0, np.sqrt(0.6)])
_GW = np.array([5.0, 8.0, 5.0]) / 9.0


def _quintic(w, L, sig, va, vb, sa, sb):
    """Quintic Hermite start (offset from the root's start, clamped to it)."""
    t = sig
    t2, t3, t4, t5 = t ** 2, t ** 3, t ** 4, t ** 5
    H5 = 10 * t3 - 15 * t4 + 6 * t5
    H1 = t - 6 * t3 + 8 * t4 - 3 * t5
    H4 = -4 * t3 + 7 * t4 - 3 * t5
    H2 = 0.5 * t2 - 1.5 * t3 + 1.5 * t4 - 0.5 * t5
    H3 = 0.5 * t3 - t4 + 0.5 * t5
    xa2, xb2 = -sa / va ** 3, -sb / vb ** 3
    return np.clip(w * H5 + L * (H1 / va + H4 / vb) + L * L * (H2 * xa2 + H3 * xb2), 0.0, w)


def _cpu_start_log2(batch, k):
    """Per interval of a 2^k-root table, log2 of the worst quintic-start error
    (relative to root width) at 1/4, 1/2, 3/4 of each root's length:
    * rule: against the converged solution of the single 3-point Gauss-Legendre
      partial integral the sampler's FP64 step solves, with the root length
      from the two half-root rules, on the span's power-basis coefficients
      (the library's host routine, rounded to float as the FP32 table holds
      them), all in FP64 numpy;
    * exact: against the exact inverse arc-length map (dense Simpson
      cumulative of scipy's derivative), with exact end speeds and slopes;
    and the interval's min/max speed ratio."""
    rule, exact, ratio = [], [], []
    sig = np.array([0.25, 0.5, 0.75])
    for c in range(batch.n_curve):
        p, D = int(batch.degree[c]), int(batch.dim[c])
        T = batch.knots[batch.knot_off[c]:batch.knot_off[c + 1]]
        P = batch.ctrl[batch.ctrl_off[c]:batch.ctrl_off[c + 1], :D]
        d1 = BSpline(T, P, p).derivative()
        d2 = d1.derivative() if p >= 2 else None
        for s in range(p, len(T) - p - 1):
            u0, h = T[s], T[s + 1] - T[s]
            cf, _, _ = G.host_bspline_span_poly(p, D, s, T, P)
            cf = cf.astype(np.float32).astype(np.float64)
            polys = [cf[d][::-1] for d in range(D)]
            dpolys = [np.polyder(q) for q in polys]

            def v_rule(x):
                return np.sqrt(sum(np.polyval(q, x) ** 2 for q in polys))

            def s_rule(x):
                return sum(np.polyval(q, x) * np.polyval(dq, x) for q, dq in zip(polys, dpolys)) / v_rule(x)

            def quad(a, d):
                return d / 2 * np.sum(_GW * v_rule(a + d / 2 + d / 2 * _GX))

            def slope(x):
                if d2 is None:
                    return 0.0
                # the span's own one-sided value at its right end (C'' jumps
                # at the knots of a quadratic)
                u = min(u0 + h * x, np.nextafter(T[s + 1], -np.inf))
                c1, c2 = d1(u), d2(u)
                return h * h * float(np.dot(c1, c2)) / np.linalg.norm(c1)

            xs = np.linspace(0, 1, 257)
            vs = np.linalg.norm(d1(u0 + h * xs), axis=1)
            ratio.append(vs.min() / vs.max())
            we, wr = 0.0, 0.0
            for r in range(1 << k):
                a, b = r / (1 << k), (r + 1) / (1 << k)
                w = b - a
                # rule
                L = quad(a, w / 2) + quad(a + w / 2, w / 2)
                va, vb = v_rule(a), v_rule(b)
                for t, d0 in zip(sig, _quintic(w, L, sig, va, vb, s_rule(a), s_rule(b))):
                    tau, d, lo, hi = t * L, d0, 0.0, w
                    for _ in range(60):
                        F = quad(a, d) - tau
                        lo, hi = (lo, d) if F > 0 else (d, hi)
                        v = v_rule(a + d)
                        dn = d - F / v if v > 0 else 0.5 * (lo + hi)
                        dn = dn if lo <= dn <= hi else 0.5 * (lo + hi)
                        done = abs(dn - d) <= w * 2.0 ** -45
                        d = dn
                        if done:
                            break
                    wr = max(wr, abs(d0 - d) / w)
                # exact
                xf = np.linspace(a, b, 4001)
                vf = np.linalg.norm(d1(u0 + h * xf), axis=1) * h
                if vf.min() <= 1e-6 * vf.max():
                    we = np.nan
                    continue
                sf = np.concatenate([[0], np.cumsum((vf[1:] + vf[:-1]) / 2 * np.diff(xf))])
                Le = sf[-1]
                ytrue = (np.interp(sig * Le, sf, xf) - a) / w
                d = _quintic(w, Le, sig, vf[0], vf[-1], slope(a), slope(b))
                we = max(we, float(np.max(np.abs(d / w - ytrue))))
            rule.append(np.log2(max(wr, 2.0 ** -40)))
            exact.append(np.log2(max(we, 2.0 ** -40)) if we == we else np.nan)
    return np.array(rule), np.array(exact), np.array(ratio)


def test_start_labels_match_cpu(ctx):
    b = W.bspline_batch(10, seed=71, degrees=(2, 3, 5), dims=(2, 3), n_ctrl_range=(12, 40), hairpin_frac=0.3)
    k = 3
    with G.ArclengthCalculateGpu(ctx, b, normalize=True) as eng:
        eng.createArclengthTableUniform(k)
        lab = eng.start_labels()
        deep = eng.tables()[2]
    rule, exact, ratio = _cpu_start_log2(b, k)
    assert lab.shape == rule.shape and np.all(np.isfinite(lab))
    assert np.all(lab <= 0.0) and np.all(lab >= -40.0)
    # refined roots count as error 1 (label 0)
    assert np.all(lab[deep != 0] == 0.0)
    # the same definition restated in FP64 numpy: where the error is well
    # above the FP32 start arithmetic the two agree to a factor of 1.5
    m = (deep == 0) & (rule > -19) & (lab > -19)
    assert m.sum() >= 20, m.sum()
    assert np.max(np.abs(lab[m] - rule[m])) <= 0.6, np.max(np.abs(lab[m] - rule[m]))
    small = (deep == 0) & (rule < -24)
    assert np.all(lab[small] < -19)
    # on intervals whose speed varies by less than 2x the rule's partial
    # integrals track the exact arc length, and so do the labels
    sm = (deep == 0) & (ratio > 0.5) & np.isfinite(exact) & (exact > -19) & (lab > -19)
    assert sm.sum() >= 10, sm.sum()
    assert np.max(np.abs(lab[sm] - exact[sm])) <= 0.6, np.max(np.abs(lab[sm] - exact[sm]))


def test_start_labels_reject_other_tables(ctx):
    with G.ArclengthCalculateGpu(ctx, W.ellipse_batch(4, seed=3), normalize=True) as eng:
        eng.createArclengthTable()
        with pytest.raises(G.GcvError):
            eng.start_labels()


def _pretrained(ctx):
    if not os.path.exists(S.WEIGHTS):
        pytest.skip("no trained weights in the tree")
    return S.SeedMLP.pretrained(ctx)


def test_learned_start_gain(ctx):
    # a C4-like batch: degrees 2/3/5 in 2-D/3-D with near-collinear loops
    b = W.bspline_batch(240, seed=73, degrees=(2, 3, 5), dims=(2, 3), n_ctrl_range=(16, 400), hairpin_frac=0.2)
    counts = 16 * b.intervals_per_curve()
    n = int(counts.sum())
    res = {}
    with _pretrained(ctx) as m, G.ArclengthCalculateGpu(ctx, b, normalize=True) as eng:
        eng.createArclengthTableLearned(m)
        _, raw = m.predict(eng)
        d_it = ctx.empty(n, np.int8)
        eng.set_iters_out(d_it)
        for mode in (G.GCV_START_PLAIN, G.GCV_START_LEARNED, G.GCV_START_QUINTIC):
            eng.set_newton_start(mode)
            u, seg = eng.arclengthReverse(counts, None, True, return_segments=True)
            res[mode] = (u, d_it.download()[:n].astype(np.int64))
        eng.set_iters_out(None)
        d_it.free()
    ioff = np.concatenate([[0], np.cumsum(b.intervals_per_curve())])
    seg_int = ioff[np.repeat(np.arange(b.n_curve), counts)] + seg
    up, ip = res[G.GCV_START_PLAIN]
    ul, il = res[G.GCV_START_LEARNED]
    uq, iq = res[G.GCV_START_QUINTIC]
    n32 = {k: (v[1] & 15) for k, v in res.items()}
    n64 = {k: (v[1] >> 4) for k, v in res.items()}
    # every sample takes at least one FP64 step; the plain start always takes FP32 steps
    for k in res:
        assert np.all(n64[k] >= 1)
    assert np.all(n32[G.GCV_START_PLAIN] >= 1)
    # samples of unmarked intervals take the plain start and steps
    marked = raw[seg_int, 3] > 0
    assert marked.mean() >= 0.2, marked.mean()
    assert np.all(n32[G.GCV_START_LEARNED][~marked] >= 1)
    direct_l = n32[G.GCV_START_LEARNED] == 0
    direct_q = n32[G.GCV_START_QUINTIC] == 0
    fb_l = n64[G.GCV_START_LEARNED] == 2
    fb_q = n64[G.GCV_START_QUINTIC] == 2
    print(f"marked {marked.mean():.3f}; FP32 steps per sample plain {n32[G.GCV_START_PLAIN].mean():.3f} "
          f"learned {n32[G.GCV_START_LEARNED].mean():.3f} quintic {n32[G.GCV_START_QUINTIC].mean():.3f}; "
          f"direct learned {direct_l.mean():.3f} quintic {direct_q.mean():.3f}; "
          f"fallback learned {fb_l.mean():.4f} quintic {fb_q.mean():.4f}; "
          f"fallback share of quintic attempts: marked {fb_q[marked].mean():.4f} unmarked {fb_q[~marked].mean():.4f}")
    # the marked intervals' quintic starts seldom fail their check
    assert fb_l.sum() <= 0.05 * max(1, (direct_l | fb_l).sum()), (fb_l.sum(), direct_l.sum())
    assert fb_q[marked].mean() <= 0.05
    # the gain: fewer FP32 Newton steps per sample than the plain start
    mean_p = n32[G.GCV_START_PLAIN].mean()
    mean_l = n32[G.GCV_START_LEARNED].mean()
    assert mean_l <= 0.95 * mean_p, (mean_l, mean_p)
    # the starts agree far inside the 1e-6 bar
    assert np.max(np.abs(ul - up)) <= 1e-8, np.max(np.abs(ul - up))
    assert np.max(np.abs(uq - up)) <= 1e-8, np.max(np.abs(uq - up))
    # and all meet it against the long-double truth on a subset of curves
    sub = np.arange(0, b.n_curve, 8)
    off = O.offsets(counts)
    ko, co = b.knot_off, b.ctrl_off
    sb = G.BSplineBatch(b.degree[sub], b.dim[sub],
                      np.concatenate([[0], np.cumsum(ko[sub + 1] - ko[sub])]),
                      np.concatenate([[0], np.cumsum(co[sub + 1] - co[sub])]),
                      np.concatenate([b.knots[ko[c]:ko[c + 1]] for c in sub]),
                      np.concatenate([b.ctrl[co[c]:co[c + 1]] for c in sub]))
    cs = counts[sub]
    ut, _ = O.Truth(sb).reverse(cs, O.uniform_positions(cs), True)
    idx = np.concatenate([np.arange(off[c], off[c + 1]) for c in sub])
    for u in (up, ul, uq):
        assert np.max(np.abs(u[idx] - ut)) <= 1e-6
