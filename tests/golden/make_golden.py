#!/usr/bin/env python3
"""Writes tests/golden/golden.json: closed-form and library-routine vectors
that pin the oracle (the reference holds no test vectors of its own).

  * Gauss-Legendre nodes/weights of orders 1..5 (numpy leggauss).
  * Ellipse perimeters 4 a E(1 - b^2/a^2) and partial arcs
    s(t) = a [E(m) - E(pi/2 - t | m)] (mpmath, 30 digits).
  * Cubic / quadratic / quintic B-spline first derivatives at fixed
    parameters (scipy.interpolate.BSpline.derivative).
  * B-spline lengths by mpmath.quad split at the knots (tight tolerance).
Run from the repository root: python tests/golden/make_golden.py
"""
import json
import os

import mpmath as mp
import numpy as np
from scipy.interpolate import BSpline

mp.mp.dps = 30
HERE = os.path.dirname(os.path.abspath(__file__))


def ellipse_perimeter(a, b):
    a, b = mp.mpf(a), mp.mpf(b)
    if a >= b:
        return 4 * a * mp.ellipe(1 - (b / a) ** 2)
    return 4 * b * mp.ellipe(1 - (a / b) ** 2)


def ellipse_partial(a, b, t):
    """Arc length of (a cos s, b sin s) for s in [0, t]."""
    a, b, t = mp.mpf(a), mp.mpf(b), mp.mpf(t)
    return mp.quad(lambda s: mp.sqrt((a * mp.sin(s)) ** 2 + (b * mp.cos(s)) ** 2),
                   mp.linspace(0, t, 9))


def bspline_case(p, D, n, seed):
    rng = np.random.default_rng(seed)
    inner = np.sort(rng.uniform(0.05, 0.95, n - p - 1))
    knots = np.r_[np.zeros(p + 1), inner, np.ones(p + 1)]
    ctrl = rng.normal(size=(n, D))
    knots = knots.astype(np.float32).astype(np.float64)
    ctrl = ctrl.astype(np.float32).astype(np.float64)
    spl = BSpline(knots, ctrl, p)
    der = spl.derivative()
    us = [0.013, 0.21, 0.5, 0.77, 0.999]
    vals = [der(u).tolist() for u in us]

    def speed(u):
        u = float(u)
        return mp.sqrt(sum(mp.mpf(float(v)) ** 2 for v in der(u)))

    brk = sorted(set(knots.tolist()))
    L = mp.quad(speed, brk)
    return {"p": p, "D": D, "knots": knots.tolist(), "ctrl": ctrl.tolist(), "u": us, "deriv": vals,
            "length": mp.nstr(L, 20)}


def main():
    out = {"gl": {}, "ellipse": [], "ellipse_partial": [], "bspline": []}
    for n in range(1, 6):
        x, w = np.polynomial.legendre.leggauss(n)
        out["gl"][str(n)] = {"x": x.tolist(), "w": w.tolist()}
    for a, b in [(3.0, 1.0), (1.0, 1.0), (1.0, 0.5), (10.0, 0.1), (0.5, 0.005), (2.5, 2.0)]:
        out["ellipse"].append({"a": a, "b": b, "L": mp.nstr(ellipse_perimeter(a, b), 25)})
    for a, b, t in [(3.0, 1.0, 1.0), (3.0, 1.0, 4.0), (1.0, 0.01, 0.3)]:
        out["ellipse_partial"].append({"a": a, "b": b, "t": t, "s": mp.nstr(ellipse_partial(a, b, t), 25)})
    seed = 11
    for p, D, n in [(1, 2, 6), (2, 2, 7), (3, 3, 9), (3, 2, 12), (5, 3, 10), (4, 2, 8)]:
        out["bspline"].append(bspline_case(p, D, n, seed))
        seed += 1
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", os.path.join(HERE, "golden.json"))


if __name__ == "__main__":
    main()
