"""Derivation of the polynomial coefficients of abs_sincos_2pi
(include/gcurval/detail/curves.h).

With r = 4x, q = rint(r), f = r - q in [-1/2, 1/2]:
  sin(pi f / 2) = f S(f^2),  cos(pi f / 2) = C(f^2),  g = f^2 in [0, 1/4].
FP32: S (degree 3) and C (degree 4) by iteratively reweighted least squares
toward minimax relative error on a Chebyshev grid in g, then rounded to float.
FP64: Taylor coefficients of sin(pi f/2)/f (9 terms) and cos(pi f/2) (10 terms).
Run: python -m gcurval_amd.fit_sincos
"""
import mpmath as mp
import numpy as np

mp.mp.dps = 40


def fit_fp32():
    g = (np.cos(np.linspace(0, np.pi, 3000)) + 1) / 2 * 0.25
    g = g[g > 0]
    f = np.sqrt(g)
    ys = np.array([float(mp.sin(mp.pi / 2 * mp.mpf(x)) / mp.mpf(x)) for x in f])
    yc = np.array([float(mp.cos(mp.pi / 2 * mp.mpf(x))) for x in f])
    out = []
    for y, deg in ((ys, 3), (yc, 4)):
        w = np.ones_like(g)
        A = np.vander(g, deg + 1, increasing=True)
        for _ in range(40):
            c, *_ = np.linalg.lstsq(A * w[:, None], y * w, rcond=None)
            r = np.abs(A @ c - y) / np.abs(y)
            w *= 1 + 5 * r / r.max()
        out.append([float(np.float32(v)) for v in c])
    return out


def taylor_fp64():
    h = mp.pi / 2
    S = [float((-1) ** k * h ** (2 * k + 1) / mp.factorial(2 * k + 1)) for k in range(9)]
    C = [float((-1) ** k * h ** (2 * k) / mp.factorial(2 * k)) for k in range(10)]
    return S, C


if __name__ == "__main__":
    s32, c32 = fit_fp32()
    s64, c64 = taylor_fp64()
    print("FP32 S", s32)
    print("FP32 C", c32)
    print("FP64 S", s64)
    print("FP64 C", c64)
