"""Derivation of the polynomial coefficients of sin2_quarter
(include/gcurval/detail/curves.h), the squared sine behind the ellipse speed.

With r = 4x, q = rint(r), f = r - q in [-1/2, 1/2] and g = f^2 in [0, 1/4]:
  sin^2(pi f / 2) = g U(g),   U(g) = sin^2(pi f / 2) / f^2,
and sin^2(2 pi x), cos^2(2 pi x) are g U(g) and 1 - g U(g), swapped when q is
odd. g U(g) <= 1/2, so both squares keep full relative accuracy near zero.
FP32: U of degree 4 by iteratively reweighted least squares toward minimax
relative error on a Chebyshev grid in g, then rounded to float.
FP64: Taylor coefficients of U (11 terms; the first omitted term is below
2e-17 relative on g <= 1/4).
Run: python -m gcurval_amd.fit_sincos
"""
import mpmath as mp
import numpy as np

mp.mp.dps = 40


def u_exact(g):
    f = mp.sqrt(mp.mpf(g))
    return mp.sin(mp.pi / 2 * f) ** 2 / f ** 2


def fit_fp32(deg=4):
    g = (np.cos(np.linspace(0, np.pi, 3000)) + 1) / 2 * 0.25
    g = g[g > 0]
    y = np.array([float(u_exact(x)) for x in g])
    w = np.ones_like(g)
    A = np.vander(g, deg + 1, increasing=True)
    for _ in range(40):
        c, *_ = np.linalg.lstsq(A * w[:, None], y * w, rcond=None)
        r = np.abs(A @ c - y) / np.abs(y)
        w *= 1 + 5 * r / r.max()
    return [float(np.float32(v)) for v in c]


def taylor_fp64(terms=11):
    # sin^2 t = sum_{k>=1} (-1)^(k+1) 2^(2k-1) t^(2k) / (2k)!,  t^2 = (pi/2)^2 g
    h2 = (mp.pi / 2) ** 2
    return [float((-1) ** (k + 1) * 2 ** (2 * k - 1) * h2 ** k / mp.factorial(2 * k)) for k in range(1, terms + 1)]


def max_rel_error(coefs, dtype):
    g = np.linspace(1e-6, 0.25, 20001)
    acc = np.zeros_like(g, dtype=dtype)
    for c in coefs[::-1]:
        acc = acc * g.astype(dtype) + dtype(c)
    ref = np.array([float(u_exact(x)) for x in g[::50]])
    return float(np.max(np.abs(acc[::50].astype(np.float64) - ref) / ref))


if __name__ == "__main__":
    u32 = fit_fp32()
    u64 = taylor_fp64()
    print("FP32 U", u32, "max rel err", max_rel_error(u32, np.float32))
    print("FP64 U", u64, "max rel err", max_rel_error(u64, np.float64))
