"""Derivation of the polynomial coefficients of sin2_quarter
(include/gcurval/detail/curves.h), the squared sine behind the ellipse speed.

With r = 4x, q = rint(r), f = r - q in [-1/2, 1/2] and g = f^2 in [0, 1/4]:
  sin^2(pi f / 2) = g U(g),   U(g) = sin^2(pi f / 2) / f^2,
and sin^2(2 pi x), cos^2(2 pi x) are g U(g) and 1 - g U(g), swapped when q is
odd. g U(g) <= 1/2, so both squares keep full relative accuracy near zero.
FP32: U of degree 4 by iteratively reweighted least squares toward minimax
relative error on a Chebyshev grid in g, then rounded to float.
FP64: U of degree 7, the Chebyshev expansion of U on g in [0, 1/4] truncated
after T_7 and converted to powers of g in 40-digit arithmetic (the first
omitted Chebyshev coefficient is 3e-17 of U; the 11-term Taylor series reaches
the same accuracy with three more terms).
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


def cheb_fp64(deg=7, n=40):
    # Chebyshev coefficients of U((t + 1) / 8), t in [-1, 1], then T_j(8g - 1) in powers of g
    xs = [mp.cos(mp.pi * (k + mp.mpf(1) / 2) / n) for k in range(n)]
    vals = [u_exact((x + 1) / 8) for x in xs]
    cheb = [(2 if j else 1) * sum(vals[k] * mp.cos(mp.pi * j * (k + mp.mpf(1) / 2) / n) for k in range(n)) / n
            for j in range(deg + 1)]
    poly = [mp.mpf(0)] * (deg + 1)       # result, coefficients of g^i
    t_prev, t_cur = [mp.mpf(1)], [mp.mpf(-1), mp.mpf(8)]  # T_0, T_1(8g - 1)
    for j in range(deg + 1):
        tj = t_prev if j == 0 else t_cur
        if j >= 2:
            # T_j = 2 (8g - 1) T_{j-1} - T_{j-2}
            nxt = [mp.mpf(0)] * (j + 1)
            for i, v in enumerate(t_cur):
                nxt[i] -= 2 * v
                nxt[i + 1] += 16 * v
            for i, v in enumerate(t_prev):
                nxt[i] -= v
            t_prev, t_cur = t_cur, nxt
            tj = t_cur
        for i, v in enumerate(tj):
            poly[i] += cheb[j] * v
    return [float(v) for v in poly]


def max_rel_error(coefs, dtype):
    g = np.linspace(1e-6, 0.25, 20001)
    acc = np.zeros_like(g, dtype=dtype)
    for c in coefs[::-1]:
        acc = acc * g.astype(dtype) + dtype(c)
    ref = np.array([float(u_exact(x)) for x in g[::50]])
    return float(np.max(np.abs(acc[::50].astype(np.float64) - ref) / ref))


if __name__ == "__main__":
    u32 = fit_fp32()
    u64 = cheb_fp64()
    print("FP32 U", u32, "max rel err", max_rel_error(u32, np.float32))
    print("FP64 U", u64, "max rel err", max_rel_error(u64, np.float64))
