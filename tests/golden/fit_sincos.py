"""Derivation of the polynomial coefficients of sin2_2pi
(include/gcurval/detail/curves.h), the squared sine behind the ellipse speed.

With y = 2x + shift (shift 0 or -1/2), q = rint(y) and e = y - q in
[-1/2, 1/2], h = e^2 in [0, 1/4]:
  sin^2(pi y) = sin^2(pi e) = h V(h),   V(h) = sin^2(pi sqrt(h)) / h,
V decreasing from pi^2 at h = 0 to 4 at h = 1/4. With shift = 0 this is
sin^2(2 pi x); with shift = -1/2 it is cos^2(2 pi x). h V(h) carries full
relative accuracy near its zeros and the speed
  |C'|^2 = base + D h V(h),   base = min(A^2, B^2), D = |A^2 - B^2|
adds only non-negative terms.
FP32: V of degree 5 by iteratively reweighted least squares toward minimax
relative error on a Chebyshev grid in h, then rounded to float.
FP64: V of degree 9, the Chebyshev expansion of V on h in [0, 1/4] truncated
after T_9 and converted to powers of h in 40-digit arithmetic (the first
omitted Chebyshev coefficient is 6.7e-17 of min V); sin2_2pi_polish, the
final Newton step's form, truncates after T_7 (1.9e-12).
Run: python tests/golden/fit_sincos.py
"""
import mpmath as mp
import numpy as np

mp.mp.dps = 40


def v_exact(h):
    h = mp.mpf(h)
    if h == 0:
        return mp.pi ** 2
    e = mp.sqrt(h)
    return mp.sin(mp.pi * e) ** 2 / h


def fit_fp32(deg=5):
    h = (np.cos(np.linspace(0, np.pi, 3000)) + 1) / 2 * 0.25
    y = np.array([float(v_exact(x)) for x in h])
    w = np.ones_like(h)
    A = np.vander(h, deg + 1, increasing=True)
    for _ in range(40):
        c, *_ = np.linalg.lstsq(A * w[:, None], y * w, rcond=None)
        r = np.abs(A @ c - y) / np.abs(y)
        w *= 1 + 5 * r / r.max()
    return [float(np.float32(v)) for v in c]


def cheb_fp64(deg=9, n=60):
    # Chebyshev coefficients of V((t + 1) / 8), t in [-1, 1], then T_j(8h - 1) in powers of h
    xs = [mp.cos(mp.pi * (k + mp.mpf(1) / 2) / n) for k in range(n)]
    vals = [v_exact((x + 1) / 8) for x in xs]
    cheb = [(2 if j else 1) * sum(vals[k] * mp.cos(mp.pi * j * (k + mp.mpf(1) / 2) / n) for k in range(n)) / n
            for j in range(deg + 2)]
    poly = [mp.mpf(0)] * (deg + 1)       # result, coefficients of h^i
    t_prev, t_cur = [mp.mpf(1)], [mp.mpf(-1), mp.mpf(8)]  # T_0, T_1(8h - 1)
    for j in range(deg + 1):
        tj = t_prev if j == 0 else t_cur
        if j >= 2:
            # T_j = 2 (8h - 1) T_{j-1} - T_{j-2}
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


def first_omitted(deg=9, n=60):
    """|c_{deg+1}| / min V: the truncation bound of cheb_fp64(deg)."""
    xs = [mp.cos(mp.pi * (k + mp.mpf(1) / 2) / n) for k in range(n)]
    vals = [v_exact((x + 1) / 8) for x in xs]
    j = deg + 1
    c = 2 * sum(vals[k] * mp.cos(mp.pi * j * (k + mp.mpf(1) / 2) / n) for k in range(n)) / n
    return float(abs(c) / 4)


def max_rel_error(coefs, dtype):
    h = np.linspace(0.0, 0.25, 20001)
    acc = np.zeros_like(h, dtype=dtype)
    for c in coefs[::-1]:
        acc = acc * h.astype(dtype) + dtype(c)
    ref = np.array([float(v_exact(x)) for x in h[::50]])
    return float(np.max(np.abs(acc[::50].astype(np.float64) - ref) / ref))


if __name__ == "__main__":
    v32 = fit_fp32()
    v64 = cheb_fp64()
    v64p = cheb_fp64(7)
    print("FP32 V", v32, "max rel err", max_rel_error(v32, np.float32))
    print("FP64 V", v64, "max rel err", max_rel_error(v64, np.float64), "truncation", first_omitted())
    print("FP64 polish V", v64p, "max rel err", max_rel_error(v64p, np.float64), "truncation", first_omitted(7))
