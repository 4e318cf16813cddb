"""The squared-sine polynomials of the ellipse speed (include/gcurval/detail/
curves.h, sincos2_2pi) against 40-digit sin^2 / cos^2: the constants in the
device source are parsed and evaluated in the same order (Horner in g, then
the quadrant swap), in FP64 and FP32. Bars: 4e-16 relative for FP64 (the FP64
polish step and FP64 builds; about 2 ulp, the rounding of g = f^2 and of the
Horner steps, the same for the degree-7 form as for the 11-term Taylor series)
and 2e-7 for FP32 (quadrature nodes)."""
import os
import re

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "..", "include", "gcurval", "detail", "curves.h")


def _constants(kind):
    text = open(SRC).read()
    start = text.index(f"__device__ __forceinline__ void sincos2_2pi({kind} x")
    body = text[start:text.index("\n}\n", start)]
    return [float(v) for v in re.findall(r"-?\d+\.\d+(?:e-?\d+)?", body)]


def _eval(x, dtype):
    consts = _constants("double" if dtype == np.float64 else "float")
    x = np.asarray(x, dtype)
    r = dtype(4) * x
    q = np.rint(r)
    f = r - q
    g = f * f
    if dtype == np.float64:
        # u = c0; u = fma(u, g, c1) ...: the leading coefficient comes first
        # after the constants 4.0 / 1.0 used elsewhere in the body
        coefs = [c for c in consts if c not in (4.0, 1.0)]
        u = np.full_like(g, coefs[0])
        for c in coefs[1:]:
            u = u * g + dtype(c)
    else:
        coefs = [c for c in consts if c not in (4.0, 1.0)]
        # fmaf(fmaf(fmaf(fmaf(c4, g, c3), g, c2), g, c1), g, c0) lists c4 .. c0 in order
        u = np.full_like(g, dtype(coefs[0]))
        for c in coefs[1:]:
            u = (u * g + dtype(c)).astype(dtype)
    lo = g * u
    hi = dtype(1) - lo
    odd = (q.astype(np.int64) & 1) == 1
    return np.where(odd, hi, lo), np.where(odd, lo, hi)


def _truth(x):
    mp.mp.dps = 40
    s = np.array([float(mp.sin(2 * mp.pi * mp.mpf(float(v))) ** 2) for v in x])
    c = np.array([float(mp.cos(2 * mp.pi * mp.mpf(float(v))) ** 2) for v in x])
    return s, c


def test_fp64_polynomial_within_2_ulp():
    x = np.concatenate([np.linspace(0, 1, 2001), np.array([0.125, 0.375, 1e-9, 0.25 - 1e-9])])
    s2, c2 = _eval(x, np.float64)
    st, ct = _truth(x)
    for got, ref in ((s2, st), (c2, ct)):
        m = ref > 1e-30  # exact zeros (x a multiple of 1/4) are matched exactly
        assert np.max(np.abs(got[m] - ref[m]) / ref[m]) <= 4e-16
        assert np.all(got[~m] == 0.0)


def test_fp32_polynomial_within_2e7():
    x = np.linspace(0, 1, 1501, dtype=np.float32)
    s2, c2 = _eval(x, np.float32)
    st, ct = _truth(x)
    for got, ref in ((s2, st), (c2, ct)):
        m = ref > 1e-30
        assert np.max(np.abs(got[m].astype(np.float64) - ref[m]) / ref[m]) <= 2e-7
        assert np.all(got[~m] == 0.0)


def test_constants_are_the_derivation():
    from gcurval_amd import fit_sincos as F
    c64 = [c for c in _constants("double") if c not in (4.0, 1.0)]
    np.testing.assert_array_equal(np.array(c64[::-1]), np.array(F.cheb_fp64()))
