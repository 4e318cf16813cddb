"""The squared-sine polynomials of the ellipse speed (include/gcurval/detail/
curves.h, sin2_2pi) against 40-digit sin^2 / cos^2: the constants in the
device source are parsed and evaluated in the same order (y = 2x + shift,
e = y - rint(y), Horner in h = e^2, then h V(h)), in FP64 and FP32, for
shift 0 (sin^2(2 pi x)) and shift -1/2 (cos^2(2 pi x)).
Bars: 8e-16 relative for FP64 (the FP64 polish step and FP64 builds; 5.6e-16
measured, the rounding of h, the Horner steps and h V) and 3e-7 for FP32
(quadrature nodes; 1.5e-7 / 1.8e-7 measured). Exact zeros (x a multiple of 1/4) must come
out exactly zero. The derivation pins the FP64 constants, and a planted
wrong coefficient in the oracle of the polynomial fails the accuracy bar."""
import os
import re
import sys

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "..", "include", "gcurval", "detail", "curves.h")


def _constants(kind, name="sin2_2pi"):
    text = open(SRC).read()
    start = text.index(f"__device__ __forceinline__ {kind} {name}({kind} x")
    body = text[start:text.index("\n}\n", start)]
    vals = [float(v) for v in re.findall(r"-?\d+\.\d+(?:e-?\d+)?", body)]
    # fma(2.0, x, shift) comes first; the rest are V's coefficients, highest first
    assert vals[0] == 2.0
    return vals[1:]


def _fma64(a, b, c):
    # a * b + c rounded once to double (40-digit intermediate)
    mp.mp.dps = 40
    return float(mp.mpf(float(a)) * mp.mpf(float(b)) + mp.mpf(float(c)))


def _eval(x, dtype, shift, coefs=None):
    coefs = _constants("double" if dtype == np.float64 else "float") if coefs is None else coefs
    x = np.asarray(x, dtype)
    if dtype == np.float64:
        y = np.array([_fma64(2.0, v, shift) for v in x])
    else:
        y = (np.float64(2) * x.astype(np.float64) + np.float64(shift)).astype(np.float32)  # one rounding, as fmaf
    e = (y - np.rint(y)).astype(dtype)
    h = (e * e).astype(dtype)
    v = np.full_like(h, dtype(coefs[0]))
    for c in coefs[1:]:
        if dtype == np.float32:
            v = (v.astype(np.float64) * h.astype(np.float64) + np.float64(c)).astype(np.float32)  # fmaf
        else:
            v = np.array([_fma64(a, b, c) for a, b in zip(v, h)])
    return (h * v).astype(dtype)


def _truth(x, fn):
    mp.mp.dps = 40
    return np.array([float(fn(2 * mp.pi * mp.mpf(float(v))) ** 2) for v in x])


def _max_rel(got, ref):
    m = ref > 1e-30
    assert np.all(got[~m] == 0.0)
    return float(np.max(np.abs(got[m].astype(np.float64) - ref[m]) / ref[m]))


X64 = np.concatenate([np.linspace(0, 1, 1201), np.array([0.125, 0.375, 1e-9, 0.25 - 1e-9, 0.75 + 1e-12])])
X32 = np.concatenate([np.linspace(0, 1, 1501, dtype=np.float32),
                      np.random.default_rng(7).random(1500).astype(np.float32)])


def test_fp64_polynomial_within_a_few_ulp():
    assert _max_rel(_eval(X64, np.float64, 0.0), _truth(X64, mp.sin)) <= 8e-16
    assert _max_rel(_eval(X64, np.float64, -0.5), _truth(X64, mp.cos)) <= 8e-16


def test_fp32_polynomial_within_3e7():
    assert _max_rel(_eval(X32, np.float32, 0.0), _truth(X32, mp.sin)) <= 3e-7
    assert _max_rel(_eval(X32, np.float32, -0.5), _truth(X32, mp.cos)) <= 3e-7


def test_polish_polynomial_within_3e12():
    # the degree-7 form of the final FP64 Newton step of FP32 builds
    c = _constants("double", "sin2_2pi_polish")
    assert len(c) == 8
    assert _max_rel(_eval(X64, np.float64, 0.0, c), _truth(X64, mp.sin)) <= 3e-12
    assert _max_rel(_eval(X64, np.float64, -0.5, c), _truth(X64, mp.cos)) <= 3e-12
    # and a truncation one degree lower would not meet that bar
    assert _max_rel(_eval(X64, np.float64, 0.0, c[1:]), _truth(X64, mp.sin)) > 3e-12


def test_constants_are_the_derivation():
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    import fit_sincos as F
    np.testing.assert_array_equal(np.array(_constants("double")[::-1]), np.array(F.cheb_fp64()))
    np.testing.assert_array_equal(np.array(_constants("double", "sin2_2pi_polish")[::-1]),
                                  np.array(F.cheb_fp64(7)))


def test_planted_coefficient_error_is_caught():
    # a dropped last term, or a sign error, in the FP32 constants fails the bar
    c = _constants("float")
    for bad in (c[:-1] + [c[-1] * (1 + 1e-6)], [-c[0]] + c[1:]):
        assert _max_rel(_eval(X32[:400], np.float32, 0.0, bad), _truth(X32[:400], mp.sin)) > 3e-7
