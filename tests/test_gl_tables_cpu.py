"""The device Gauss-Legendre tables (include/gcurval/detail/device.h, GL<N>)
against numpy's leggauss: nodes and weights within 2e-15 relative (the
tables hold the textbook decimals; leggauss's own values are a few ulp off
the correctly rounded ones), the
reference's node order for N <= 5 (centre node first for odd N), and GL<8>,
the cross-check rule of GL-5 seeded builds, exact on polynomials of degree 15."""
import os
import re

import numpy as np
import pytest

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "gcurval", "detail",
                   "device.h")


def _table(n):
    text = open(HDR).read()
    m = re.search(r"template <> struct GL<%d> \{(.*?)\n\};" % n, text, re.S)
    assert m, n
    body = m.group(1)
    def arr(name):
        a = re.search(r"%s\[%d\] = \{(.*?)\}" % (name, n), body, re.S).group(1)
        return np.array([float(v) for v in a.replace("\n", " ").split(",")])
    return arr("x"), arr("w")


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 8])
def test_device_gl_table_matches_leggauss(n):
    x, w = _table(n)
    xr, wr = np.polynomial.legendre.leggauss(n)
    o = np.argsort(x)
    np.testing.assert_allclose(x[o], xr, rtol=0, atol=2e-15)
    np.testing.assert_allclose(w[o], wr, rtol=2e-15, atol=0)
    if n % 2 == 1:
        assert x[0] == 0.0  # centre node first, as the reference orders odd rules


def test_gl8_exact_to_degree_15():
    x, w = _table(8)
    for d in range(16):
        exact = 0.0 if d % 2 else 2.0 / (d + 1)
        assert abs(np.sum(w * x ** d) - exact) <= 1e-15, d
    assert abs(np.sum(w * x ** 16) - 2.0 / 17) > 1e-10
