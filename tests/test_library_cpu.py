"""CPU-only checks of the C ABI library: it loads without a GPU, exports every
symbol include/gcurval_c.h declares, and its host build of the B-spline setup
routine (power-basis derivative per span) matches scipy."""
import json
import os
import re

import numpy as np
import pytest

import gcurval_amd as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "golden.json")))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "gcurval_c.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)  # drop comments
    decl = re.compile(r"^\s*(?:const\s+)?(?:int|void|char)\s*\*?\s*(gcv_[a-z0-9_]+)\s*\(", re.M)
    return sorted(set(decl.findall(txt)))


def test_library_loads_and_exports_all_declared_symbols():
    L = G.load_library()
    syms = declared_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # the ctypes mirror binds exactly the declared entry points
    assert sorted(G.exported_symbols()) == syms


def test_version_and_defaults():
    L = G.load_library()
    assert b"gfx950" in L.gcv_version()
    o = G.Opts()
    L.gcv_opts_default(o)
    assert (o.precision, o.gl_order, o.n_subdivided, o.max_depth, o.tolerance, o.normalize) == (0, 3, 64, 0, 0.0, 0)


@pytest.mark.parametrize("k", range(6))
def test_host_span_poly_matches_scipy_derivative(k):
    case = GOLD["bspline"][k]
    p, D = case["p"], case["D"]
    knots = np.array(case["knots"])
    ctrl = np.zeros((len(case["ctrl"]), 3))
    ctrl[:, :D] = np.array(case["ctrl"])
    for u, d in zip(case["u"], case["deriv"]):
        s = int(np.searchsorted(knots, u, side="right") - 1)
        coef, u0, h = G.host_bspline_span_poly(p, D, s, knots, ctrl)
        x = (u - u0) / h
        v = np.array([np.polyval(coef[dd][::-1], x) for dd in range(D)]) / h
        np.testing.assert_allclose(v, d, rtol=1e-11, atol=1e-12)


def test_host_span_poly_rejects_bad_degree():
    with pytest.raises(G.GcvError):
        G.host_bspline_span_poly(6, 3, 6, np.zeros(20), np.zeros((13, 3)))


def test_degenerate_span_gives_zero_polynomial():
    knots = np.array([0, 0, 0, 0, 0.5, 0.5, 1, 1, 1, 1.0])
    ctrl = np.random.default_rng(0).normal(size=(6, 3))
    with pytest.raises(G.GcvError):  # span 4 has zero width
        G.host_bspline_span_poly(3, 3, 4, knots, ctrl)


def test_sample_offsets_rejects_bad_counts_and_positions():
    off, pos = G.sample_offsets([2, 0, 3], 3)
    assert off.tolist() == [0, 2, 2, 5] and pos is None
    off, pos = G.sample_offsets([1, 2], 2, [0.0, 0.5, 1.0])
    assert pos.dtype == np.float64 and pos.size == 3
    with pytest.raises(ValueError, match="negative"):
        G.sample_offsets([2, -1, 3], 3)
    with pytest.raises(ValueError, match="per-curve counts"):
        G.sample_offsets([2, 3], 3)
    with pytest.raises(ValueError, match="positions expected"):
        G.sample_offsets([2, 3], 2, np.zeros(4))
