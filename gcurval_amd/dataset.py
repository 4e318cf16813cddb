"""Loader of gcurval's B-spline dataset format (dataset/BSpline/Bundle_*.txt).

Format, per curve (whitespace separated): ``degree n_ctrl n_knots``, then the
n_ctrl x coordinates, the n_ctrl y coordinates and the n_knots knots.
As the reference's loader does (samples/sample2_bspline/dataset_loader.h:40-126),
each curve's control points are moved to their centroid and divided by their
largest |x| or |y|; 3-D output sets z = 0; values are held in the engine's
data type (``f32=True`` rounds them to float, the reference's FP32 build).
The reference reads at most 1000 curves per file, in file order.
"""
from __future__ import annotations

import os

import numpy as np

from . import BSplineBatch

CURVES_PER_FILE = 1000


def parse_curves(text: str):
    """Raw curves of one file: list of (degree, x, y, knots) float64 arrays."""
    tok = text.split()
    out = []
    i = 0
    while i < len(tok):
        p, nc, nk = int(tok[i]), int(tok[i + 1]), int(tok[i + 2])
        i += 3
        vals = np.array(tok[i:i + 2 * nc + nk], dtype=np.float64)
        if vals.size != 2 * nc + nk:
            raise ValueError("truncated curve record")
        i += 2 * nc + nk
        if nk != nc + p + 1:
            raise ValueError(f"curve with {nc} control points and {nk} knots is not a B-spline of degree {p}")
        out.append((p, vals[:nc], vals[nc:2 * nc], vals[2 * nc:]))
    return out


def normalise(x: np.ndarray, y: np.ndarray, dtype=np.float64):
    """Centroid to the origin, then divide by the largest |x| or |y|, in the
    data type's arithmetic (dataset_loader.h:70-96)."""
    x = x.astype(dtype)
    y = y.astype(dtype)
    cx = x.sum(dtype=dtype) / dtype(x.size)
    cy = y.sum(dtype=dtype) / dtype(y.size)
    x = x - cx
    y = y - cy
    m = max(np.max(np.abs(x)), np.max(np.abs(y)))
    if m > 0:
        x = x / m
        y = y / m
    return x, y


def load_bspline_dataset(paths, n_curve: int | None = None, dim: int = 2, f32: bool = True,
                         normalize: bool = True) -> BSplineBatch:
    """Curves of the given files, up to 1000 per file, at most n_curve in all,
    as a BSplineBatch with ctrl_stride = dim (2, or 3 with z = 0)."""
    if dim not in (2, 3):
        raise ValueError("dim must be 2 or 3")
    dtype = np.float32 if f32 else np.float64
    deg, knots, ctrl, koff, coff = [], [], [], [0], [0]
    for path in paths:
        if n_curve is not None and len(deg) >= n_curve:
            break
        with open(path) as f:
            curves = parse_curves(f.read())[:CURVES_PER_FILE]
        for p, x, y, t in curves:
            if n_curve is not None and len(deg) >= n_curve:
                break
            if normalize:
                x, y = normalise(x, y, dtype)
            P = np.zeros((x.size, dim), np.float64)
            P[:, 0] = np.asarray(x, dtype).astype(np.float64)
            P[:, 1] = np.asarray(y, dtype).astype(np.float64)
            deg.append(p)
            knots.append(np.asarray(t, dtype).astype(np.float64))
            ctrl.append(P)
            koff.append(koff[-1] + t.size)
            coff.append(coff[-1] + x.size)
    if not deg:
        raise ValueError("no curves read")
    return BSplineBatch(np.array(deg, np.int32), np.full(len(deg), dim, np.int32), np.array(koff, np.int64),
                        np.array(coff, np.int64), np.concatenate(knots), np.concatenate(ctrl))


def bundle_paths(root: str, n_files: int = 100):
    """dataset/BSpline/Bundle_0.txt .. Bundle_{n-1}.txt under root, in order."""
    return [os.path.join(root, f"Bundle_{i}.txt") for i in range(n_files)]
