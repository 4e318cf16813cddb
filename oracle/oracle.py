"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY. Loaded by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the reference results are compared with (or as the timed CPU baseline);
never by the product package gcurval_amd.

Engines (see gcv_oracle.cpp):
  RefCpu       ArclengthCalculateCpu (gcurval/detail/cpu/arclength_calculate.h)
  RefGpuPlain  ArclengthCalculateCuda semantics, serial
  Truth        long-double adaptive 20-point Gauss-Legendre + Newton inverse
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# GCV_ORACLE_LIB selects another build of the same source (the sanitized one)
LIB_PATH = os.environ.get("GCV_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")


class OrcBatch(C.Structure):
    _fields_ = [
        ("kind", C.c_int), ("n_curve", C.c_int),
        ("a", C.c_void_p), ("b", C.c_void_p),
        ("degree", C.c_void_p), ("dim", C.c_void_p),
        ("knot_off", C.c_void_p), ("ctrl_off", C.c_void_p),
        ("knots", C.c_void_p), ("ctrl", C.c_void_p),
        ("ctrl_stride", C.c_int),
    ]


class OrcOpts(C.Structure):
    _fields_ = [
        ("precision", C.c_int), ("gl_order", C.c_int), ("n_sub", C.c_int), ("max_depth", C.c_int),
        ("tol", C.c_double), ("normalize", C.c_int),
    ]


_lib = None
VP = C.c_void_p
_SIGS = [
    ("orc_gl_table", None, [C.c_int, VP, VP]),
    ("orc_eps_device", C.c_double, [C.c_int, C.c_double, C.c_int, C.c_int]),
    ("orc_n_intervals", C.c_int64, [C.POINTER(OrcBatch)]),
    ("orc_curve_scale", C.c_double, [C.POINTER(OrcBatch), C.c_int]),
    ("orc_speed", C.c_double, [C.POINTER(OrcBatch), C.c_int, C.c_int, C.c_double, C.c_int]),
    ("orc_deriv", None, [C.POINTER(OrcBatch), C.c_int, C.c_int, C.c_double, VP]),
    ("orc_truth_deriv", None, [C.POINTER(OrcBatch), C.c_int, C.c_int, C.c_double, VP]),
    ("orc_cpu_create", VP, [C.POINTER(OrcBatch), C.POINTER(OrcOpts)]),
    ("orc_cpu_destroy", None, [VP]),
    ("orc_cpu_build", C.c_double, [VP, C.c_int]),
    ("orc_cpu_lengths", None, [VP, C.c_int, VP]),
    ("orc_cpu_table_size", C.c_int64, [VP, C.c_int]),
    ("orc_cpu_table", None, [VP, C.c_int, VP, VP, VP]),
    ("orc_cpu_reverse", C.c_double, [VP, C.c_int, VP, VP, C.c_int, VP, C.c_int]),
    ("orc_gpu_create", VP, [C.POINTER(OrcBatch), C.POINTER(OrcOpts)]),
    ("orc_gpu_destroy", None, [VP]),
    ("orc_gpu_build", C.c_double, [VP]),
    ("orc_gpu_n_tasks", C.c_int64, [VP]),
    ("orc_gpu_tables", None, [VP, VP, VP, VP, VP]),
    ("orc_gpu_reverse", None, [VP, C.c_int, VP, VP, C.c_int, VP, VP]),
    ("orc_truth_create", VP, [C.POINTER(OrcBatch)]),
    ("orc_truth_destroy", None, [VP]),
    ("orc_truth_lengths", None, [VP, VP, VP]),
    ("orc_truth_partial", C.c_double, [VP, C.c_int, C.c_int, C.c_double]),
    ("orc_truth_reverse", None, [VP, C.c_int, VP, VP, C.c_int, VP, VP, C.c_int]),
]


def build():
    r = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        for n, res, args in _SIGS:
            f = getattr(L, n)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data if a is not None else None


class Batch:
    """Keeps the numpy arrays alive behind an OrcBatch."""

    def __init__(self, batch):
        self.src = batch
        if batch.kind == "ellipse":
            self.arrs = [np.ascontiguousarray(batch.a, np.float64), np.ascontiguousarray(batch.b, np.float64)]
            self.c = OrcBatch(0, batch.n_curve, _p(self.arrs[0]), _p(self.arrs[1]), None, None, None, None, None,
                              None, 0)
            self.n_int = batch.n_curve
        else:
            ctrl = np.ascontiguousarray(batch.ctrl, np.float64)
            self.arrs = [np.ascontiguousarray(batch.degree, np.int32), np.ascontiguousarray(batch.dim, np.int32),
                         np.ascontiguousarray(batch.knot_off, np.int64), np.ascontiguousarray(batch.ctrl_off, np.int64),
                         np.ascontiguousarray(batch.knots, np.float64), ctrl]
            self.c = OrcBatch(1, batch.n_curve, None, None, *[_p(a) for a in self.arrs], ctrl.shape[1])
            self.n_int = int(lib().orc_n_intervals(C.byref(self.c)))
        self.n_curve = batch.n_curve


def opts(precision=0, gl_order=3, n_sub=64, max_depth=None, tol=None, normalize=False):
    if max_depth is None:
        max_depth = 14 if precision == 0 else 23
    if tol is None:
        tol = 1e-6 if precision == 0 else 1e-14
    return OrcOpts(precision, gl_order, n_sub, max_depth, tol, 1 if normalize else 0)


def offsets(counts):
    c = np.ascontiguousarray(counts, np.int64)
    off = np.zeros(c.size + 1, np.int64)
    np.cumsum(c, out=off[1:])
    return off


def uniform_positions(counts, closed=True):
    """Positions j/(N-1) (closed) or j/N per curve, concatenated, as the
    reference drivers build them (samples/sample2_bspline/main.cpp:59,
    samples/sample1_ellipse/main.cpp:75) but in double."""
    out = []
    for n in np.asarray(counts, np.int64):
        n = int(n)
        j = np.arange(n, dtype=np.float64)
        if closed:
            out.append(j * (1.0 / (n - 1)) if n > 1 else np.zeros(n))
        else:
            out.append(j * (1.0 / n) if n > 0 else np.zeros(0))
    return np.concatenate(out) if out else np.zeros(0)


class RefCpu:
    def __init__(self, batch, o: OrcOpts):
        self.b = batch if isinstance(batch, Batch) else Batch(batch)
        self.o = o
        self.h = lib().orc_cpu_create(C.byref(self.b.c), C.byref(o))

    def build(self, threads=1) -> float:
        return lib().orc_cpu_build(self.h, threads)

    def lengths(self):
        out = np.empty(self.b.n_curve)
        lib().orc_cpu_lengths(self.h, self.b.n_curve, _p(out))
        return out

    def table(self, c):
        n = lib().orc_cpu_table_size(self.h, c)
        u, s, g = np.empty(n), np.empty(n), np.empty(n, np.int32)
        lib().orc_cpu_table(self.h, c, _p(u), _p(s), _p(g))
        return u, s, g

    def reverse(self, counts, pos, is_relative=True, threads=1):
        off = offsets(counts)
        pos = np.ascontiguousarray(pos, np.float64)
        out = np.empty(off[-1])
        t = lib().orc_cpu_reverse(self.h, self.b.n_curve, _p(off), _p(pos), 1 if is_relative else 0, _p(out), threads)
        return out, t

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_cpu_destroy(self.h)
            self.h = None


class RefGpuPlain:
    def __init__(self, batch, o: OrcOpts):
        self.b = batch if isinstance(batch, Batch) else Batch(batch)
        self.o = o
        self.h = lib().orc_gpu_create(C.byref(self.b.c), C.byref(o))

    def build(self) -> float:
        return lib().orc_gpu_build(self.h)

    def n_tasks(self):
        return lib().orc_gpu_n_tasks(self.h)

    def tables(self):
        n_int = self.b.n_int
        L, sc, sp, rd = (np.empty(self.b.n_curve), np.empty(n_int), np.empty(n_int * self.o.n_sub),
                         np.empty(n_int * self.o.n_sub, np.int32))
        lib().orc_gpu_tables(self.h, _p(L), _p(sc), _p(sp), _p(rd))
        return L, sc, sp, rd

    def reverse(self, counts, pos, is_relative=True):
        off = offsets(counts)
        pos = np.ascontiguousarray(pos, np.float64)
        out = np.empty(off[-1])
        seg = np.empty(off[-1], np.int32)
        lib().orc_gpu_reverse(self.h, self.b.n_curve, _p(off), _p(pos), 1 if is_relative else 0, _p(out), _p(seg))
        return out, seg

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_gpu_destroy(self.h)
            self.h = None


class Truth:
    def __init__(self, batch):
        self.b = batch if isinstance(batch, Batch) else Batch(batch)
        self.h = lib().orc_truth_create(C.byref(self.b.c))

    def lengths(self):
        L, cum = np.empty(self.b.n_curve), np.empty(self.b.n_int)
        lib().orc_truth_lengths(self.h, _p(L), _p(cum))
        return L, cum

    def partial(self, c, segid, u):
        return lib().orc_truth_partial(self.h, c, segid, u)

    def reverse(self, counts, pos, is_relative=True, threads=8):
        off = offsets(counts)
        pos = np.ascontiguousarray(pos, np.float64)
        out = np.empty(off[-1])
        seg = np.empty(off[-1], np.int32)
        lib().orc_truth_reverse(self.h, self.b.n_curve, _p(off), _p(pos), 1 if is_relative else 0, _p(out), _p(seg),
                                threads)
        return out, seg

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_truth_destroy(self.h)
            self.h = None


def gl_table(order):
    x, w = np.zeros(order), np.zeros(order)
    lib().orc_gl_table(order, _p(x), _p(w))
    return x, w


def eps_device(precision, tol, n_sub, level):
    return lib().orc_eps_device(precision, tol, n_sub, level)


def speed(batch: Batch, c, segid, u, precision=1):
    return lib().orc_speed(C.byref(batch.c), c, segid, u, precision)


def deriv(batch: Batch, c, segid, u, truth=False):
    out = np.zeros(3)
    (lib().orc_truth_deriv if truth else lib().orc_deriv)(C.byref(batch.c), c, segid, u, _p(out))
    return out
