"""gcurval_amd — MI355X (gfx950) arc-length calculation and arc-length sampling.

Python host mirror of gcurval's engine interface over the C ABI declared in
``include/gcurval_c.h`` (ctypes; no torch). The class ``ArclengthCalculateGpu``
follows ``gcurval::ArclengthCalculateGpu`` (gcurval/arclength_calculate.h:10-18,
gcurval/detail/cuda/arclength_calculate.h:141-237): construct with a curve batch,
``createArclengthTable()``, then ``getAllCurveArclengths()`` /
``getArclength(cid)`` and ``arclengthReverse(...)``.

The HIP library is required: importing works without it, but every call that
would reach the device raises ``RuntimeError`` when ``libgcurval_amd.so`` is
missing. There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

__all__ = [
    "lib", "load_library", "GcvError", "Opts", "Context", "DeviceArray",
    "EllipseBatch", "BSplineBatch", "ArclengthCalculateGpu",
    "gcurvalDevice", "gcurvalHost", "gcurvalReverseReturnUs", "gcurvalReverseReturnPoints",
    "GCV_FP32", "GCV_FP64", "GCV_START_PLAIN", "GCV_START_LEARNED", "GCV_START_QUINTIC",
]

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgcurval_amd.so")

# GcurvalEnum (gcurval/base/util.h:71-77)
gcurvalDevice, gcurvalHost, gcurvalReverseReturnUs, gcurvalReverseReturnPoints = 0, 1, 2, 3
GCV_FP32, GCV_FP64 = 0, 1
# Newton start of the B-spline sampler (gcv_table_set_newton_start)
GCV_START_PLAIN, GCV_START_LEARNED, GCV_START_QUINTIC = 0, 1, 2

_STATUS = {0: "GCV_OK", 1: "GCV_ERR_INVALID", 2: "GCV_ERR_OOM", 3: "GCV_ERR_HIP",
           4: "GCV_ERR_UNSUPPORTED", 5: "GCV_ERR_STATE", 6: "GCV_ERR_DATA"}


class GcvError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_STATUS.get(code, code)}: {msg}")
        self.code = code


class Opts(C.Structure):
    _fields_ = [
        ("precision", C.c_int),
        ("gl_order", C.c_int),
        ("n_subdivided", C.c_int),
        ("max_depth", C.c_int),
        ("tolerance", C.c_double),
        ("normalize", C.c_int),
        ("newton_iters", C.c_int),
        ("skip_fp64_step", C.c_int),
        ("reserved", C.c_int * 5),
    ]


class _BSplineBatchC(C.Structure):
    _fields_ = [
        ("n_curve", C.c_int),
        ("ctrl_stride", C.c_int),
        ("degree", C.c_void_p),
        ("dim", C.c_void_p),
        ("knot_off", C.c_void_p),
        ("ctrl_off", C.c_void_p),
        ("knots", C.c_void_p),
        ("ctrl", C.c_void_p),
    ]


_lib = None

# (name, restype, argtypes)
_SIGS = [
    ("gcv_version", C.c_char_p, []),
    ("gcv_opts_default", None, [C.POINTER(Opts)]),
    ("gcv_ctx_create", C.c_int, [C.c_int, C.c_size_t, C.POINTER(C.c_void_p)]),
    ("gcv_ctx_destroy", C.c_int, [C.c_void_p]),
    ("gcv_last_error", C.c_char_p, [C.c_void_p]),
    ("gcv_alloc", C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    ("gcv_free", C.c_int, [C.c_void_p, C.c_void_p]),
    ("gcv_arena_stats", C.c_int, [C.c_void_p, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    ("gcv_copy_h2d", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    ("gcv_copy_d2h", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    ("gcv_copy_d2d", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    ("gcv_valu_calibration", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_int]),
    ("gcv_sync", C.c_int, [C.c_void_p]),
    ("gcv_device_info", C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.c_char_p, C.c_int]),
    ("gcv_timer_start", C.c_int, [C.c_void_p]),
    ("gcv_timer_stop", C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    ("gcv_profile_enable", C.c_int, [C.c_void_p, C.c_int]),
    ("gcv_profile_reset", C.c_int, [C.c_void_p]),
    ("gcv_profile_query", C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    ("gcv_table_create_ellipse", C.c_int,
     [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.POINTER(Opts), C.POINTER(C.c_void_p)]),
    ("gcv_table_create_bspline", C.c_int,
     [C.c_void_p, C.POINTER(_BSplineBatchC), C.POINTER(Opts), C.POINTER(C.c_void_p)]),
    ("gcv_table_build", C.c_int, [C.c_void_p]),
    ("gcv_table_build_uniform", C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    ("gcv_table_build_seeded", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("gcv_table_root_layout", C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_void_p),
                                        C.POINTER(C.c_void_p)]),
    ("gcv_table_seed_floor", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    ("gcv_table_destroy", C.c_int, [C.c_void_p]),
    ("gcv_table_info", C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
    ("gcv_table_lengths", C.c_int, [C.c_void_p, C.c_void_p]),
    ("gcv_table_device_views", C.c_int,
     [C.c_void_p] + [C.POINTER(C.c_void_p)] * 5),
    ("gcv_table_stats", C.c_int, [C.c_void_p] + [C.POINTER(C.c_uint64)] * 9),
    ("gcv_sample", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    ("gcv_sample_noalloc", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    ("gcv_ctx_stream", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    ("gcv_points", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    ("gcv_graph_begin", C.c_int, [C.c_void_p]),
    ("gcv_graph_end", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    ("gcv_graph_launch", C.c_int, [C.c_void_p]),
    ("gcv_graph_destroy", C.c_int, [C.c_void_p]),
    ("gcv_table_autotune", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    ("gcv_table_set_launch", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    ("gcv_table_set_scan", C.c_int, [C.c_void_p, C.c_int]),
    ("gcv_table_set_newton_start", C.c_int, [C.c_void_p, C.c_int]),
    ("gcv_table_start_labels", C.c_int, [C.c_void_p, C.c_void_p]),
    ("gcv_table_set_iters_out", C.c_int, [C.c_void_p, C.c_void_p]),
    ("gcv_mlp_create", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    ("gcv_mlp_destroy", C.c_int, [C.c_void_p]),
    ("gcv_mlp_set_params", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("gcv_mlp_get_params", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("gcv_mlp_features", C.c_int, [C.c_void_p, C.c_void_p]),
    ("gcv_mlp_predict_k", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    ("gcv_table_build_mlp", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    ("gcv_table_seed_k", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    ("gcv_mlp_forward", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    ("gcv_mlp_grad", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_float, C.c_void_p, C.c_void_p]),
    ("gcv_mlp_train_step", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_float, C.c_float, C.c_void_p]),
    ("gcv_host_bspline_span_poly", C.c_int,
     [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_void_p,
      C.POINTER(C.c_double), C.POINTER(C.c_double)]),
]


def load_library(path: str | None = None):
    """Load libgcurval_amd.so (raises RuntimeError if it is not built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"gcurval_amd: HIP library not built ({p}); run python -m gcurval_amd.build")
    L = C.CDLL(p)
    for name, res, args in _SIGS:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def lib():
    return load_library()


def exported_symbols():
    return [s[0] for s in _SIGS]


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class DeviceArray:
    """A device allocation from the context arena with a numpy dtype."""

    def __init__(self, ctx: "Context", count: int, dtype):
        self.ctx = ctx
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = self.count * self.dtype.itemsize
        p = C.c_void_p()
        ctx._check(lib().gcv_alloc(ctx.h, max(self.nbytes, 1), C.byref(p)))
        self.ptr = p.value

    def upload(self, host: np.ndarray) -> "DeviceArray":
        h = np.ascontiguousarray(host, dtype=self.dtype)
        assert h.size == self.count, (h.size, self.count)
        self.ctx._check(lib().gcv_copy_h2d(self.ctx.h, self.ptr, _ptr(h), self.nbytes))
        return self

    def download(self) -> np.ndarray:
        out = np.empty(self.count, dtype=self.dtype)
        self.ctx._check(lib().gcv_copy_d2h(self.ctx.h, _ptr(out), self.ptr, self.nbytes))
        return out

    def free(self):
        if self.ptr is not None and self.ctx.h is not None:
            lib().gcv_free(self.ctx.h, self.ptr)
        self.ptr = None


ARENA_POOL = (1 << 64) - 1  # GCV_ARENA_POOL: an arena that grows by hipMalloc chunks


class Context:
    """One HIP device, one stream, one arena: by default a pool that grows by
    hipMalloc chunks (ARENA_POOL); arena_bytes=0: hipMalloc per allocation;
    any other value: one fixed chunk of that size."""

    def __init__(self, device: int = 0, arena_bytes: int = ARENA_POOL):
        L = lib()
        h = C.c_void_p()
        rc = L.gcv_ctx_create(device, arena_bytes, C.byref(h))
        if rc != 0:
            raise GcvError(rc, f"gcv_ctx_create(device={device}) failed")
        self.h = h.value

    def _check(self, rc: int):
        if rc != 0:
            raise GcvError(rc, lib().gcv_last_error(self.h).decode())

    def array(self, host: np.ndarray, dtype=None) -> DeviceArray:
        a = np.ascontiguousarray(host, dtype=dtype or np.asarray(host).dtype)
        return DeviceArray(self, a.size, a.dtype).upload(a)

    def empty(self, count: int, dtype) -> DeviceArray:
        return DeviceArray(self, count, dtype)

    def sync(self):
        self._check(lib().gcv_sync(self.h))

    def copy_d2d(self, dst: DeviceArray, src: DeviceArray, nbytes: int):
        self._check(lib().gcv_copy_d2d(self.h, dst.ptr, src.ptr, nbytes))

    def valu_calibration(self, kind: int, buf: DeviceArray, n_blocks: int, iters: int):
        """Enqueue the VALU rate kernel of `kind` (0 FP32 FMA, 1 FP64 FMA,
        2 FP32 sqrt) over n_blocks x 256 lanes; buf holds n_blocks * 256
        doubles."""
        if buf.nbytes < n_blocks * 256 * 8:
            raise ValueError("valu_calibration: buffer smaller than n_blocks * 256 doubles")
        self._check(lib().gcv_valu_calibration(self.h, kind, buf.ptr, n_blocks, iters))

    def device_info(self):
        n = C.c_int()
        buf = C.create_string_buffer(256)
        self._check(lib().gcv_device_info(self.h, C.byref(n), buf, 256))
        return n.value, buf.value.decode()

    def timer_start(self):
        self._check(lib().gcv_timer_start(self.h))

    def timer_stop(self) -> float:
        ms = C.c_double()
        self._check(lib().gcv_timer_stop(self.h, C.byref(ms)))
        return ms.value

    def profile(self, on: bool):
        self._check(lib().gcv_profile_enable(self.h, 1 if on else 0))

    def profile_reset(self):
        self._check(lib().gcv_profile_reset(self.h))

    def profile_query(self, prefix: str):
        t = C.c_double()
        n = C.c_int64()
        self._check(lib().gcv_profile_query(self.h, prefix.encode(), C.byref(t), C.byref(n)))
        return t.value, n.value

    def graph_begin(self):
        """Start capturing this context's stream into a graph."""
        self._check(lib().gcv_graph_begin(self.h))

    def graph_end(self) -> "Graph":
        g = C.c_void_p()
        self._check(lib().gcv_graph_end(self.h, C.byref(g)))
        return Graph(self, g.value)

    def arena_stats(self):
        u, c = C.c_size_t(), C.c_size_t()
        self._check(lib().gcv_arena_stats(self.h, C.byref(u), C.byref(c)))
        return u.value, c.value

    def close(self):
        if self.h is not None:
            lib().gcv_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Graph:
    """A captured sequence of library calls, replayed with one launch."""

    def __init__(self, ctx: "Context", h: int):
        self.ctx = ctx
        self.h = h

    def launch(self):
        self.ctx._check(lib().gcv_graph_launch(self.h))

    def close(self):
        if self.h:
            lib().gcv_graph_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class _Tune(C.Structure):
    _fields_ = [("sub_waves_per_block", C.c_int), ("sub_grid_pct", C.c_int), ("best_build_ms", C.c_double),
                ("default_build_ms", C.c_double), ("n_tried", C.c_int), ("bin_streams", C.c_int)]


def make_opts(precision=GCV_FP32, gl_order=3, n_subdivided=64, max_depth=0, tolerance=0.0,
              normalize=False, newton_iters=0, fp64_step=True) -> Opts:
    o = Opts()
    lib().gcv_opts_default(C.byref(o))
    o.precision = precision
    o.gl_order = gl_order
    o.n_subdivided = n_subdivided
    o.max_depth = max_depth
    o.tolerance = tolerance
    o.normalize = 1 if normalize else 0
    o.newton_iters = newton_iters
    o.skip_fp64_step = 0 if fp64_step else 1
    return o


class EllipseBatch:
    """Ellipses x = a cos t, y = b sin t (samples/sample1_ellipse/ellipse.h:17-34)."""

    kind = "ellipse"

    def __init__(self, a, b):
        self.a = np.ascontiguousarray(a, dtype=np.float64)
        self.b = np.ascontiguousarray(b, dtype=np.float64)
        assert self.a.shape == self.b.shape and self.a.ndim == 1
        self.n_curve = self.a.size

    @property
    def n_intervals(self):
        return self.n_curve


class BSplineBatch:
    """Clamped B-splines in CSR layout (samples/sample2_bspline/structs.h:6-25).

    knots[knot_off[c]:knot_off[c+1]] and ctrl[ctrl_off[c]:ctrl_off[c+1], :dim[c]]
    for curve c; ctrl has ctrl_stride (2 or 3) columns.
    """

    kind = "bspline"

    def __init__(self, degree, dim, knot_off, ctrl_off, knots, ctrl):
        self.degree = np.ascontiguousarray(degree, dtype=np.int32)
        self.dim = np.ascontiguousarray(dim, dtype=np.int32)
        self.knot_off = np.ascontiguousarray(knot_off, dtype=np.int64)
        self.ctrl_off = np.ascontiguousarray(ctrl_off, dtype=np.int64)
        self.knots = np.ascontiguousarray(knots, dtype=np.float64)
        ctrl = np.ascontiguousarray(ctrl, dtype=np.float64)
        assert ctrl.ndim == 2 and ctrl.shape[1] in (2, 3)
        self.ctrl = ctrl
        self.ctrl_stride = ctrl.shape[1]
        self.n_curve = self.degree.size

    @property
    def n_intervals(self):
        nk = np.diff(self.knot_off)
        return int(np.sum(nk - 2 * self.degree - 1))

    def intervals_per_curve(self):
        return np.diff(self.knot_off) - 2 * self.degree.astype(np.int64) - 1


def sample_offsets(n_ptrs_for_curves, n_curve: int, sample_position=None):
    """Sample offsets (the partial sum of the per-curve counts, n_curve + 1
    entries, as gcv_sample takes them) and the positions as a flat float64
    array (or None), after checking what the device cannot: one non-negative
    count per curve, and one position per sample."""
    n = np.ascontiguousarray(n_ptrs_for_curves, dtype=np.int64)
    if n.ndim != 1 or n.size != n_curve:
        raise ValueError(f"n_ptrs_for_curves: {n_curve} per-curve counts expected, got shape {n.shape}")
    if np.any(n < 0):
        raise ValueError("n_ptrs_for_curves: negative sample count")
    off = np.zeros(n_curve + 1, dtype=np.int64)
    np.cumsum(n, out=off[1:])
    pos = None
    if sample_position is not None:
        pos = np.ascontiguousarray(sample_position, dtype=np.float64).reshape(-1)
        if pos.size != int(off[-1]):
            raise ValueError(f"sample_position: {int(off[-1])} positions expected (the sum of the counts), "
                             f"got {pos.size}")
    return off, pos


class ArclengthCalculateGpu:
    """Arc-length table + sampler over one curve batch on one GPU.

    Mirrors gcurval::ArclengthCalculateGpu (gcurval/detail/cuda/arclength_calculate.h:141-237):
    the constructor uploads the batch, createArclengthTable() builds, then the
    getters and arclengthReverse() use the table.
    """

    def __init__(self, ctx: Context, batch, precision=GCV_FP32, gl_order=3, n_subdivided=64, max_depth=0,
                 tolerance=0.0, normalize=False, newton_iters=0, fp64_step=True):
        self.ctx = ctx
        self.batch = batch
        self.opts = make_opts(precision, gl_order, n_subdivided, max_depth, tolerance, normalize, newton_iters,
                              fp64_step)
        self._bufs = []
        self.h = None
        h = C.c_void_p()
        try:
            if batch.kind == "ellipse":
                da, db = ctx.array(batch.a), ctx.array(batch.b)
                self._bufs += [da, db]
                ctx._check(lib().gcv_table_create_ellipse(ctx.h, batch.n_curve, da.ptr, db.ptr,
                                                          C.byref(self.opts), C.byref(h)))
            else:
                bufs = [ctx.array(batch.degree), ctx.array(batch.dim), ctx.array(batch.knot_off),
                        ctx.array(batch.ctrl_off), ctx.array(batch.knots), ctx.array(batch.ctrl.reshape(-1))]
                self._bufs += bufs
                bc = _BSplineBatchC(batch.n_curve, batch.ctrl_stride, *[b.ptr for b in bufs])
                ctx._check(lib().gcv_table_create_bspline(ctx.h, C.byref(bc), C.byref(self.opts), C.byref(h)))
        except Exception:
            # a rejected batch (bad options or data) leaves no device buffers behind
            for buf in self._bufs:
                buf.free()
            self._bufs = []
            raise
        self.h = h.value
        nc, ni, ns = C.c_int(), C.c_int64(), C.c_int()
        ctx._check(lib().gcv_table_info(self.h, C.byref(nc), C.byref(ni), C.byref(ns)))
        self.n_curve, self.n_intervals, self.n_subdivided = nc.value, ni.value, ns.value
        self._lengths = None

    # -- table build (createArclengthTable) --
    def createArclengthTable(self):
        self.ctx._check(lib().gcv_table_build(self.h))
        self._lengths = None

    def createArclengthTableUniform(self, log2_roots: int, d_cost: "DeviceArray | None" = None):
        """Build with 2^log2_roots roots per knot interval; optional per-interval quadrature counts."""
        self.ctx._check(lib().gcv_table_build_uniform(self.h, log2_roots, d_cost.ptr if d_cost is not None else None))
        self._lengths = None

    def createArclengthTableSeeded(self, d_k: "DeviceArray", d_cost: "DeviceArray | None" = None):
        """Build with 2^k_i roots for knot interval i (d_k: device int8 per interval)."""
        self.ctx._check(lib().gcv_table_build_seeded(self.h, d_k.ptr, d_cost.ptr if d_cost is not None else None))
        self._lengths = None

    def seed_floor(self) -> np.ndarray:
        """Per-interval minimum k of the last seeded build."""
        p = C.c_void_p()
        self.ctx._check(lib().gcv_table_seed_floor(self.h, C.byref(p)))
        out = np.empty(self.n_intervals, np.int8)
        if out.nbytes:
            self.ctx._check(lib().gcv_copy_d2h(self.ctx.h, _ptr(out), p.value, out.nbytes))
        return out

    def root_layout(self):
        """(seeded, log2 roots of a uniform build, device int8 log2 root count
        per interval of a seeded build, device int64 first root of each
        interval of a seeded build: roots dense in interval order)."""
        sd, ls, rk, ro = C.c_int(), C.c_int(), C.c_void_p(), C.c_void_p()
        self.ctx._check(lib().gcv_table_root_layout(self.h, C.byref(sd), C.byref(ls), C.byref(rk), C.byref(ro)))
        return bool(sd.value), ls.value, rk.value, ro.value

    def createArclengthTableLearned(self, mlp, hedge: bool = False, d_cost: "DeviceArray | None" = None):
        """Seeded build whose per-interval root counts come from the learned predictor."""
        self.ctx._check(lib().gcv_table_build_mlp(self.h, mlp.h, 1 if hedge else 0,
                                                  d_cost.ptr if d_cost is not None else None))
        self._lengths = None

    def seed_k(self) -> np.ndarray:
        """k used by the last learned-seed build (before the safety floor)."""
        p = C.c_void_p()
        self.ctx._check(lib().gcv_table_seed_k(self.h, C.byref(p)))
        out = np.empty(self.n_intervals, np.int8)
        if out.nbytes:
            self.ctx._check(lib().gcv_copy_d2h(self.ctx.h, _ptr(out), p.value, out.nbytes))
        return out

    def autotune(self, mlp=None, reps: int = 3) -> dict:
        """Pick the persistent subdivision kernel's launch shape for this table
        by timing builds (learned-seed builds when `mlp` is given)."""
        t = _Tune()
        self.ctx._check(lib().gcv_table_autotune(self.h, mlp.h if mlp is not None else None, reps, C.byref(t)))
        self._lengths = None
        return {k: getattr(t, k) for k, _ in _Tune._fields_}

    def set_launch(self, waves_per_block: int = 0, grid_pct: int = 100, bin_streams: int = 0):
        self.ctx._check(lib().gcv_table_set_launch(self.h, waves_per_block, grid_pct, bin_streams))

    def set_scan(self, mode: int = 0):
        """Cumulative-scan kernel: 0 by curve length, 1 a wave per curve, 2 the look-back scan."""
        self.ctx._check(lib().gcv_table_set_scan(self.h, mode))

    def set_newton_start(self, mode: int = GCV_START_LEARNED):
        """Newton start of B-spline samples: GCV_START_PLAIN (cubic Hermite, FP32
        steps, FP64 step; the table's default), GCV_START_LEARNED (the
        predictor's per-interval choice of the quintic start on learned
        builds) or GCV_START_QUINTIC."""
        self.ctx._check(lib().gcv_table_set_newton_start(self.h, mode))

    def start_labels(self) -> np.ndarray:
        """Per interval of the built table: log2 of the worst quintic-start error
        over its roots (relative to root width)."""
        n = self.n_intervals
        d = self.ctx.empty(max(n, 1), np.float32)
        try:
            self.ctx._check(lib().gcv_table_start_labels(self.h, d.ptr))
            return d.download()[:n]
        finally:
            d.free()

    def set_iters_out(self, d_iters: "DeviceArray | None"):
        """Per-sample Newton step counts of the following samplings (FP32 steps
        in the low nibble, FP64 steps in the high one); None turns it off."""
        self.ctx._check(lib().gcv_table_set_iters_out(self.h, d_iters.ptr if d_iters is not None else None))

    def getAllCurveArclengths(self) -> np.ndarray:
        out = np.empty(self.n_curve, dtype=np.float64)
        self.ctx._check(lib().gcv_table_lengths(self.h, _ptr(out)))
        self._lengths = out
        return out.copy()

    def getArclength(self, cid: int) -> float:
        if self._lengths is None:
            self.getAllCurveArclengths()
        return float(self._lengths[cid])

    def clearArclength(self):
        self._lengths = None

    def stats(self):
        """Build statistics of the last build: quadratures, refined roots,
        deepest task stack, overflow flag (must be 0), refinement tasks,
        leaves that still failed the error test at the depth cap, tiles
        re-run with the full task stack (the counts include both passes),
        64-root wave tiles, and refinement passes over popped tasks."""
        v = [C.c_uint64() for _ in range(9)]
        self.ctx._check(lib().gcv_table_stats(self.h, *[C.byref(x) for x in v]))
        keys = ("quadratures", "deep_roots", "max_stack", "overflow", "tasks", "depth_capped", "reruns", "tiles",
                "passes")
        return {k: x.value for k, x in zip(keys, v)}

    def device_views(self):
        v = [C.c_void_p() for _ in range(5)]
        self.ctx._check(lib().gcv_table_device_views(self.h, *[C.byref(x) for x in v]))
        return [x.value for x in v]

    def tables(self):
        """Host copies: (int_cum[n_int], root prefixes, deep[n_int], curve_len, int_off). The root
        prefixes are interval after interval, 2^log2_roots per interval (uniform builds) or 2^k_i
        for interval i (seeded builds)."""
        ic, rp, dp, cl, io = self.device_views()
        L = lib()
        n_int = self.n_intervals
        n_roots = int(np.sum(1 << self.root_counts().astype(np.int64)))
        out = [np.empty(n_int, np.float64), np.empty(n_roots, np.float64), np.empty(n_int, np.uint64),
               np.empty(self.n_curve, np.float64), np.empty(self.n_curve + 1, np.int64)]
        for arr, p in zip(out, (ic, rp, dp, cl, io)):
            if arr.nbytes:
                self.ctx._check(L.gcv_copy_d2h(self.ctx.h, _ptr(arr), p, arr.nbytes))
        return tuple(out)

    def root_offsets(self) -> np.ndarray:
        """First root of each interval in the root-prefix view."""
        seeded, ls, _, ro = self.root_layout()
        if not seeded:
            return np.arange(self.n_intervals, dtype=np.int64) << ls
        out = np.empty(self.n_intervals, np.int64)
        if out.nbytes:
            self.ctx._check(lib().gcv_copy_d2h(self.ctx.h, _ptr(out), ro, out.nbytes))
        return out

    def root_counts(self) -> np.ndarray:
        """log2 root count per interval of the last build."""
        seeded, ls, rk, _ = self.root_layout()
        k = np.full(self.n_intervals, ls, np.int8)
        if seeded and k.size:
            self.ctx._check(lib().gcv_copy_d2h(self.ctx.h, _ptr(k), rk, k.nbytes))
        return k

    # -- sampling (arclengthReverse) --
    def sample_device(self, d_off: DeviceArray, n_total: int, d_pos: DeviceArray | None, is_relative: bool,
                      d_out: DeviceArray, d_seg: DeviceArray | None = None, uniform_closed: bool = True):
        """Enqueue sampling on device buffers (no host sync)."""
        self.ctx._check(lib().gcv_sample(self.h, d_off.ptr, n_total, d_pos.ptr if d_pos is not None else None,
                                         1 if uniform_closed else 0, 1 if is_relative else 0, d_out.ptr,
                                         d_seg.ptr if d_seg is not None else None))

    def arclengthReverse(self, n_ptrs_for_curves, sample_position=None, is_relative=True,
                         where=gcurvalHost, what=gcurvalReverseReturnUs, uniform_closed=True,
                         return_segments=False):
        """Sample parameters for per-curve sample counts.

        sample_position: host array of positions concatenated per curve, or
        None for uniform positions j/(N_c-1) (uniform_closed) or j/N_c.
        where=gcurvalHost returns numpy arrays; gcurvalDevice returns the
        DeviceArray (owned by the caller, free with .free()).
        """
        points = what == gcurvalReverseReturnPoints
        off, pos = sample_offsets(n_ptrs_for_curves, self.n_curve, sample_position)
        total = int(off[-1])
        d_off = self.ctx.array(off)
        d_pos = self.ctx.array(pos) if pos is not None else None
        d_out = self.ctx.empty(max(total, 1), np.float64)
        d_seg = self.ctx.empty(max(total, 1), np.int32) if (return_segments or points) else None
        try:
            self.sample_device(d_off, total, d_pos, is_relative, d_out, d_seg, uniform_closed)
            if points:
                dim = 2 if self.batch.kind == "ellipse" else int(self.batch.ctrl_stride)
                d_pts = self.ctx.empty(max(total, 1) * dim, np.float64)
                self.ctx._check(lib().gcv_points(self.h, d_off.ptr, total, d_out.ptr, d_seg.ptr, dim, d_pts.ptr))
                d_out.free()
                d_seg.free()
                if where == gcurvalDevice:
                    return d_pts
                pts = d_pts.download()[:total * dim].reshape(total, dim)
                d_pts.free()
                return pts
            if where == gcurvalDevice:
                return (d_out, d_seg) if return_segments else d_out
            u = d_out.download()[:total]
            seg = d_seg.download()[:total] if d_seg is not None else None
            d_out.free()
            if d_seg is not None:
                d_seg.free()
            return (u, seg) if return_segments else u
        finally:
            d_off.free()
            if d_pos is not None:
                d_pos.free()

    def close(self):
        if getattr(self, "h", None) is not None:
            self.ctx._check(lib().gcv_table_destroy(self.h))  # refused during graph capture
            self.h = None
        for b in self._bufs:
            b.free()
        self._bufs = []

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def host_bspline_span_poly(p, D, s, knots, ctrl, scale_inv=1.0):
    """Power-basis coefficients of dC/dx on span s, computed by the library's
    host build of the setup routine (no GPU needed)."""
    knots = np.ascontiguousarray(knots, np.float64)
    ctrl = np.ascontiguousarray(ctrl, np.float64)
    out = np.zeros(p * D, np.float64)
    u0, h = C.c_double(), C.c_double()
    rc = lib().gcv_host_bspline_span_poly(p, D, s, _ptr(knots), _ptr(ctrl), ctrl.shape[1], scale_inv, _ptr(out),
                                          C.byref(u0), C.byref(h))
    if rc != 0:
        raise GcvError(rc, "bspline_span_poly")
    return out.reshape(D, p), u0.value, h.value
