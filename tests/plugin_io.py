"""Input/output files of tests/cpp/plugin_engines.hip (the plug-in headers'
GPU and CPU engine subclasses) and the builds of that driver."""
from __future__ import annotations

import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "plugin_engines.hip")
HEADERS = [os.path.join(ROOT, "tests", "cpp", h) for h in ("plugin_ellipse.h", "plugin_bspline.h")] + [
    os.path.join(ROOT, "include", "gcurval", h) for h in ("base.h", "arclength_calculate_cpu.h", "gcurval.h")]
CPU_EXE = os.path.join(ROOT, "tests", "cpp", "plugin_engines_cpu")
GPU_EXE = os.path.join(ROOT, "tests", "cpp", "plugin_engines")

# host-only build with the sanitizers; -ffp-contract=off keeps float rounding
# in written expression order, as the oracle's build does
CPU_CMD = ["g++", "-std=c++17", "-O1", "-g", "-ffp-contract=off", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-Wall", "-Wextra", "-Werror",
           "-Wno-maybe-uninitialized", "-I" + os.path.join(ROOT, "include"), "-x", "c++", SRC, "-o", CPU_EXE,
           "-lpthread"]

SAN_ENV = {"ASAN_OPTIONS": "detect_leaks=1:halt_on_error=1:abort_on_error=0",
           "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"}


def build_cpu(force: bool = False) -> str:
    """Build the sanitized host-only driver (CPU engine) unless it is newer
    than its sources."""
    if not force and os.path.exists(CPU_EXE):
        t = os.path.getmtime(CPU_EXE)
        if all(os.path.getmtime(p) <= t for p in [SRC] + HEADERS):
            return CPU_EXE
    r = subprocess.run(CPU_CMD + [], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ build of plugin_engines failed:\n" + r.stdout + r.stderr)
    return CPU_EXE


def write_ellipse(path, a, b, counts, pos=None):
    with open(path, "wb") as f:
        np.array([len(a)], np.int32).tofile(f)
        np.asarray(a, np.float64).tofile(f)
        np.asarray(b, np.float64).tofile(f)
        np.asarray(counts, np.int32).tofile(f)
        np.array([1 if pos is not None else 0], np.int32).tofile(f)
        if pos is not None:
            np.asarray(pos, np.float64).tofile(f)


def cubic_layout(batch):
    """knots_offset / knot_vector / control points (x, y, z) of a cubic 3-D
    BSplineBatch in the plug-in's layout (control points of curve c start at
    knots_offset[c] - 4 c)."""
    assert np.all(batch.degree == 3) and batch.ctrl_stride == 3
    ko = batch.knot_off.astype(np.int32)
    assert np.array_equal(batch.ctrl_off, batch.knot_off - 4 * np.arange(batch.n_curve + 1))
    return ko, batch.knots, batch.ctrl


def write_bspline(path, batch, counts):
    ko, knots, ctrl = cubic_layout(batch)
    with open(path, "wb") as f:
        np.array([batch.n_curve], np.int32).tofile(f)
        ko.tofile(f)
        np.asarray(knots, np.float64).tofile(f)
        np.ascontiguousarray(ctrl, np.float64).tofile(f)
        np.asarray(counts, np.int32).tofile(f)


def read_out(path, n, total, dim):
    v = np.fromfile(path, np.float64)
    assert v.size == n + total + total * dim, (v.size, n, total, dim)
    return v[:n], v[n:n + total], v[n + total:].reshape(total, dim)


def run(exe, prec, family, engine, inp, out, threads=1, env_extra=None, timeout=600):
    env = dict(os.environ)
    if env_extra:
        env.update(env_extra)
    r = subprocess.run([exe, prec, family, engine, inp, out, str(threads)], capture_output=True, text=True,
                       env=env, timeout=timeout)
    return r
