"""Build recipe for the gfx950 engine library (libgcurval_amd.so).

Compiles every ``csrc/*.hip`` unit with hipcc for gfx950 in parallel and links
them into ``gcurval_amd/libgcurval_amd.so`` (in-tree, so the built library
travels with the source tree). Run ``python -m gcurval_amd.build``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgcurval_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed FP32 VALU forms (v_pk_*_f32) are disabled for all device code: some
# of their forms have produced silently wrong results on MI355X hosts, so the
# FP32 kernels run on the non-packed VALU.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # matrix-core products accumulate in ordinary VGPRs: AGPRs are allocated
    # on top of a kernel's VGPR peak, so in kernels that mix interval work
    # and the learned seed they would cost occupancy
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-I" + os.path.join(ROOT, "include"),
    "-Wno-unused-result",
]


def _sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, src[:-4] + ".o")
    cmd = [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True, jobs: int | None = None) -> str:
    """Compile all units and link the shared library; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    t0 = time.time()
    srcs = _sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"[gcurval_amd] built {LIB} from {len(srcs)} units in {time.time() - t0:.1f}s", flush=True)
    return LIB


TEST_DRIVERS = [("tests/cpp/template_surface.hip", "tests/cpp/template_surface"),
                ("tests/cpp/plugin_engines.hip", "tests/cpp/plugin_engines")]


def build_tests(verbose: bool = True) -> list:
    """Build the C++ drivers of the header-only template surface used by the
    GPU tests (linked against the library, found through an rpath)."""
    out = []
    for src, exe in TEST_DRIVERS:
        s, e = os.path.join(ROOT, src), os.path.join(ROOT, exe)
        rpath = os.path.relpath(HERE, os.path.dirname(e))
        cmd = [HIPCC, *HIP_FLAGS, s, "-o", e + ".tmp", "-L" + HERE, "-lgcurval_amd",
               "-Wl,-rpath,$ORIGIN/" + rpath]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
        os.replace(e + ".tmp", e)
        out.append(e)
        if verbose:
            print(f"[gcurval_amd] built {e}", flush=True)
    return out


def build_oracle(verbose: bool = True) -> str:
    """Build the CPU oracle (test infrastructure) with its Makefile."""
    odir = os.path.join(ROOT, "oracle")
    r = subprocess.run(["make", "-s", "-B", "-C", odir], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    path = os.path.join(odir, "liboracle.so")
    if verbose:
        print(f"[gcurval_amd] built {path}", flush=True)
    return path


if __name__ == "__main__":
    build()
    if "--no-tests" not in sys.argv:
        build_tests()
    if "--no-oracle" not in sys.argv:
        build_oracle()
