#!/usr/bin/env python3
"""Benchmark: adaptive arc-length table build + arc-length sampling on one MI355X.

A step is one pass of the hot path over the whole synthetic batch: the table
build (setup, adaptive Gauss-Legendre subdivision, cumulative scan) followed by
arc-length-uniform sampling of every curve, with all inputs resident in HBM.
`value` is curve-intervals per second of step time with the build path the
workload is defined on: the plain 64-way build for the 1M-ellipse workload,
the learned-seed build (the MLP predicts each knot interval's initial root
count; the adaptive error test still verifies every root) for the B-spline
workloads. Samples per second of the same steps are reported beside it, and
the other build path is timed on the same inputs ("seed_comparison").

Default workload: 1,000,000 ellipses, a ~ U[0.5, 10], b = a U[0.01, 1], one
knot interval each, FP32 quadrature with FP64 scan, 64..4096 samples per curve
(log-uniform), uniform positions j/(N-1). The same line carries the learned vs
plain comparison on the three B-spline workloads ("c2_seed_comparison": 20,000
cubic 3-D curves, ~7.3M intervals; "c3_seed_comparison": the same distribution
at ~4M intervals; "c4_seed_comparison": 50,000 mixed-degree 2-D/3-D curves,
~12M intervals), the latency of the single-ellipse case ("c0_latency"), and the
reference CPU algorithm on one thread and on all host cores ("cpu_baseline").
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3   # MI355X published FP32 (packed / matrix) peak
PEAK_HBM_GBS = 8000.0      # MI355X published HBM3E bandwidth

# Algorithmic FLOP per speed evaluation (add/sub/mul = 1, FMA = 2,
# div/sqrt/sin/cos = 1): ellipse 12; B-spline 2D(p-1) + (2D-1) + 1 + 2 + 2.
def f_node_ellipse():
    return 12


def f_node_bspline(p, D):
    return 2 * D * (p - 1) + (2 * D - 1) + 1 + 2 + 2


# kernels reported per step; a name is a prefix of the launch names it sums
KERNELS = ("k_subdivide", "k_curve_scan", "k_seg_scan", "k_bspline_curve_setup", "k_bspline_coef", "k_sample", "k_fill_stats",
           "k_seed_floor", "k_bucket", "k_mlp_predict", "k_block_curve")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def workload(name):
    from gcurval_amd import workloads as W
    if name == "c1":
        n = int(os.environ.get("GCV_BENCH_CURVES", 1_000_000))
        batch = W.ellipse_batch(n, seed=413)
        counts = W.ellipse_sample_counts(n, seed=414)
        cfg = {"workload": "c1_ellipses", "n_curve": n, "a": "U[0.5,10]", "b": "a*U[0.01,1]",
               "samples_per_curve": "floor(exp(U[ln 64, ln 4096]))", "quadrature": "fp32 GL-3, fp64 scan",
               "n_subdivided": 64, "tolerance": 1e-6, "normalize": True}
        return batch, counts, cfg
    if name == "c2":
        n = int(os.environ.get("GCV_BENCH_CURVES", 20_000))
        batch = W.bspline_c2(n)
        counts = 32 * batch.intervals_per_curve()
        cfg = {"workload": "c2_cubic3d_bsplines", "n_curve": n, "n_ctrl": "round(exp(U[ln 8, ln 2048]))",
               "samples_per_interval": 32, "quadrature": "fp32 GL-3, fp64 scan", "n_subdivided": 64,
               "tolerance": 1e-6, "normalize": True}
        return batch, counts, cfg
    if name == "c3":
        # the cubic 3-D distribution above at ~4M knot intervals (different curves: seed 419)
        n = int(os.environ.get("GCV_BENCH_CURVES", 11_000))
        batch = W.bspline_batch(n, 419, degrees=(3,), dims=(3,), n_ctrl_range=(8, 2048), cusp_frac=0.1)
        counts = 32 * batch.intervals_per_curve()
        cfg = {"workload": "c3_cubic3d_bsplines_4M", "n_curve": n, "n_ctrl": "round(exp(U[ln 8, ln 2048]))",
               "samples_per_interval": 32, "quadrature": "fp32 GL-3, fp64 scan", "n_subdivided": 64,
               "tolerance": 1e-6, "normalize": True}
        return batch, counts, cfg
    if name == "c4":
        n = int(os.environ.get("GCV_BENCH_CURVES", 50_000))
        batch = W.bspline_c4(n)
        counts = np.full(n, 4000, np.int64)
        cfg = {"workload": "c4_mixed_degree_bsplines", "n_curve": n, "degrees": [2, 3, 5], "dims": [2, 3],
               "samples_per_curve": 4000, "quadrature": "fp32 GL-3, fp64 scan", "n_subdivided": 64,
               "tolerance": 1e-6, "normalize": True}
        return batch, counts, cfg
    raise SystemExit(f"unknown config {name}")


def cpu_baseline(batch, counts, threads, budget_s=20.0):
    """Reference CPU algorithm (oracle restatement of ArclengthCalculateCpu)
    on a bounded prefix of the same workload; rates scale to the metric units."""
    from oracle import oracle as O
    from gcurval_amd import BSplineBatch, EllipseBatch

    def subset(k):
        if batch.kind == "ellipse":
            return EllipseBatch(batch.a[:k], batch.b[:k]), counts[:k]
        ko, co = batch.knot_off[:k + 1], batch.ctrl_off[:k + 1]
        return BSplineBatch(batch.degree[:k], batch.dim[:k], ko, co, batch.knots[:ko[-1]],
                            batch.ctrl[:co[-1]]), counts[:k]

    # normalised inputs, as the reference drivers feed it
    k = 16
    while True:
        sb, sc = subset(min(k, batch.n_curve))
        ob = O.Batch(sb)
        eng = O.RefCpu(ob, O.opts(0, 3, normalize=True))
        t_build = eng.build(threads)
        pos = O.uniform_positions(sc, closed=True)
        _, t_rev = eng.reverse(sc, pos, True, threads)
        t = t_build + t_rev
        if t >= budget_s / 2.5 or k >= batch.n_curve:
            break
        k = min(batch.n_curve, int(k * max(2.0, min(8.0, (budget_s / 2) / max(t, 1e-3)))))
    n_int = ob.n_int
    n_smp = int(np.sum(sc))
    return {
        "n_curve": sb.n_curve, "n_intervals": n_int, "n_samples": n_smp,
        "build_s": t_build, "sample_s": t_rev,
        "intervals_per_s": n_int / t, "samples_per_s": n_smp / t,
    }


def cpu_info():
    """(logical CPUs of the host, CPU model name)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return os.cpu_count(), model


def c0_latency(G, ctx, reps=20, warmup=3):
    """Single ellipse a = 3, b = 1 as one knot interval, FP64, GL-5, absolute
    tolerance 1e-8, then 1,000 arc-length-uniform samples at j / N: the
    latency of one build + sample call pair (median of `reps` event-timed
    runs, and of as many replays of the pair captured in a graph)."""
    from gcurval_amd import EllipseBatch
    b = EllipseBatch(np.array([3.0]), np.array([1.0]))
    eng = G.ArclengthCalculateGpu(ctx, b, precision=G.GCV_FP64, gl_order=5, n_subdivided=64, tolerance=1e-8,
                                  normalize=False)
    n = 1000
    d_off = ctx.array(np.array([0, n], np.int64))
    d_out = ctx.empty(n, np.float64)

    def step():
        eng.createArclengthTable()
        eng.sample_device(d_off, n, None, True, d_out, None, uniform_closed=False)

    for _ in range(warmup):
        step()
    ctx.sync()
    times = []
    for _ in range(reps):
        ctx.timer_start()
        step()
        times.append(ctx.timer_stop())
    ctx.graph_begin()
    step()
    graph = ctx.graph_end()
    graph.launch()
    ctx.sync()
    gtimes = []
    for _ in range(reps):
        ctx.timer_start()
        graph.launch()
        gtimes.append(ctx.timer_stop())
    graph.close()
    L = float(eng.getAllCurveArclengths()[0])
    u = d_out.download()
    rec = {"workload": "c0_single_ellipse", "a": 3.0, "b": 1.0, "precision": "fp64", "gl_order": 5,
           "tolerance": 1e-8, "n_samples": n, "positions": "j/N",
           "median_ms": float(np.median(times)), "min_ms": float(np.min(times)),
           "graph_median_ms": float(np.median(gtimes)), "length": L,
           "samples_monotone": bool(np.all(np.diff(u) > 0))}
    try:
        from scipy.special import ellipe
        exact = 4 * 3.0 * float(ellipe(1.0 - 1.0 / 9.0))
        rec["length_rel_err"] = abs(L - exact) / exact
    except ImportError:
        pass
    eng.close()
    d_off.free()
    d_out.free()
    return rec


def copy_calibration(G, mib, steps):
    """HBM ceiling of a plain streaming copy (read + write) and the byte
    count rocprof's FETCH_SIZE/WRITE_SIZE must report for it."""
    nbytes = mib << 20
    with G.Context(0) as ctx:
        src = ctx.empty(nbytes // 8, np.float64)
        dst = ctx.empty(nbytes // 8, np.float64)
        for _ in range(2):
            ctx.copy_d2d(dst, src, nbytes)
        ctx.profile(True)
        ctx.profile_reset()
        ctx.sync()
        ctx.timer_start()
        for _ in range(steps):
            ctx.copy_d2d(dst, src, nbytes)
        ms = ctx.timer_stop()
        t, n = ctx.profile_query("k_copy16")
        print(json.dumps({"copy_bytes_per_launch": nbytes, "read_kib": nbytes / 1024, "write_kib": nbytes / 1024,
                          "avg_launch_ms": t / n, "gbs_read_plus_write": 2 * nbytes / (t / n * 1e-3) / 1e9,
                          "timed_ms": ms, "steps": steps}), flush=True)


def valu_calibration(G, steps):
    """Sustained VALU rates of the library's chain kernels (FP32 FMA, FP64
    FMA, FP32 sqrt): 8 independent chains per lane over a grid of 16384
    blocks x 256 lanes (8 waves on every SIMD of 256 CUs, 8 rounds), timed
    by HIP events on the library stream."""
    n_blocks, iters = 16384, 2048
    with G.Context(0) as ctx:
        n_cu = ctx.device_info()[0]
        buf = ctx.empty(n_blocks * 256, np.float64)
        out = {}
        for kind, name in enumerate(("k_valu_fma32", "k_valu_fma64", "k_valu_sqrt32")):
            for _ in range(2):
                ctx.valu_calibration(kind, buf, n_blocks, iters)
            ctx.sync()
            ctx.profile(True)
            ctx.profile_reset()
            for _ in range(steps):
                ctx.valu_calibration(kind, buf, n_blocks, iters)
            ctx.sync()
            t, n = ctx.profile_query(name)
            ctx.profile(False)
            ms = t / n
            wave_ops = n_blocks * 4 * iters * 8           # wave64 instructions of the chains
            per_simd = wave_ops / (4 * n_cu)
            rec = {"avg_launch_ms": ms, "wave64_ops": wave_ops,
                   "ns_per_op_per_simd": ms * 1e6 / per_simd,
                   "cycles_per_op_at_2p4ghz": ms * 1e-3 * 2.4e9 / per_simd}
            if kind < 2:
                rec["tflops"] = wave_ops * 64 * 2 / (ms * 1e-3) / 1e12
            out[name] = rec
        buf.free()
        print(json.dumps({"valu_calibration": out, "n_cu": n_cu, "steps": steps}), flush=True)


METRIC = "curve-intervals/sec and samples/sec at \u22641e-6 rel error: MLP-seeded vs plain, 1 GPU"


def measure(G, ctx, batch, counts, steps, warmup, paths, weights=None):
    """Time `steps` hot-path steps (table build + uniform sampling of every
    curve) of each build path after `warmup` untimed steps; per-kernel HIP
    events on the library stream."""
    eng = G.ArclengthCalculateGpu(ctx, batch, precision=G.GCV_FP32, gl_order=3, n_subdivided=64, normalize=True)
    n_total = int(np.sum(counts))
    off = np.zeros(batch.n_curve + 1, np.int64)
    np.cumsum(counts, out=off[1:])
    d_off = ctx.array(off)
    d_out = ctx.empty(n_total, np.float64)
    mlp = None
    if "learned" in paths:
        from gcurval_amd.seed import SeedMLP
        mlp = SeedMLP.pretrained(ctx, weights) if weights else SeedMLP.pretrained(ctx)
    out = {}
    for path in paths:
        def build():
            if path == "learned":
                eng.createArclengthTableLearned(mlp)
            else:
                eng.createArclengthTable()

        def step():
            build()
            eng.sample_device(d_off, n_total, None, True, d_out, None, uniform_closed=True)

        for _ in range(warmup):
            step()
        ctx.sync()
        ctx.profile(True)
        ctx.profile_reset()
        ctx.sync()
        ctx.timer_start()
        for _ in range(steps):
            step()
        ms = ctx.timer_stop()
        ctx.sync()
        kernels = {}
        for name in KERNELS:
            t, n = ctx.profile_query(name)
            if n:
                kernels[name] = {"total_ms": t, "launches": n, "avg_ms": t / n, "ms_per_step": t / steps}
        t_smp, _ = ctx.profile_query("k_sample")
        ctx.profile(False)
        # build wall time on its own (a learned build may run its per-bin
        # launches concurrently, so kernel times need not add up to it)
        ctx.sync()
        ctx.timer_start()
        for _ in range(steps):
            build()
        build_ms = ctx.timer_stop() / steps
        st = eng.stats()
        lengths = eng.getAllCurveArclengths()
        # the same step captured once and replayed as one graph launch
        ctx.graph_begin()
        step()
        graph = ctx.graph_end()
        graph.launch()
        ctx.sync()
        ctx.timer_start()
        for _ in range(steps):
            graph.launch()
        graph_ms = ctx.timer_stop()
        graph.close()
        log(f"[bench] {path}: {steps} steps in {ms:.3f} ms (graph replays {graph_ms:.3f} ms); build stats {st}")
        out[path] = {"ms": ms, "ms_per_step": ms / steps, "graph_ms_per_step": graph_ms / steps,
                     "kernels": kernels, "stats": st, "build_ms": build_ms,
                     "sample_ms": t_smp / steps, "lengths": lengths}
        if batch.kind != "ellipse":
            # sampling alone on this table under each Newton start (the
            # step above used the table's default, the plain start)
            by_start = {}
            modes = (("plain", G.GCV_START_PLAIN), ("learned", G.GCV_START_LEARNED),
                     ("quintic", G.GCV_START_QUINTIC))
            for name, mode in modes:
                if path != "learned" and name == "learned":
                    continue
                eng.set_newton_start(mode)
                eng.sample_device(d_off, n_total, None, True, d_out, None, uniform_closed=True)
                ctx.sync()
                ctx.profile(True)
                ctx.profile_reset()
                for _ in range(steps):
                    eng.sample_device(d_off, n_total, None, True, d_out, None, uniform_closed=True)
                ctx.sync()
                t, n = ctx.profile_query("k_sample")
                ctx.profile(False)
                by_start[name] = t / n if n else None
            eng.set_newton_start(G.GCV_START_PLAIN)
            out[path]["sample_ms_by_newton_start"] = by_start
            log(f"[bench] {path}: k_sample ms by Newton start {by_start}")
    eng.close()
    d_off.free()
    d_out.free()
    if mlp is not None:
        mlp.close()
    return out


def comparison(runs, n_int, n_total):
    pl, le = runs["plain"], runs["learned"]

    def side(r):
        return {"build_ms": r["build_ms"], "step_ms": r["ms_per_step"], "sample_ms": r["sample_ms"],
                "graph_step_ms": r["graph_ms_per_step"],
                "quadratures": r["stats"]["quadratures"], "deep_roots": r["stats"]["deep_roots"],
                "build_intervals_per_sec": n_int / (r["build_ms"] * 1e-3),
                "step_intervals_per_sec": n_int / (r["ms_per_step"] * 1e-3),
                "step_samples_per_sec": n_total / (r["ms_per_step"] * 1e-3),
                "kernel_ms_per_step": {k: v["ms_per_step"] for k, v in r["kernels"].items()},
                "sample_ms_by_newton_start": r.get("sample_ms_by_newton_start")}

    return {"plain": side(pl), "learned": side(le),
            "build_speedup": pl["build_ms"] / le["build_ms"],
            "step_speedup": pl["ms_per_step"] / le["ms_per_step"],
            "quadrature_ratio": pl["stats"]["quadratures"] / max(1, le["stats"]["quadratures"]),
            "max_rel_length_diff": float(np.max(np.abs(le["lengths"] - pl["lengths"]) / pl["lengths"]))}


def roofline_of(run, batch, counts, cfg_name, steps):
    """Dominant kernel (FP32 VALU bound): algorithmic FLOP per step from the
    counted quadratures (subdivision) or samples (sampler), divided by that
    kernel's event-timed duration per step."""
    kernels, st = run["kernels"], run["stats"]
    n_total = int(np.sum(counts))
    if batch.kind == "ellipse":
        fnode = f_node_ellipse()
    else:
        ps = np.repeat(batch.degree, batch.intervals_per_curve())
        ds = np.repeat(batch.dim, batch.intervals_per_curve())
        fnode = float(np.mean([f_node_bspline(int(p), int(d)) for p, d in zip(ps[:100000], ds[:100000])]))
    ngl = 3
    sub_ms = kernels.get("k_subdivide", {"ms_per_step": 0.0})["ms_per_step"]
    smp_ms = kernels.get("k_sample", {"ms_per_step": 0.0})["ms_per_step"]
    quad_flop = st["quadratures"] * (ngl * fnode + 5)
    # per sample, cubic-start path: 2 end speeds (Hermite start), one NGL
    # quadrature and its end speed (the FP32 Newton step), one NGL quadrature
    # in FP64; quintic path (ellipse curves with axis ratio >= 0.2 at uniform
    # positions): the quintic start (30 FLOP: powers, basis, combination), the
    # speed at the start and the FP64 quadrature
    cubic = (2 + ngl + 1 + ngl) * fnode + 2 * 5
    quintic = (1 + ngl) * fnode + 5 + 30
    frac_q = 0.0
    if batch.kind == "ellipse":
        ratio = np.minimum(np.abs(batch.a), np.abs(batch.b)) / np.maximum(np.maximum(np.abs(batch.a), np.abs(batch.b)), 1e-300)
        frac_q = float(np.sum(counts[ratio >= 0.2])) / max(1, n_total)
    smp_flop = n_total * (frac_q * quintic + (1.0 - frac_q) * cubic)
    if smp_ms >= sub_ms:
        dom, flop, dur = "k_sample", smp_flop, smp_ms
    else:
        dom, flop, dur = "k_subdivide", quad_flop, sub_ms
    achieved = flop / (dur * 1e-3) / 1e12 if dur > 0 else 0.0
    traffic, traffic_source = None, None
    tf = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tf):
        rec = json.load(open(tf)).get(cfg_name, {}).get("kernels", {}).get(dom)
        if rec:
            traffic = rec["read_bytes"] + rec["write_bytes"]
            traffic_source = ("profiles/traffic.json: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of this "
                              "workload, scaled by the copy-kernel calibration (not measured in this run)")
    # the roof that binds: the quadratures and the sampler's Newton steps run
    # on the FP32 vector ALU (FP64 parts at half rate; non-packed v_fma_f32
    # issues at the published FP32 rate, bench.py --valu-calibration); no
    # matrix cores are involved
    out = {"kernel": dom, "bound": "valu", "compute_unit": "FP32/FP64 vector ALU (no matrix cores)",
           "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
           "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic, "traffic_source": traffic_source,
           "flop_per_step": flop, "sample_flop_cubic_path": cubic, "sample_flop_quintic_path": quintic,
           "quintic_sample_fraction": frac_q, "kernel_ms_per_step": dur,
           "hbm_gbs_of_traffic": (traffic / (dur * 1e-3) / 1e9) if (traffic and dur > 0) else None}
    if dom == "k_sample" and dur > 0:
        # the sampler's HBM side: 8 B per sample (FP64 parameter out) + 4 B per
        # knot interval (count / offset), the algorithmic output bytes
        n_int = int(batch.n_intervals)
        hbm_bytes = 8 * n_total + 4 * n_int
        out["hbm"] = {"algorithmic_bytes": hbm_bytes, "achieved_gbs": hbm_bytes / (dur * 1e-3) / 1e9,
                      "peak_gbs": PEAK_HBM_GBS, "frac": hbm_bytes / (dur * 1e-3) / 1e9 / PEAK_HBM_GBS}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default=os.environ.get("GCV_BENCH_CONFIG", "c1"))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--path", choices=("plain", "learned"), default=None,
                    help="build path whose throughput is `value` (the other is timed beside it); default: "
                         "plain for the 1M-ellipse workload, which is defined on the plain method, "
                         "learned for the B-spline workloads")
    ap.add_argument("--no-compare", action="store_true", help="time only --path")
    ap.add_argument("--no-c3", "--no-seed-comparisons", dest="no_seed_comparisons", action="store_true",
                    help="skip the learned-vs-plain comparisons on the B-spline workloads (c2, c3, c4)")
    ap.add_argument("--no-c0", action="store_true", help="skip the single-ellipse latency entry")
    ap.add_argument("--seed-weights", default=None, help="learned-seed weights (.npz); default: the packaged ones")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only time the library's streaming copy kernel over MIB mebibytes")
    ap.add_argument("--valu-calibration", action="store_true",
                    help="only time the library's VALU rate kernels (FP32 FMA, FP64 FMA, FP32 sqrt)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (curves are independent and would "
              "shard across GPUs by predicted work). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import gcurval_amd as G

    if args.copy_calibration:
        copy_calibration(G, args.copy_calibration, args.steps)
        return
    if args.valu_calibration:
        valu_calibration(G, args.steps)
        return

    t_gen = time.time()
    batch, counts, cfg = workload(args.config)
    n_total = int(np.sum(counts))
    n_int = batch.n_intervals
    log(f"[bench] workload {cfg['workload']}: {batch.n_curve} curves, {n_int} intervals, "
        f"{n_total} samples (generated in {time.time() - t_gen:.1f}s)")

    if args.path is None:
        args.path = "plain" if args.config == "c1" else "learned"
    ctx = G.Context(0)
    n_cu, arch = ctx.device_info()
    other = "plain" if args.path == "learned" else "learned"
    paths = [args.path] if args.no_compare else [args.path, other]
    runs = measure(G, ctx, batch, counts, args.steps, args.warmup, paths, args.seed_weights)
    main_run = runs[args.path]
    ms = main_run["ms"]

    result = {
        "metric": METRIC,
        "value": n_int * args.steps / (ms * 1e-3),
        "unit": "curve-intervals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "graph_ms_per_step": main_run["graph_ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 quadrature, f64 scan and final Newton step",
        "data": "synthetic (seeded)",
        "config": cfg,
        "path": args.path,
        "samples_per_sec": n_total * args.steps / (ms * 1e-3),
        "n_intervals": n_int,
        "n_samples": n_total,
        "build_ms": main_run["build_ms"],
        "sample_ms": main_run["sample_ms"],
        "build_intervals_per_sec": n_int / (main_run["build_ms"] * 1e-3),
        "sample_only_samples_per_sec": n_total / (main_run["sample_ms"] * 1e-3),
        "kernels": main_run["kernels"],
        "build_stats": main_run["stats"],
        "roofline": roofline_of(main_run, batch, counts, cfg["workload"], args.steps),
        "device": {"arch": arch, "n_cu": n_cu},
    }
    if not args.no_compare:
        result["seed_comparison"] = comparison(runs, n_int, n_total)
    del runs, main_run

    if not args.no_seed_comparisons:
        speedups = {}
        for name in ("c2", "c3", "c4"):
            if name == args.config:
                if "seed_comparison" in result:
                    speedups[name] = result["seed_comparison"]["build_speedup"]
                continue
            bk, ck, cfgk = workload(name)
            rk = measure(G, ctx, bk, ck, 3, 1, ["plain", "learned"], args.seed_weights)
            comp = comparison(rk, bk.n_intervals, int(np.sum(ck)))
            comp["workload"] = cfgk
            comp["n_intervals"] = int(bk.n_intervals)
            result[f"{name}_seed_comparison"] = comp
            speedups[name] = comp["build_speedup"]
            log(f"[bench] {name} learned vs plain: build {comp['build_speedup']:.2f}x, "
                f"step {comp['step_speedup']:.2f}x")
            del rk, bk, ck
        result["seed_build_speedups"] = speedups

    if not args.no_c0:
        result["c0_latency"] = c0_latency(G, ctx)
        log(f"[bench] c0 latency {result['c0_latency']['median_ms']:.3f} ms "
            f"(graph {result['c0_latency']['graph_median_ms']:.3f} ms)")

    if not args.no_cpu_baseline:
        nproc, model = cpu_info()
        threads = args.cpu_threads or min(16, nproc or 1)

        def describe(cb, th):
            return (f"first {cb['n_curve']} curves of the same workload ({cb['n_intervals']} intervals, "
                    f"{cb['n_samples']} samples), oracle restatement of ArclengthCalculateCpu, "
                    f"build {cb['build_s']:.2f}s + sampling {cb['sample_s']:.2f}s on {th} thread(s)")
        try:
            cb1 = cpu_baseline(batch, counts, 1, budget_s=12.0)
            cbn = cpu_baseline(batch, counts, threads, budget_s=15.0) if threads > 1 else cb1
            result["cpu_baseline"] = {
                "value": cbn["intervals_per_s"], "unit": "curve-intervals/s", "cores": threads,
                "kind": "port", "samples_per_s": cbn["samples_per_s"], "sample": describe(cbn, threads),
                "single_thread": {"value": cb1["intervals_per_s"], "samples_per_s": cb1["samples_per_s"],
                                  "cores": 1, "sample": describe(cb1, 1)},
                "nproc": nproc, "cpu_model": model,
            }
        except Exception as e:  # the baseline is reported, never required
            result["cpu_baseline"] = {"value": None, "unit": "curve-intervals/s", "cores": 0, "kind": "port",
                                      "sample": f"unavailable: {e}", "nproc": nproc, "cpu_model": model}
    print(json.dumps(result), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
