#!/usr/bin/env python3
"""Benchmark: 3D 256^3 order-3 FP64 build + 2^26 evaluations on one MI355X.

Workload (the reference configuration the metric is quoted on): values
sin(x) exp(-(y-1/2)^2/0.05) cos(4 pi z) + N(0, 1e-3^2) on a 256^3 mesh over
[0,2pi) x [0,1] x [0,1], dim 0 periodic, dims 1-2 clamped; 2^26 uniform
query points; outputs value, d/dx, d/dy, d/dz.  Seeded synthetic data
(bsplineinterpolation_amd/datagen.py).

One step = the full constructor (knots, collocation LU, control-point
solve; intp_build) + one evaluation call over all 2^26 points, with the
values and query points already resident in HBM.  ``value`` = query points
per second over the whole step.

The eval kernel dominates the step; its roofline entry uses the
algorithmic bytes 2^26*(3*8 + 4*8) + 256^3*8 (coordinates in, outputs out,
table read once) over its average duration, timed by the library with HIP
events around the kernel on the stream it runs on (INTP_EVAL_TIMED).  Its
measured HBM traffic per launch comes from the committed rocprofv3
FETCH_SIZE/WRITE_SIZE summary (profiles/bench_3d256_k3_hbm_traffic.md).
The CPU baseline times the compiled reference
(oracle/_ref/libintp_ref.so; the oracle restatement if that is absent):
the constructor at full size with the reference's 8 pool workers, and
evaluation of a sample of the same queries with one thread per host core.
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

HBM_PEAK_GBS = 8000.0
EVAL_KERNEL = "eval_patch_kernel<3, 3, true, true, double>"
TRAFFIC_PROFILE = os.path.join("profiles", "bench_3d256_k3_hbm_traffic.md")


def profiled_traffic(kernel: str):
    """HBM bytes per launch of `kernel` (calibrated read + write) from the
    committed rocprofv3 summary, or None."""
    path = os.path.join(ROOT, TRAFFIC_PROFILE)
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 6 and cells[0] == "`%s`" % kernel:
            return (float(cells[4]) + float(cells[5])) * 1e9
    return None


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def _compiler() -> str:
    """The compiler that built the CPU baseline library (g++ here), as
    `g++ --version` reports it."""
    import subprocess
    try:
        return subprocess.run(["g++", "--version"], capture_output=True, text=True, timeout=10).stdout.splitlines()[0]
    except (OSError, IndexError, subprocess.SubprocessError):
        return "unknown"


def _cores() -> int:
    n = os.cpu_count() or 1
    for var in ("OMP_NUM_THREADS",):
        v = os.environ.get(var)
        if v and v.isdigit():
            n = min(n, int(v))
    try:
        n = min(n, len(os.sched_getaffinity(0)))
    except AttributeError:
        pass
    return max(1, n)


def rel_stats(got, want):
    d = got - want
    return dict(norm=float(math.sqrt(np.sum(d * d) / np.sum(want * want))),
                max=float(np.max(np.abs(d)) / np.max(np.abs(want))),
                rms=float(math.sqrt(np.mean(d * d))))


def measure_refined_form(amd, torch, steps: int, warmup: int, n: int, nq: int, sweeps: int = 1):
    """The refined-FP32 form against the FP64 form on the same workload:
    3D n^3 order-5, all clamped, nq value queries (FP32 results for the
    refined form), plus the 1D periodic order-6 n=8192 spline whose solve
    keeps FP64 factors.  Errors are the refined results against the FP64
    form's (which match the FP64 reference to 1e-12): max|d|/max|ref| and
    rms(d)/rms(ref)."""
    from bsplineinterpolation_amd import datagen
    w = datagen.workload_3d_256_quintic(n=n, nq=nq)
    dev = torch.device("cuda:0")
    f_d = torch.from_numpy(w["f"]).to(dev)
    q_d = torch.from_numpy(np.ascontiguousarray(w["queries"].T)).to(dev)
    outs = {"f64": torch.empty((1, nq), dtype=torch.float64, device=dev),
            "f32_refined": torch.empty((1, nq), dtype=torch.float32, device=dev)}
    stream = torch.cuda.current_stream()
    res = {}
    for form in ("f64", "f32_refined"):
        def step(ev=None):
            if ev is not None:
                ev[0].record(stream)
            s = amd.InterpolationFunction(5, f_d, periodic=w["periodic"], ranges=w["ranges"], precision=form,
                                          sweeps=sweeps)
            if ev is not None:
                ev[1].record(stream)
            s.eval(q_d, w["orders"], out=outs[form], timed=ev is not None)
            if ev is not None:
                ev[2].record(stream)
            return s
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        amd.eval_timings()
        events = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
        t0 = time.perf_counter()
        for i in range(steps):
            step(events[i])
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        tim = amd.eval_timings()
        res[form] = {"ms_per_step": ms, "value": nq / (ms * 1e-3),
                     "build_ms": float(np.mean([e[0].elapsed_time(e[1]) for e in events])),
                     "eval_ms": float(np.mean([e[1].elapsed_time(e[2]) for e in events])),
                     "eval_grouping_ms": float(np.mean([t[0] for t in tim])),
                     "eval_kernel_ms": float(np.mean([t[1] for t in tim]))}
    ref = outs["f64"].double()
    d = outs["f32_refined"].double() - ref
    err = {"max": float(d.abs().max() / ref.abs().max()),
           "rms": float(torch.sqrt((d * d).mean() / (ref * ref).mean()))}
    # 1D periodic order 6, n = 8192 (FP64 factors in the refined solve)
    w1 = datagen.workload_1d_periodic_sextic()
    q1 = torch.from_numpy(np.ascontiguousarray(w1["queries"])).to(dev)
    s64 = amd.InterpolationFunction(6, w1["f"], periodic=w1["periodic"], ranges=w1["ranges"])
    s32 = amd.InterpolationFunction(6, w1["f"], periodic=w1["periodic"], ranges=w1["ranges"],
                                    precision="f32_refined", sweeps=sweeps)
    r64 = s64.eval(q1)
    d1 = s32.eval(q1) - r64
    err1 = {"max": float(d1.abs().max() / r64.abs().max()), "rms": float(torch.sqrt((d1 * d1).mean() / (r64 * r64).mean()))}
    b = res["f32_refined"]
    N = n ** 3
    # algorithmic bytes (S sweeps): FP32 solves (S+1) 2 D N 4, FP64 residuals
    # S (2 D N 8 + 8 N), FP32 updates added on store (S-1) N 4, the gradient
    # table (c and the last correction in, 16-byte cells out: N (8 + 16))
    # and the last update (N 12); evaluation nq (24 + 4) + N 4
    b_build = refined_build_bytes(3, N, sweeps)
    b_eval = nq * (24 + 4) + N * 4
    return {"workload": "3D %d^3 order-5 clamped, %d queries (value); FP32 control points, FP32 results" % (n, nq),
            "sweeps": sweeps, "refined_f32": res["f32_refined"], "f64_form": res["f64"],
            "speedup_vs_f64_form": res["f64"]["ms_per_step"] / b["ms_per_step"],
            "error_vs_f64": err,
            "build_gbs": b_build / (b["build_ms"] * 1e-3) / 1e9, "build_algorithmic_bytes": b_build,
            "eval_kernel_gbs": b_eval / (b["eval_kernel_ms"] * 1e-3) / 1e9, "eval_algorithmic_bytes": b_eval,
            "periodic_1d_order6_n8192": {"error_vs_f64": err1, "note": "periodic order >= 5: FP64 factors"}}


SERVE_RATES = (1e2, 3e2, 1e3, 3e3, 1e4)


def measure_copy_gbs(torch, nbytes: int = 1 << 31, reps: int = 5) -> float:
    """The device's streaming bandwidth as torch measures it: a copy of
    `nbytes` (read + written) timed with events, in GB/s."""
    a = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda")
    a.fill_(1.0)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    return 2 * nbytes / (ms * 1e-3) / 1e9


def refined_build_bytes(dim: int, N: int, sweeps: int) -> int:
    """Algorithmic HBM bytes of a refined build with S sweeps (as the
    serving loop counts them): FP32 solves, FP64 residuals, the FP32
    updates added on store, the gradient table and the last update."""
    S = sweeps
    cell = 8 if dim == 1 else 16
    return ((S + 1) * 2 * dim * N * 4 + S * (2 * dim * N * 8 + 8 * N) + max(S - 1, 0) * N * 4
            + N * ((8 if S > 0 else 4) + cell) + (12 * N if S > 0 else 0))


def measure_refined_headline(amd, torch, steps: int, warmup: int, w, f_d, q_d, out64, f64_ms: float, sweeps: int = 1):
    """The refined-FP32 form on the headline workload (3D 256^3 cubic,
    value + 3 partials at 2^26 points, FP32 results): the same step as the
    FP64 headline, timed the same way, with max/RMS relative error of every
    output against the FP64 results of the headline run."""
    dev = q_d.device
    out32 = torch.empty((out64.shape[0], out64.shape[1]), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()

    def step(ev=None):
        if ev is not None:
            ev[0].record(stream)
        s = amd.InterpolationFunction(3, f_d, periodic=w["periodic"], ranges=w["ranges"], precision="f32_refined",
                                      sweeps=sweeps)
        if ev is not None:
            ev[1].record(stream)
        s.eval(q_d, w["orders"], out=out32.t(), timed=ev is not None)
        if ev is not None:
            ev[2].record(stream)
        return s

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    # the same step again with events: build, eval call, and the eval's
    # grouping and kernel (not part of the timed loop above)
    amd.eval_timings()
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
    for i in range(steps):
        step(events[i])
    torch.cuda.synchronize()
    tim = amd.eval_timings()
    breakdown = {"build": float(np.mean([e[0].elapsed_time(e[1]) for e in events])),
                 "eval_call": float(np.mean([e[1].elapsed_time(e[2]) for e in events])),
                 "eval_grouping": float(np.mean([t[0] for t in tim])),
                 "eval_kernel": float(np.mean([t[1] for t in tim]))}
    errs = {}
    names = ["value", "d/dx", "d/dy", "d/dz"]
    for o in range(out64.shape[1]):
        ref = out64[:, o]
        d = out32[:, o].double() - ref
        errs[names[o]] = {"max": float(d.abs().max() / ref.abs().max()),
                          "rms": float(torch.sqrt((d * d).mean() / (ref * ref).mean()))}
    del stream
    return {"workload": "the headline step (3D 256^3 cubic build + 2^26 x value, d/dx, d/dy, d/dz) in the refined-FP32 "
                        "form, %d sweep(s), FP32 results" % sweeps,
            "ms_per_step": ms, "breakdown_ms": breakdown, "f64_ms_per_step": f64_ms, "speedup_vs_f64": f64_ms / ms,
            "error_vs_f64": errs,
            "partials": "value and first partials from the gradient table (FP32 backward differences of the "
                        "FP64-corrected control points)"}


def measure_small(amd, torch, steps: int):
    """The small 2D workload: 64x64 order-3 clamped FP64, 1e6 queries of
    value and d/dx.  Launch-latency bound, so reported in microseconds:
    build (full constructor), eval call, and the two together, each timed
    with events over ``steps`` back-to-back repetitions on one stream."""
    from bsplineinterpolation_amd import datagen
    w = datagen.workload_2d_64_cubic()
    dev = torch.device("cuda:0")
    f = torch.from_numpy(w["f"]).to(dev)
    q = torch.from_numpy(np.ascontiguousarray(w["queries"].T)).to(dev)
    out = torch.empty((2, q.shape[0]), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()

    def timed(fn):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) * 1e3 / steps

    g = amd.InterpolationFunction(3, f, periodic=w["periodic"], ranges=w["ranges"])
    build_us = timed(lambda: amd.InterpolationFunction(3, f, periodic=w["periodic"], ranges=w["ranges"]))
    eval_us = timed(lambda: g.eval(q, orders=w["orders"], out=out))
    step_us = timed(lambda: amd.InterpolationFunction(3, f, periodic=w["periodic"], ranges=w["ranges"]).eval(
        q, orders=w["orders"], out=out))
    B_solve, B_eval = 2 * 2 * 4096 * 8, 1_000_000 * (16 + 16) + 4096 * 8
    return dict(workload="2D 64x64 order-3 clamped FP64 build + 1e6 queries x (value, d/dx)",
                build_us=build_us, eval_us=eval_us, step_us=step_us, points_per_s=1e6 / (step_us * 1e-6),
                eval_gbs=B_eval / (eval_us * 1e-6) / 1e9, algorithmic_bytes=B_solve + B_eval)


def measure_batched(amd, torch, steps: int):
    """The batched form: 4096 mixed 1D/2D/3D FP64 splines (orders 1..5,
    ~25% periodic dimensions), built and evaluated at 256 points each per
    run; runs replay one captured hipGraph after an autotuned first run.
    Throughput counts control points solved and points evaluated."""
    from bsplineinterpolation_amd import datagen
    w = datagen.workload_batch_mixed()
    dev = torch.device("cuda:0")
    vals = torch.from_numpy(w["values"]).to(dev)
    pts = torch.from_numpy(np.ascontiguousarray(w["points"])).to(dev)
    po = w["point_offsets"]
    out = torch.empty(int(po[-1]), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b = amd.Batch(w["specs"])
    create_ms = (time.perf_counter() - t0) * 1e3
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        t0 = time.perf_counter()
        b.run(vals, pts, po, out=out, autotune=True)
        stream.synchronize()
        tune_ms = (time.perf_counter() - t0) * 1e3
        b.run(vals, pts, po, out=out, graph=True)  # capture
        stream.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        t0 = time.perf_counter()
        for _ in range(steps):
            b.run(vals, pts, po, out=out, graph=True)
        e1.record(stream)
        stream.synchronize()
        wall_ms = (time.perf_counter() - t0) * 1e3 / steps
    run_ms = e0.elapsed_time(e1) / steps
    # the same batch with long 1D lines split into parallel segments
    # (INTP_BATCH_SPLIT_LINES: within 1e-12 of the sequential solve)
    seq_out = out.clone()
    with torch.cuda.stream(stream):
        b.run(vals, pts, po, out=out, graph=True, split_lines=True)  # capture
        stream.synchronize()
        e0.record(stream)
        for _ in range(steps):
            b.run(vals, pts, po, out=out, graph=True, split_lines=True)
        e1.record(stream)
        stream.synchronize()
    split_ms = e0.elapsed_time(e1) / steps
    d = (out - seq_out).abs().max().item() / seq_out.abs().max().item()
    N = int(b.total)
    Q = int(po[-1])
    dims = [len(sp["shape"]) for sp in w["specs"]]
    sizes = np.diff(b.offsets)
    b_solve = int(sum(2 * D * n * 8 for D, n in zip(dims, sizes)))
    b_eval = int(sum((D * 8 + 8) * q for D, q in zip(dims, np.diff(po)))) + N * 8
    return {"workload": "4096 FP64 splines: 2048 1D (n 16..4096), 1536 2D (16..256 per side), 512 3D (8..64); "
                        "orders 1..5, ~25% periodic dims; 256 value queries each",
            "control_points": N, "points": Q, "run_ms": run_ms, "wall_ms_per_run": wall_ms,
            "splines_per_s": len(dims) / (run_ms * 1e-3), "control_points_per_s": N / (run_ms * 1e-3),
            "points_per_s": Q / (run_ms * 1e-3), "algorithmic_bytes": b_solve + b_eval,
            "gbs": (b_solve + b_eval) / (run_ms * 1e-3) / 1e9, "create_ms": create_ms, "autotune_ms": tune_ms,
            "tuning": b.tuning().split(" | ")[0], "graph": True,
            "split_lines": {"run_ms": split_ms, "points_per_s": Q / (split_ms * 1e-3),
                            "gbs": (b_solve + b_eval) / (split_ms * 1e-3) / 1e9,
                            "max_rel_diff_vs_sequential": d},
            "parity": "control points and values of all 4096 splines bit-identical to the single-spline path, "
                      "autotuned and graph-replayed (tests/test_gpu_configs.py::test_gpu_config3_batch_4096_mixed)"}


def measure_serving(amd, torch, seconds: float, rates, burst: int, batch_gbs=None):
    """The serving workload: for each arrival rate, Poisson requests for
    ``seconds`` (30% builds of 2D/3D grids of 32..1024 / 32..192 per side,
    orders 3..5, FP64 or refined FP32; 70% evaluations of 1e3..1e6 points,
    half with first partials) submitted from this thread at their arrival
    times to one server (16 GiB HBM budget, 4 streams); per rate:
    requests/s, algorithmic bytes/s and p50/p99 latency per type.  Then a
    burst of ``burst`` requests submitted at once (saturation), and the same
    requests run directly back to back on one stream (the offline batch of
    the same work): serving / offline throughput is the batching ratio.
    Submissions use the C entry points with arguments prepared beforehand,
    so that the client thread keeps up with 1e4 requests/s."""
    import ctypes as C
    from bsplineinterpolation_amd import _lib, datagen
    lib = _lib.load()
    dev = torch.device("cuda:0")
    max_n, max_q = 192 ** 3, 1_000_000
    g = torch.Generator(device="cpu").manual_seed(7)
    pool_v = torch.randn(max_n, generator=g, dtype=torch.float64).to(dev)
    pool_q = torch.rand((max_q, 3), generator=g, dtype=torch.float64).to(dev)
    nbuf = 256
    outs = torch.empty((nbuf, 4 * max_q), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    dp = C.POINTER(C.c_double)
    vptr, qptr = C.cast(pool_v.data_ptr(), dp), C.cast(pool_q.data_ptr(), dp)
    optr = [C.cast(outs[i].data_ptr(), dp) for i in range(nbuf)]

    def prepare(reqs):
        args = []
        for i, rq in enumerate(reqs):
            if rq["kind"] == "build":
                D = len(rq["shape"])
                desc = _lib.GridDesc()
                desc.dim, desc.order = D, rq["order"]
                for d in range(D):
                    desc.n[d], desc.xmin[d], desc.xmax[d] = rq["shape"][d], 0.0, 1.0
                prec = _lib.INTP_PREC_F64 if rq["precision"] == "f64" else _lib.INTP_PREC_F32_REFINED
                args.append((desc, prec))
            else:
                D = len(reqs[rq["target"]]["shape"])
                ords = [0] * D + (sum(([int(k == d) for k in range(D)] for d in range(D)), []) if rq["derivs"] else [])
                nout = len(ords) // D
                args.append(((C.c_int32 * len(ords))(*ords), nout, optr[i % nbuf]))
        return args

    rid = C.c_int64()

    def submit(srv, reqs, args, ids, i):
        rq = reqs[i]
        if rq["kind"] == "build":
            desc, prec = args[i]
            _lib.check(lib.intp_serve_build(srv.h, C.byref(desc), vptr, prec, 1, C.byref(rid)))
        else:
            ords, nout, out = args[i]
            rc = lib.intp_serve_eval(srv.h, ids[rq["target"]], rq["npts"], qptr, 1, 3, nout, ords, out,
                                     rq["npts"], 1, 0, C.byref(rid))
            if rc == _lib.INTP_E_EVICTED:  # target evicted before this submit
                ids[i] = None
                return
            _lib.check(rc)
        ids[i] = rid.value

    def summarize(st, wall):
        res = {}
        for k in ("build", "eval"):
            t = st[k]
            res[k] = dict(completed=t["completed"], failed=t["failed"], per_s=t["completed"] / wall,
                          gbs=t["bytes"] / wall / 1e9, p50_ms=t["p50_ms"], p99_ms=t["p99_ms"], max_ms=t["max_ms"])
        res["gbs"] = (st["build"]["bytes"] + st["eval"]["bytes"]) / wall / 1e9
        res.update(evictions=st["evictions"], admission_waits=st["admission_waits"], wall_s=wall)
        return res

    srv = amd.Server(hbm_budget=16 << 30, streams=4, max_inflight=256)
    phases = []
    for ph, rate in enumerate(rates):
        reqs = datagen.workload_serving(rate=rate, seconds=seconds, phase=ph + 1)
        args = prepare(reqs)
        ids = [None] * len(reqs)
        t0 = time.perf_counter()
        lag = 0.0
        for i, rq in enumerate(reqs):
            while time.perf_counter() - t0 < rq["t"]:
                pass
            lag = max(lag, time.perf_counter() - t0 - rq["t"])
            submit(srv, reqs, args, ids, i)
        srv.drain()
        st = srv.stats(reset=True)
        res = summarize(st, st["last_done_s"] - st["first_submit_s"])
        res.update(rate=rate, requests=len(reqs), max_submit_lag_ms=lag * 1e3)
        phases.append(res)
        print("bench: serving rate %g/s: %d requests, build p50/p99 %.2f/%.2f ms, eval p50/p99 %.2f/%.2f ms, "
              "submit lag <= %.2f ms" % (rate, len(reqs), res["build"]["p50_ms"], res["build"]["p99_ms"],
                                         res["eval"]["p50_ms"], res["eval"]["p99_ms"], lag * 1e3),
              file=sys.stderr, flush=True)
    # saturation: every request of the burst queued at once; run twice and
    # the second measured, as the offline run below is (its first pass
    # warms the memory pool with blocks of these shapes)
    reqs = datagen.workload_serving(rate=None, count=burst, phase=len(rates) + 1)
    args = prepare(reqs)
    for _ in range(2):
        ids = [None] * len(reqs)
        t0 = time.perf_counter()
        for i in range(len(reqs)):
            submit(srv, reqs, args, ids, i)
        submit_s = time.perf_counter() - t0
        srv.drain()
        st = srv.stats(reset=True)
        sat = summarize(st, st["last_done_s"] - st["first_submit_s"])
        sat["submit_s"] = submit_s
    sat["note"] = "second of two identical bursts (the first warms the pool, as for the offline run)"
    srv.close()
    # The same requests as an offline batch: direct calls with no dispatcher
    # at full concurrency, spread round-robin over the server's four
    # streams, each evaluation ordered after its build by an event; and,
    # for reference, the same calls serialized on one stream.
    ctx = amd.Context.default()

    def offline(nstreams):
        streams = [torch.cuda.Stream() for _ in range(nstreams)]
        handles = [C.c_void_p(st.cuda_stream) for st in streams]
        hs = [None] * len(reqs)
        built = [None] * len(reqs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i, rq in enumerate(reqs):
            k = i % nstreams
            sh = handles[k]
            if rq["kind"] == "build":
                desc, prec = args[i]
                h = C.c_void_p()
                if prec == _lib.INTP_PREC_F64:
                    _lib.check(lib.intp_build(ctx.h, C.byref(desc), vptr, _lib.INTP_MEM_DEVICE, C.byref(h), sh))
                else:
                    _lib.check(lib.intp_build_refined(ctx.h, C.byref(desc), vptr, _lib.INTP_MEM_DEVICE, 1,
                                                      C.byref(h), sh))
                hs[i] = h
                built[i] = torch.cuda.Event()
                built[i].record(streams[k])
            else:
                ords, nout, out = args[i]
                streams[k].wait_event(built[rq["target"]])
                _lib.check(lib.intp_eval(hs[rq["target"]], rq["npts"], qptr, 1, 3, nout, ords, out, rq["npts"], 1,
                                         0, sh))
        torch.cuda.synchronize()
        secs = time.perf_counter() - t0
        for h in hs:
            if h is not None:
                lib.intp_spline_release(h)
        return secs

    offline(4)  # warm: grids and pool blocks of these shapes
    off_s = offline(4)
    off1_s = offline(1)
    nbytes = 0.0
    for rq in reqs:
        if rq["kind"] == "build":
            D, N = len(rq["shape"]), int(np.prod(rq["shape"]))
            nbytes += (2 * D * N * 8 if rq["precision"] == "f64" else
                       2 * 2 * D * N * 4 + (2 * D * N * 8 + 8 * N) + N * 4)
        else:
            b = reqs[rq["target"]]
            D, N = len(b["shape"]), int(np.prod(b["shape"]))
            nbytes += rq["npts"] * (D * 8 + (1 + (D if rq["derivs"] else 0)) * 8) + N * (
                8 if b["precision"] == "f64" else 4)
    offline = dict(wall_s=off_s, gbs=nbytes / off_s / 1e9, streams=4,
                   one_stream=dict(wall_s=off1_s, gbs=nbytes / off1_s / 1e9))
    out = dict(workload="Poisson arrivals, %g s per rate; 30%% builds (2D sides 32..1024 or 3D sides 32..192, "
                        "orders 3..5, FP64 or refined FP32 with 1 sweep), 70%% evaluations of 1e3..1e6 points "
                        "(half with first partials) on the 16 latest builds; 16 GiB HBM budget, 4 streams" % seconds,
               rates=phases, saturation=dict(requests=burst, **sat), offline_same_requests=offline,
               ratio_vs_offline=sat["gbs"] / offline["gbs"],
               ratio_vs_offline_note="served burst over the same requests run directly at full concurrency "
                                     "(4 streams, no dispatcher) on the same algorithmic bytes")
    if batch_gbs:
        out["ratio_vs_batched_form_gbs"] = sat["gbs"] / batch_gbs
    return out


def cpu_baseline(w, gpu_out_sample, sample_mt: int, sample_st: int):
    """Reference CPU path, bounded sample.  Returns (dict, error stats)."""
    import oracle
    from oracle import oracle as omod
    kind = "reference" if omod.ref_available() else "port"
    Cls = oracle.RefSpline if kind == "reference" else oracle.OracleSpline
    kw = {} if kind == "reference" else {"nthreads": 8}
    t0 = time.perf_counter()
    s = Cls(3, w["f"], periodic=w["periodic"], ranges=w["ranges"], **kw)
    t_build = time.perf_counter() - t0
    q = np.ascontiguousarray(w["queries"][:, :sample_mt].T)
    cores = _cores()
    t0 = time.perf_counter()
    ref = s.eval(q, w["orders"], nthreads=cores)
    t_mt = (time.perf_counter() - t0) / sample_mt
    t0 = time.perf_counter()
    s.eval(q[:sample_st], w["orders"], nthreads=1)
    t_st = (time.perf_counter() - t0) / sample_st
    nq = w["nq"]
    step_s = t_build + t_mt * nq
    errs = [rel_stats(gpu_out_sample[i], ref[i]) for i in range(len(w["orders"]))]
    return dict(value=nq / step_s, unit="points/s", cores=cores, kind=kind, cpu_model=_cpu_model(),
                compiler=_compiler() + " -O3 -DNDEBUG -DINTP_MULTITHREAD -DINTP_PERIODIC_NO_DUMMY_POINT "
                                       "-DINTP_CELL_LAYOUT (oracle/Makefile)",
                sample=("constructor at full size (%.3f s, 8 pool workers + caller); evaluation of value + 3 "
                        "partials on %d of the 2^26 queries with %d threads (%.3f us/point), scaled to 2^26; "
                        "single-thread %.3f us/point on %d points" % (t_build, sample_mt, cores, t_mt * 1e6,
                                                                       t_st * 1e6, sample_st)),
                build_s=t_build, eval_us_per_point_mt=t_mt * 1e6, eval_us_per_point_st=t_st * 1e6,
                step_s=step_s), errs


def measure_memory_probes(torch, n: int = 1 << 26, reps: int = 5):
    """Reference rates for the eval kernel's memory pattern: 32-byte rows
    (value + 3 partials per point) copied in order and scattered to a
    random permutation (the grouped kernel's result stores)."""
    dev = torch.device("cuda:0")
    g = torch.Generator(device="cpu").manual_seed(3)
    src = torch.rand((n, 4), generator=g, dtype=torch.float64).to(dev)
    dst = torch.empty_like(src)
    perm = torch.randperm(n, generator=g).to(dev)
    probes = (("sequential_rows", lambda: dst.copy_(src)),
              ("random_row_scatter", lambda: dst.index_copy_(0, perm, src)))
    res = {}
    for name, fn in probes:
        for _ in range(2):
            fn()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        torch.cuda.synchronize()
        ev[0].record()
        for _ in range(reps):
            fn()
        ev[1].record()
        torch.cuda.synchronize()
        ms = ev[0].elapsed_time(ev[1]) / reps
        res[name] = dict(ms=ms, row_gbs=2 * n * 32 / ms / 1e6)
        print("bench: probe %s %.3f ms" % (name, ms), file=sys.stderr, flush=True)
    return dict(rows=n, row_bytes=32, probes=res)


def measure_drop_in_timing():
    """Per-call cost of the C++ drop-in's scalar members and eval_proxy
    objects (tests/cpp/drop_in_timing.cpp), with the compiled reference's
    operator() timed on the same meshes and points when oracle/_ref is
    built (the CPU comparison of this leg)."""
    import subprocess
    exe = os.path.join(ROOT, "tests", "cpp", "_build", "drop_in_timing")
    if not os.path.exists(exe):
        raise SystemExit("bench.py: %s is not built (python -c 'import __graft_entry__ as g; g.build()')" % exe)
    ref = os.path.join(ROOT, "oracle", "_ref", "libintp_ref.so")
    cmd = [exe] + ([ref] if os.path.exists(ref) else [])
    r = subprocess.run(cmd, capture_output=True, text=True, check=True)
    rows = [json.loads(line) for line in r.stdout.splitlines() if line.startswith("{")]
    return dict(workload="2^15 random points sorted by cell, 32 passes, meshes of about 2^18 points (1D 262144, "
                         "2D 512^2, 3D 64^3, 4D 24^4), even dimensions periodic; reference operator() on one thread",
                rows=rows)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--queries", type=int, default=1 << 26)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-sample", type=int, default=1 << 22)
    ap.add_argument("--no-forms", action="store_true", help="skip the refined-FP32 form measurement")
    ap.add_argument("--only-forms", action="store_true",
                    help="measure only the refined-FP32 form (profiling runs); prints its object")
    ap.add_argument("--only-refined-headline", action="store_true",
                    help="measure only the refined-FP32 form of the headline step, beside the FP64 step (profiling "
                         "runs); prints its object")
    ap.add_argument("--only-serving", action="store_true", help="measure only the serving loop; prints its object")
    ap.add_argument("--serve-seconds", type=float, default=2.0, help="seconds of arrivals per serving rate")
    ap.add_argument("--serve-burst", type=int, default=1500, help="requests of the saturation burst")
    ap.add_argument("--only-batched", action="store_true",
                    help="measure only the batched form (4096 mixed splines); prints its object")
    ap.add_argument("--drop-in-timing", action="store_true",
                    help="per-call cost of the C++ drop-in's scalar members and proxies; prints their object")
    ap.add_argument("--memory-probes", action="store_true",
                    help="time 32-byte row copies in order and scattered (2^26 rows); prints their object")
    ap.add_argument("--calibration-copy", action="store_true",
                    help="copy 256 MiB once before timing (scales rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: runs on one GPU only; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.drop_in_timing:
        print(json.dumps(measure_drop_in_timing()))
        return
    import torch
    import bsplineinterpolation_amd as amd
    from bsplineinterpolation_amd import datagen

    n, nq = args.n, args.queries
    if args.memory_probes:
        print(json.dumps(measure_memory_probes(torch)))
        return
    if args.only_batched:
        print(json.dumps(measure_batched(amd, torch, args.steps)))
        return
    if args.only_serving:
        print(json.dumps(measure_serving(amd, torch, args.serve_seconds, SERVE_RATES, args.serve_burst)))
        return
    if args.only_forms:
        print(json.dumps(measure_refined_form(amd, torch, args.steps, args.warmup, n, nq)))
        return
    if args.only_refined_headline:
        w = datagen.workload_3d_256_cubic(n=n, nq=nq)
        dev = torch.device("cuda:0")
        f_d = torch.from_numpy(w["f"]).to(dev)
        q_d = torch.from_numpy(np.ascontiguousarray(w["queries"].T)).to(dev)
        out64 = torch.empty((nq, 4), dtype=torch.float64, device=dev)

        def step64():
            s = amd.InterpolationFunction(3, f_d, periodic=w["periodic"], ranges=w["ranges"])
            s.eval(q_d, w["orders"], out=out64.t())
            return s

        for _ in range(args.warmup):
            step64()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step64()
        torch.cuda.synchronize()
        f64_ms = (time.perf_counter() - t0) * 1e3 / args.steps
        print(json.dumps(measure_refined_headline(amd, torch, args.steps, args.warmup, w, f_d, q_d, out64, f64_ms)))
        return
    print("bench: generating inputs (%d^3 mesh, %d queries)" % (n, nq), file=sys.stderr, flush=True)
    w = datagen.workload_3d_256_cubic(n=n, nq=nq)
    w["nq"] = nq
    dev = torch.device("cuda:0")
    f_d = torch.from_numpy(w["f"]).to(dev)
    # one record per query, like the reference's per-point calls:
    # coordinates [nq, 3] in, (value, d/dx, d/dy, d/dz) [nq, 4] out
    q_d = torch.from_numpy(np.ascontiguousarray(w["queries"].T)).to(dev)
    out_d = torch.empty((nq, 4), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    ctx = amd.Context(0)
    per, rg, orders = w["periodic"], w["ranges"], w["orders"]
    stream = torch.cuda.current_stream()

    def step(ev=None):
        if ev is not None:
            ev[0].record(stream)
        s = amd.InterpolationFunction(3, f_d, periodic=per, ranges=rg, ctx=ctx)
        if ev is not None:
            ev[1].record(stream)
        s.eval(q_d, orders, out=out_d.t())
        if ev is not None:
            ev[2].record(stream)
        return s

    if args.calibration_copy:
        a = torch.ones(1 << 25, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    print("bench: warmup %d" % args.warmup, file=sys.stderr, flush=True)
    keep = None
    for _ in range(args.warmup):
        keep = step()
    torch.cuda.synchronize()
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    print("bench: timing %d steps" % args.steps, file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        keep = step(events[i])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    build_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in events]))
    eval_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in events]))
    # grouping and kernel alone: timed calls run them one after the other
    # (untimed calls overlap the grouping of one segment with the kernels
    # of the previous one), library events on the call's stream
    amd.eval_timings()
    for _ in range(3):
        keep.eval(q_d, orders, out=out_d.t(), timed=True)
    timings = amd.eval_timings()
    assert len(timings) == 3
    group_ms = float(np.mean([t[0] for t in timings]))
    kernel_ms = float(np.mean([t[1] for t in timings]))

    # control-point solve alone: template built once, interpolate timed
    tmpl = amd.InterpolationFunctionTemplate(3, (n, n, n), periodic=per, ranges=rg, ctx=ctx)
    tmpl.interpolate(f_d)
    torch.cuda.synchronize()
    sev = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(5)]
    for e in sev:
        e[0].record(stream)
        sp = tmpl.interpolate(f_d)
        e[1].record(stream)
    torch.cuda.synchronize()
    solve_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in sev]))
    del sp

    N = n ** 3
    b_eval = nq * (3 * 8 + 4 * 8) + N * 8
    b_solve = 2 * 3 * N * 8
    eval_gbs = b_eval / (eval_ms * 1e-3) / 1e9
    kernel_gbs = b_eval / (kernel_ms * 1e-3) / 1e9
    solve_gbs = b_solve / (solve_ms * 1e-3) / 1e9
    value = nq / (ms_per_step * 1e-3)

    result = {
        "metric": "3D 256^3 cubic: evals/s + solve GB/s vs HBM peak; served p50/p99 latency",
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SplitMix64/Box-Muller generator, bsplineinterpolation_amd/datagen.py)",
        "config": {"workload": "3D %d^3 order-3 FP64 build (dim 0 periodic, dims 1-2 clamped) + %d queries x "
                               "(value, d/dx, d/dy, d/dz); each step runs the full constructor (control-point "
                               "solve from the values) while the grid's knots and LU factors come from the "
                               "library's cache of recent grids after the first step" % (n, nq),
                   "mesh": [n, n, n], "order": 3, "periodic": per, "queries": nq, "outputs": 4},
        "breakdown_ms": {"build_full_constructor": build_ms, "solve_only": solve_ms, "eval_call": eval_ms,
                         "eval_grouping": group_ms, "eval_kernel": kernel_ms},
        "solve": {"gbs": solve_gbs, "frac_hbm": solve_gbs / HBM_PEAK_GBS, "algorithmic_bytes": b_solve},
        "eval_call": {"gbs": eval_gbs, "frac_hbm": eval_gbs / HBM_PEAK_GBS, "algorithmic_bytes": b_eval},
        "roofline": {"bound": "hbm", "kernel": EVAL_KERNEL, "achieved": kernel_gbs,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": kernel_gbs / HBM_PEAK_GBS,
                     "algorithmic_bytes": b_eval, "avg_launch_ms": kernel_ms,
                     "traffic": profiled_traffic(EVAL_KERNEL) if (n, nq) == (256, 1 << 26) else None,
                     "traffic_source": TRAFFIC_PROFILE},
    }
    # the same figures against the box's measured streaming bandwidth
    copy_gbs = measure_copy_gbs(torch)
    result["measured_copy_gbs"] = copy_gbs
    result["roofline"]["frac_of_measured_copy"] = kernel_gbs / copy_gbs
    result["solve"]["frac_of_measured_copy"] = solve_gbs / copy_gbs
    result["eval_call"]["frac_of_measured_copy"] = eval_gbs / copy_gbs

    if not args.no_forms:
        print("bench: refined-FP32 form (3D order 5)", file=sys.stderr, flush=True)
        del keep
        keep = None
        result["forms"] = measure_refined_form(amd, torch, max(3, args.steps // 2), 2, n, nq)
        result["forms"]["small_2d64"] = measure_small(amd, torch, 20)
        print("bench: refined-FP32 form on the headline workload", file=sys.stderr, flush=True)
        result["forms"]["refined_headline"] = measure_refined_headline(amd, torch, max(3, args.steps // 2), 2, w, f_d,
                                                                       q_d, out_d, ms_per_step)
        print("bench: batched form (4096 mixed splines)", file=sys.stderr, flush=True)
        result["forms"]["batched"] = measure_batched(amd, torch, max(5, args.steps))
        print("bench: serving loop", file=sys.stderr, flush=True)
        result["forms"]["serving"] = measure_serving(amd, torch, args.serve_seconds, SERVE_RATES, args.serve_burst,
                                                     result["forms"]["batched"]["gbs"])

    if not args.no_cpu:
        m = min(args.cpu_sample, nq)
        print("bench: CPU baseline (sample %d points)" % m, file=sys.stderr, flush=True)
        sample = out_d[:m].t().cpu().numpy()
        cb, errs = cpu_baseline(w, sample, m, min(1 << 15, m))
        result["cpu_baseline"] = cb
        result["speedup_vs_cpu"] = value / cb["value"]
        result["error_vs_cpu_reference"] = {name: e for name, e in zip(["value", "d/dx", "d/dy", "d/dz"], errs)}
    del keep
    print(json.dumps(result))


if __name__ == "__main__":
    main()
