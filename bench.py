#!/usr/bin/env python3
"""Pressure-solve benchmark on one MI355X.

Workload: the urban block (128 x 128 x 64 cells, h = 2 m, 24 seeded ground boxes,
about 15 % solid, open x with a log-law inflow, outlet anchoring), right-hand
side -div(u) of a seeded perturbed log-law face velocity. A step is one full
pressure solve (PCG with the geometric multigrid V(2,2) red-black preconditioner,
FP64) from the right-hand side resident in HBM to the solution in HBM.

value = solves/s at relative residual 1e-8 (plain FP64 path); the same workload
at 1e-10 and in mixed precision (FP32 V-cycle inside FP64 PCG) is reported
beside it. roofline: the fine-level smoother half-sweep, timed with HIP events
on the grid's stream. cpu_baseline: the oracle (serial FP64 restatement of the
reference's own multigrid PCG) on the same workload, one core.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "cfd-craft_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def cpu_baseline(case, b, t, tol):
    """The oracle on the same system: the reference's multigrid PCG, serial FP64,
    pinned to one core; hierarchy built by a first (cold) solve, then one reused
    solve is timed (performance_probe.rs:110-136 protocol, one sample)."""
    import oracle

    old = os.sched_getaffinity(0)
    core = min(old)
    os.sched_setaffinity(0, {core})
    try:
        g = oracle.Grid(case)
        t0 = time.perf_counter()
        g.mg_stats()  # hierarchy build (cold set-up, not part of the reused solve)
        setup = time.perf_counter() - t0
        t0 = time.perf_counter()
        x, info = g.solve_scaled(b, t, tol, 2000)
        dt = time.perf_counter() - t0
    finally:
        os.sched_setaffinity(0, old)
    return x, dict(
        value=1.0 / dt,
        unit="solves/s",
        cores=1,
        kind="port",
        sample=f"one reused multigrid-PCG solve of the 128x128x64 urban block at rel-res {tol:g} "
        f"(reference's masked-aggregation MG, {info['iterations']} iterations, {dt:.2f} s; hierarchy set-up "
        f"{setup:.2f} s excluded)",
        iterations=info["iterations"],
    )


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the pressure solve of one grid is measured on one MI355X "
              "(see DESIGN.md for how the path would shard)", file=sys.stderr)
        sys.exit(2)

    from cfdgpu import workloads as W

    case, b, t = W.urban_block()
    tol = 1e-8
    base = None
    x_cpu = None
    if not args.no_cpu:
        x_cpu, base = cpu_baseline(case, b, t, tol)

    import cfdgpu

    g = cfdgpu.Grid(case)
    db = cfdgpu.DeviceArray.from_host(b)
    dt_ = cfdgpu.DeviceArray.from_host(t)
    dx = cfdgpu.DeviceArray(g.count)

    def run(steps, warmup, tol_, precision):
        g.set_options(precision)
        for _ in range(warmup):
            g.solve_device(db, dx, tol_, 2000, d_scale=dt_)
        cfdgpu.synchronize()
        lat = []
        infos = []
        t0 = time.perf_counter()
        for _ in range(steps):
            s0 = time.perf_counter()
            infos.append(g.solve_device(db, dx, tol_, 2000, d_scale=dt_))
            lat.append(time.perf_counter() - s0)
        cfdgpu.synchronize()
        el = time.perf_counter() - t0
        return el, lat, infos

    el, lat, infos = run(args.steps, args.warmup, tol, cfdgpu.PRECISION_FP64)
    x_gpu = dx.numpy()
    ms = 1e3 * el / args.steps
    value = args.steps / el
    # beside the headline: 1e-10 plain, and mixed precision at 1e-8 and 1e-10
    extra = {}
    for name, tol_, prec in (("plain_tol_1e-10", 1e-10, cfdgpu.PRECISION_FP64),
                             ("mixed_tol_1e-8", 1e-8, cfdgpu.PRECISION_MIXED),
                             ("mixed_tol_1e-10", 1e-10, cfdgpu.PRECISION_MIXED)):
        k = max(3, args.steps // 2)
        e2, l2, i2 = run(k, 1, tol_, prec)
        extra[name] = dict(solves_per_s=k / e2, iterations=i2[-1]["iterations"],
                           true_rel_residual=i2[-1]["relative_residual"])
    g.set_options(cfdgpu.PRECISION_FP64)
    sec, nbytes = g.profile_kernel(1, reps=50)
    achieved = nbytes / sec / 1e9
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None,
                    kernel="k_fine_gs<double,double> (fine red-black half-sweep)",
                    avg_launch_us=round(sec * 1e6, 3), algorithmic_bytes_per_launch=nbytes)
    # the whole V-cycle and the operator for context
    sv, bv = g.profile_kernel(4, reps=20)
    so, bo = g.profile_kernel(0, reps=50)
    extra["vcycle"] = dict(avg_us=round(sv * 1e6, 2), algorithmic_GBps=round(bv / sv / 1e9, 1))
    extra["operator"] = dict(avg_us=round(so * 1e6, 2), algorithmic_GBps=round(bo / so / 1e9, 1))
    lat_ms = sorted(1e3 * v for v in lat)
    p = lambda q: lat_ms[min(len(lat_ms) - 1, int(math.ceil(q * len(lat_ms))) - 1)]  # noqa: E731
    out = {
        "metric": "Poisson solves/s at rel-res 1e-8 + HBM GB/s %peak; p50/p99 latency served vs plain",
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded urban block and log-law velocity field)",
        "config": {
            "workload": "urban block 128x128x64, h=2 m, 24 seeded boxes (~15% solid), open x outlet, plain FP64 "
            "PCG + geometric MG V(2,2) red-black, rel-res 1e-8",
            "cells": case.count,
            "iterations": infos[-1]["iterations"],
            "true_rel_residual": infos[-1]["relative_residual"],
            "latency_ms": {"p50": round(p(0.5), 3), "p99": round(p(0.99), 3)},
        },
        "roofline": roofline,
        "extra": extra,
    }
    if base is not None:
        err = float(np.linalg.norm(x_gpu - x_cpu) / np.linalg.norm(x_cpu))
        base_out = {k: base[k] for k in ("value", "unit", "cores", "kind", "sample")}
        out["cpu_baseline"] = base_out
        out["config"]["field_rel_error_vs_oracle"] = err
        out["config"]["oracle_iterations"] = base["iterations"]
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
