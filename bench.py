#!/usr/bin/env python3
"""Pressure-solve benchmark on one MI355X.

Workload: the urban block (128 x 128 x 64 cells, h = 2 m, 24 seeded ground boxes,
about 15 % solid, open x with a log-law inflow, outlet anchoring), right-hand
side -div(u) of a seeded perturbed log-law face velocity. A step is one full
pressure solve (PCG with the red-black multigrid V(2,2) preconditioner on the
reference's aggregation hierarchy, FP64) from the right-hand side resident in
HBM to the solution in HBM.

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


def served(requests=512, n_grids=16, max_active=8, seed=4):
    """Served mix: resident grids of 64^3 .. 192x192x96 cells (seeded boxes, 5-25 %
    solid), tolerances 1e-6/1e-8/1e-10. The same request list runs (1) plain:
    one FP64 solve after another on the caller's thread, (2) served: FP64 and
    mixed precision through the continuous-batching server, closed loop (all
    requests queued at once: capacity) and (3) open loop: Poisson arrivals at
    80 % of the served mixed capacity, for per-request latency p50/p99 measured
    from arrival to completion. Grid set-up is excluded (resident grids)."""
    import cfdgpu
    from cfdgpu import workloads as W

    pool = [W.served_request(seed, k) for k in range(n_grids)]
    server = cfdgpu.Server(max_active=max_active)
    grids = [server.admit(c) for c, _, _, _ in pool]
    bufs = [(cfdgpu.DeviceArray.from_host(b), cfdgpu.DeviceArray.from_host(t), cfdgpu.DeviceArray(g.count))
            for (_, b, t, _), g in zip(pool, grids)]
    rng = np.random.default_rng(seed)
    reqs = [(int(rng.integers(0, n_grids)), W.SERVED_TOLERANCES[int(rng.integers(0, 3))]) for _ in range(requests)]
    cells = sum(grids[k].count for k, _ in reqs)

    def pct(v, q):
        v = sorted(v)
        return v[min(len(v) - 1, int(math.ceil(q * len(v))) - 1)]

    # warm every grid in both precisions (hierarchy and graphs ready)
    for prec in (cfdgpu.PRECISION_FP64, cfdgpu.PRECISION_MIXED):
        for k, g in enumerate(grids):
            g.set_options(prec)
            g.solve_device(bufs[k][0], bufs[k][2], 1e-8, 3000, d_scale=bufs[k][1])
    cfdgpu.synchronize()
    out = {"requests": requests, "grids": n_grids, "max_active": max_active,
           "shapes": [list(s) for s in W.SERVED_SHAPES], "tolerances": W.SERVED_TOLERANCES,
           "mean_cells": round(cells / requests)}
    # (1) plain: sequential FP64 solves
    for g in grids:
        g.set_options(cfdgpu.PRECISION_FP64)
    lat = []
    t0 = time.perf_counter()
    for k, tol in reqs:
        s0 = time.perf_counter()
        grids[k].solve_device(bufs[k][0], bufs[k][2], tol, 3000, d_scale=bufs[k][1])
        lat.append(time.perf_counter() - s0)
    el = time.perf_counter() - t0
    out["plain_fp64"] = {"solves_per_s": round(requests / el, 1), "p50_ms": round(1e3 * pct(lat, 0.5), 3),
                         "p99_ms": round(1e3 * pct(lat, 0.99), 3)}
    # (2) served, closed loop
    cap = {}
    for name, prec in (("served_fp64", cfdgpu.PRECISION_FP64), ("served_mixed", cfdgpu.PRECISION_MIXED)):
        t0 = time.perf_counter()
        tks = [server.submit(grids[k], bufs[k][0], bufs[k][2], tol, 3000, prec, d_scale=bufs[k][1]) for k, tol in reqs]
        server.drain()
        el = time.perf_counter() - t0
        for tk in tks:
            server.result(tk)
        cap[name] = requests / el
        out[name] = {"solves_per_s": round(cap[name], 1)}
    # (3) open loop at 80 % of served mixed capacity
    lam = 0.8 * cap["served_mixed"]
    arrivals = np.cumsum(rng.exponential(1.0 / lam, requests))
    tks = []
    t0 = time.perf_counter()
    i = 0
    while i < requests:
        now = time.perf_counter() - t0
        while i < requests and arrivals[i] <= now:
            k, tol = reqs[i]
            tks.append(server.submit(grids[k], bufs[k][0], bufs[k][2], tol, 3000, cfdgpu.PRECISION_MIXED,
                                     d_scale=bufs[k][1]))
            i += 1
        server.progress()
    server.drain()
    el = time.perf_counter() - t0
    lat = [server.result(tk)[1] for tk in tks]
    out["open_loop_mixed"] = {"lambda_per_s": round(lam, 1), "achieved_per_s": round(requests / el, 1),
                              "p50_ms": round(1e3 * pct(lat, 0.5), 3), "p99_ms": round(1e3 * pct(lat, 0.99), 3)}
    out["served_mixed_over_plain"] = round(cap["served_mixed"] / out["plain_fp64"]["solves_per_s"], 3)
    server.close()
    return out


def cavity48(steps=5):
    """The closed 48^3 lid cavity of the reference's CPU-runnable case: plain FP64
    IC(0)-PCG at rel-res 1e-10, right-hand side -div(u) of seeded U(-1,1) face
    velocities (the parity tests hold its iteration count to the oracle's)."""
    import cfdgpu
    from cfdgpu import workloads as W

    case, b, t = W.lid_cavity_48()
    g = cfdgpu.Grid(case)
    db = cfdgpu.DeviceArray.from_host(b)
    ds = cfdgpu.DeviceArray.from_host(t)
    dx = cfdgpu.DeviceArray(g.count)
    g.solve_device(db, dx, 1e-10, 1200, d_scale=ds)
    cfdgpu.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        info = g.solve_device(db, dx, 1e-10, 1200, d_scale=ds)
    cfdgpu.synchronize()
    el = (time.perf_counter() - t0) / steps
    return {"solves_per_s": round(1.0 / el, 2), "ms_per_solve": round(el * 1e3, 2), "iterations": info["iterations"],
            "true_rel_residual": info["relative_residual"]}


# The reference's published single-thread CPU times (ms, reused solve; AMD EPYC 9V74):
# docs/evidence/pressure-multigrid.md:34-38 (manufactured kernels, tol 1e-11, IC(0) / MG)
# and docs/evidence/native-cpu-performance.md:87 (canyon-93k IC(0) solve, median warm).
PUBLISHED_CPU_MS = {
    "cavity24": ([24, 2, 24], 1.292, 0.928),
    "cavity48": ([48, 2, 48], 13.007, 3.646),
    "cavity96": ([96, 2, 96], 102.410, 16.864),
    "cube24": ([24, 24, 24], 19.565, 13.059),
    "cube48": ([48, 48, 48], 390.576, 152.088),
}
CANYON93K_IC0_CPU_MS = 393.146


def published_fixtures(reps=3):
    """The fixtures whose reference CPU times are published, on the GPU with the
    reference's protocol: the manufactured field sin(2 pi x/Lx) + 0.3 cos(4 pi z/Lz)
    (centered), b = A x built on the device, tol 1e-11, one cold solve, then the
    median of `reps` reused solves (inputs resident in HBM). IC(0), the reference's
    multigrid cycle (its iteration counts) and the red-black cycle; canyon-93k
    with IC(0) (its published 166 iterations)."""
    import json as _json

    import cfdgpu
    from cfdgpu import workloads as W
    from cfdgpu.case import IC0, MULTIGRID, Case

    def timed(case, cycle=None):
        g = cfdgpu.Grid(case)
        if cycle is not None:
            g.set_options(cfdgpu.PRECISION_FP64, cycle=cycle)
        ex = g.center_nullspaces(W.manufactured(case, W.fluid_mask(case)))
        db = cfdgpu.DeviceArray.from_host(g.apply(ex))
        dx = cfdgpu.DeviceArray(g.count)
        g.solve_device(db, dx, 1e-11, 10000)  # cold
        cfdgpu.synchronize()
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            info = g.solve_device(db, dx, 1e-11, 10000)
            cfdgpu.synchronize()
            times.append((time.perf_counter() - t0) * 1e3)
        x = dx.numpy()
        return {"gpu_ms": round(float(np.median(times)), 3), "iterations": info["iterations"],
                "rms_error": float(np.sqrt(np.mean((x - ex) ** 2)))}

    out = {}
    for name, (cells, ic0_ms, mg_ms) in PUBLISHED_CPU_MS.items():
        row = {}
        r = timed(W.cavity(cells, IC0))
        r.update(reference_cpu_ms=ic0_ms, speedup=round(ic0_ms / r["gpu_ms"], 1))
        row["ic0"] = r
        r = timed(W.cavity(cells, MULTIGRID), cfdgpu.CYCLE_REFERENCE)
        r.update(reference_cpu_ms=mg_ms, speedup=round(mg_ms / r["gpu_ms"], 1))
        row["mg_reference_cycle"] = r
        r = timed(W.cavity(cells, MULTIGRID), cfdgpu.CYCLE_RB)
        r.update(reference_cpu_ms=mg_ms, speedup=round(mg_ms / r["gpu_ms"], 1))
        row["mg_rb_cycle"] = r
        out[name] = row
    with open(os.path.join(ROOT, "tests", "golden", "canyon-93k.json")) as f:
        case = Case.from_json(_json.load(f))
    case.preconditioner = IC0
    r = timed(case)
    r.update(reference_cpu_ms=CANYON93K_IC0_CPU_MS, speedup=round(CANYON93K_IC0_CPU_MS / r["gpu_ms"], 1))
    out["canyon93k_ic0"] = r
    return out


def anisotropic(steps=3):
    """A flat solid-free lid cavity (96x16x96 cells over 1 x 1 x 0.1 m: z couplings
    100x stiffer) at rel-res 1e-10 with the reference's own multigrid cycle (z-line
    block smoothers on semicoarsened levels, multigrid/line.rs), the zebra cycle
    (the same hierarchy and blocks swept by colour classes), the red-black cycle
    and IC(0); right-hand side A applied to sin(0.73 i) + 0.3 cos(0.17 i),
    centered, as in the reference's anisotropy suite."""
    import cfdgpu
    from cfdgpu.case import IC0, MULTIGRID, PERIODIC, WALL, Case

    out = {"cells": [96, 16, 96], "size_m": [1.0, 1.0, 0.1]}
    for name, pc, cycle in (("reference_cycle", MULTIGRID, cfdgpu.CYCLE_REFERENCE),
                            ("zebra_cycle", MULTIGRID, cfdgpu.CYCLE_ZEBRA), ("rb_cycle", MULTIGRID, cfdgpu.CYCLE_RB),
                            ("ic0", IC0, None)):
        case = Case(cells=[96, 16, 96], size_m=[1.0, 1.0, 0.1], boundaries=[WALL, PERIODIC, WALL],
                    velocity_m_s=[1.0, 0.0, 0.0], preconditioner=pc)
        g = cfdgpu.Grid(case)
        if cycle is not None:
            g.set_options(cfdgpu.PRECISION_FP64, cycle=cycle)
        i = np.arange(g.count, dtype=float)
        ex = g.center_nullspaces(np.sin(0.73 * i) + 0.3 * np.cos(0.17 * i))
        db = cfdgpu.DeviceArray.from_host(g.apply(ex))
        dx = cfdgpu.DeviceArray(g.count)
        for _ in range(2):
            g.solve_device(db, dx, 1e-10, 3000)
        cfdgpu.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            info = g.solve_device(db, dx, 1e-10, 3000)
        cfdgpu.synchronize()
        el = (time.perf_counter() - t0) / steps
        x = dx.numpy()
        out[name] = {"solves_per_s": round(1.0 / el, 2), "iterations": info["iterations"],
                     "true_rel_residual": info["relative_residual"],
                     "recovery_rms": float(np.sqrt(np.mean((x - ex) ** 2)))}
        if name == "reference_cycle":
            out["block_smoothers"] = len(g.multigrid_stats().get("block_smoothers", []))
    return out


def campaign(directions=(0.0, 90.0, 200.0, 290.0), solves=20, tol=1e-8, seed=5):
    """A bounded sample of the street-canyon wind-direction campaign at its full
    size (384x384x160, 23.6M cells, 400 convex prisms between streets, about
    24 % solid, x and y open): 4 of the 36 directions, 20 of the 200 chained
    right-hand sides each. Per direction (its own operator) the chain
    b_{k+1} = b_k + 0.02 rms(b_k) xi, xi ~ N(0,1) on the fluid cells (solid cells
    have no equation; their entries stay zero) is solved cold (x = 0) and warm
    (from the grid's resident previous solution), plain FP64 and mixed. The
    right-hand sides are resident before timing. The last warm mixed field of
    each direction is compared with a cold FP64 solve at 1e-12 (the oracle itself
    is too slow at this size; the parity tests hold the FP64 path to it)."""
    import cfdgpu
    from cfdgpu import workloads as W

    rng = np.random.default_rng(seed)
    totals = {}
    errors = []
    setups = []
    solid = None
    for d in directions:
        case, b, t = W.street_canyon(d)
        fluid = W.fluid_mask(case)
        solid = 1.0 - float(fluid.mean())
        began = time.perf_counter()
        g = cfdgpu.Grid(case)
        cfdgpu.synchronize()
        setups.append(time.perf_counter() - began)
        chain = [b]
        for _ in range(solves - 1):
            prev = chain[-1]
            rms = math.sqrt(float(np.mean(prev[fluid] ** 2)))
            chain.append(prev + np.where(fluid, 0.02 * rms * rng.standard_normal(prev.size), 0.0))
        dbs = [cfdgpu.DeviceArray.from_host(v) for v in chain]
        del chain
        dx = cfdgpu.DeviceArray(g.count)
        for name, prec, warm in (("cold_fp64", cfdgpu.PRECISION_FP64, False), ("warm_fp64", cfdgpu.PRECISION_FP64, True),
                                 ("cold_mixed", cfdgpu.PRECISION_MIXED, False),
                                 ("warm_mixed", cfdgpu.PRECISION_MIXED, True)):
            g.set_options(prec, warm_start=warm)
            g.solve_device(dbs[0], dx, tol, 3000)  # warm-up (and the first resident solution)
            cfdgpu.synchronize()
            its = 0
            t0 = time.perf_counter()
            for db in dbs:
                its += g.solve_device(db, dx, tol, 3000)["iterations"]
            cfdgpu.synchronize()
            el = time.perf_counter() - t0
            tot = totals.setdefault(name, [0.0, 0, 0])
            tot[0] += el
            tot[1] += its
            tot[2] += len(dbs)
        x_warm = dx.numpy()
        g.set_options(cfdgpu.PRECISION_FP64)
        ref = cfdgpu.DeviceArray(g.count)
        g.solve_device(dbs[-1], ref, 1e-12, 3000)
        xr = ref.numpy()
        errors.append(float(np.linalg.norm(x_warm - xr) / np.linalg.norm(xr)))
        del g, dbs, dx, ref
        print(f"[campaign] direction {d}: set-up {setups[-1]:.2f} s", file=sys.stderr, flush=True)
    out = {"directions": list(directions), "solves_per_direction": solves, "tol": tol,
           "grid": "384x384x160 street canyon, 400 prisms, x/y open", "cells": 384 * 384 * 160,
           "solid_fraction": round(solid, 4), "cold_setup_s_mean": round(float(np.mean(setups)), 3)}
    for name, (el, its, k) in totals.items():
        out[name] = {"solves_per_s": round(k / el, 2), "mean_iterations": round(its / k, 2)}
    out["last_warm_mixed_field_rel_error_max"] = max(errors)
    return out


def large(steps=3):
    """The 256x256x128 prism city (8.4M cells, 200 convex 4-8-vertex prisms, about
    20 % solid): plain FP64 and mixed precision at 1e-8 and 1e-10; each field is
    compared with the FP64 solve at 1e-12 (itself held to the oracle by the parity
    tests); fine smoother and operator bandwidth at a size beyond the 256 MB MALL."""
    import cfdgpu
    from cfdgpu import workloads as W

    case, b, t = W.urban_prisms()
    # Grid::build's topology (rasterization, face codes, components) on the host
    # and on the device, and a whole cold set-up each way
    t0 = time.perf_counter()
    cfdgpu.case_topology(case)
    topo_host = time.perf_counter() - t0
    t0 = time.perf_counter()
    cfdgpu.case_topology_device(case)
    topo_dev = time.perf_counter() - t0
    os.environ["CFD_GPU_SETUP"] = "0"
    try:
        t0 = time.perf_counter()
        gh = cfdgpu.Grid(case)
        setup_host = time.perf_counter() - t0
    finally:
        os.environ.pop("CFD_GPU_SETUP", None)
    del gh
    t0 = time.perf_counter()
    g = cfdgpu.Grid(case)
    setup = time.perf_counter() - t0
    db = cfdgpu.DeviceArray.from_host(b)
    ds = cfdgpu.DeviceArray.from_host(t)
    dx = cfdgpu.DeviceArray(g.count)
    g.set_options(cfdgpu.PRECISION_FP64)
    ref_info = g.solve_device(db, dx, 1e-12, 3000, d_scale=ds)
    ref = dx.numpy()
    out = {"cells": g.count, "fluid_cells": g.fluid_count, "solid_fraction": round(1 - g.fluid_count / g.count, 4),
           "cold_setup_s": round(setup, 3), "cold_setup_host_topology_s": round(setup_host, 3),
           "topology_host_s": round(topo_host, 3), "topology_device_s": round(topo_dev, 3),
           "reference_iterations_1e-12": ref_info["iterations"]}
    for tol in (1e-8, 1e-10):
        for name, prec in (("plain", cfdgpu.PRECISION_FP64), ("mixed", cfdgpu.PRECISION_MIXED)):
            g.set_options(prec)
            info = g.solve_device(db, dx, tol, 3000, d_scale=ds)
            cfdgpu.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                info = g.solve_device(db, dx, tol, 3000, d_scale=ds)
            cfdgpu.synchronize()
            el = time.perf_counter() - t0
            x = dx.numpy()
            out[f"{name}_tol_{tol:g}"] = {"solves_per_s": round(steps / el, 2), "iterations": info["iterations"],
                                          "true_rel_residual": info["relative_residual"],
                                          "field_rel_error": float(np.linalg.norm(x - ref) / np.linalg.norm(ref))}
    g.set_options(cfdgpu.PRECISION_FP64)
    for name, k in (("fine_gs_fp64", 1), ("operator", 0), ("axpy", 2)):
        sec, nb = g.profile_kernel(k, reps=20)
        out[name] = {"avg_us": round(sec * 1e6, 2), "algorithmic_GBps": round(nb / sec / 1e9, 1),
                     "frac": round(nb / sec / 1e9 / HBM_PEAK_GBS, 4)}
    g.set_options(cfdgpu.PRECISION_MIXED)
    sec, nb = g.profile_kernel(1, reps=20)
    out["fine_gs_mixed"] = {"avg_us": round(sec * 1e6, 2), "algorithmic_GBps": round(nb / sec / 1e9, 1),
                            "frac": round(nb / sec / 1e9 / HBM_PEAK_GBS, 4)}
    return out


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_k_rb_gs_urban_block_128x128x64.txt")


def measured_traffic():
    """HBM bytes per launch of the fine half-sweep from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE, each scaled by its copy-kernel factor)."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":")[1])
    except OSError:
        pass
    return None


def pmc_probe():
    """The kernels the --pmc passes count: 20 fine half-sweeps of the urban block
    (FP64) and 5 elementwise products c = 2 a over 64 Mi floats (256 MiB read,
    256 MiB written), which calibrate FETCH_SIZE and WRITE_SIZE."""
    import torch

    import cfdgpu
    from cfdgpu import workloads as W

    case, b, t = W.urban_block()
    g = cfdgpu.Grid(case)
    g.set_options(cfdgpu.PRECISION_FP64)
    sec, nb = g.profile_kernel(1, reps=20)
    a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
    c = torch.empty_like(a)
    for _ in range(5):
        torch.mul(a, 2.0, out=c)
    torch.cuda.synchronize()
    print(json.dumps({"fine_gs_avg_us": sec * 1e6, "fine_gs_algorithmic_bytes": nb, "copy_bytes_each_way": 4 << 26}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="only the headline solve and its roofline kernel")
    ap.add_argument("--pmc-probe", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the pressure solve of one grid is measured on one MI355X; "
              "several GPUs would each hold their own grids (one wind direction or served grid per GPU)",
              file=sys.stderr)
        sys.exit(2)

    if args.pmc_probe:
        pmc_probe()
        return

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
    variants = () if args.no_extra else (("plain_tol_1e-10", 1e-10, cfdgpu.PRECISION_FP64),
                                         ("mixed_tol_1e-8", 1e-8, cfdgpu.PRECISION_MIXED),
                                         ("mixed_tol_1e-10", 1e-10, cfdgpu.PRECISION_MIXED))
    for name, tol_, prec in variants:
        k = max(3, args.steps // 2)
        e2, l2, i2 = run(k, 2, tol_, prec)
        extra[name] = dict(solves_per_s=k / e2, iterations=i2[-1]["iterations"],
                           true_rel_residual=i2[-1]["relative_residual"])
    g.set_options(cfdgpu.PRECISION_FP64)
    sec, nbytes = g.profile_kernel(1, reps=50)
    achieved = nbytes / sec / 1e9
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=measured_traffic(),
                    kernel="k_rb_gs<double> (fine red-black half-sweep, colour-separated layout)",
                    avg_launch_us=round(sec * 1e6, 3), algorithmic_bytes_per_launch=nbytes)
    if args.no_extra:
        args.no_served = args.no_large = True
    # the whole V-cycle and the operator for context
    sv, bv = g.profile_kernel(4, reps=20)
    so, bo = g.profile_kernel(0, reps=50)
    extra["vcycle"] = dict(avg_us=round(sv * 1e6, 2), algorithmic_GBps=round(bv / sv / 1e9, 1))
    extra["operator"] = dict(avg_us=round(so * 1e6, 2), algorithmic_GBps=round(bo / so / 1e9, 1))
    # the same system with the IC(0) preconditioner (plain FP64, 1e-10)
    from cfdgpu.case import IC0

    cycles = () if args.no_extra else (("geometric_cycle", cfdgpu.CYCLE_GEOMETRIC, 5),
                                       ("reference_cycle", cfdgpu.CYCLE_REFERENCE, 1))
    case_ic = W.urban_block()[0]
    case_ic.preconditioner = IC0
    if not args.no_extra:
        gi = cfdgpu.Grid(case_ic)
        for _ in range(2):
            gi.solve_device(db, dx, 1e-10, 3000, d_scale=dt_)
        cfdgpu.synchronize()
        t0 = time.perf_counter()
        ii = gi.solve_device(db, dx, 1e-10, 3000, d_scale=dt_)
        cfdgpu.synchronize()
        extra["ic0_tol_1e-10"] = dict(solves_per_s=round(1.0 / (time.perf_counter() - t0), 3),
                                      iterations=ii["iterations"], true_rel_residual=ii["relative_residual"])
        del gi
    # the same system on the other multigrid cycles (plain FP64, 1e-8): the structured
    # geometric one and the reference's own (its iteration count is the oracle's)
    for name, cyc, k in cycles:
        g.set_options(cfdgpu.PRECISION_FP64, cycle=cyc)
        for _ in range(2):  # the second solve captures the iteration graphs it queues
            g.solve_device(db, dx, tol, 3000, d_scale=dt_)
        cfdgpu.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            ic = g.solve_device(db, dx, tol, 3000, d_scale=dt_)
        cfdgpu.synchronize()
        extra[name] = dict(solves_per_s=round(k / (time.perf_counter() - t0), 3), iterations=ic["iterations"],
                           true_rel_residual=ic["relative_residual"])
    g.set_options(cfdgpu.PRECISION_FP64)
    if not args.no_extra:
        extra["cavity_48_ic0"] = cavity48()
        extra["anisotropic_96x16x96"] = anisotropic()
        extra["published_fixtures"] = published_fixtures()
        extra["wind_campaign"] = campaign()
    if not args.no_served:
        extra["served"] = served()
    if not args.no_large:
        extra["prisms_256x256x128"] = large()
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
            "PCG + multigrid V(2,2) red-black on the split 2x2x2 aggregation hierarchy, rel-res 1e-8",
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
