#!/usr/bin/env python3
"""Pressure-solve benchmark on one MI355X.

Headline workload: one direction (200 degrees) of the street-canyon wind
campaign at its full size: 384x384x160 cells (23.6M, h = 1 m), 400 convex
4-8-vertex prisms on 20 x 20 lots between 3 m streets (about 20 % solid), x and
y open with the log-law wind, ground_open_top z. A step is one pressure solve
(PCG with the red-black multigrid V(2,2) preconditioner on the reference's
aggregation hierarchy, plain FP64, from x = 0) of one right-hand side of the
campaign's chain b_{k+1} = b_k + 0.02 rms(b_k) xi, every right-hand side
distinct and resident in HBM before timing; value = solves/s at relative
residual 1e-8. Mixed precision (FP32 V-cycle inside FP64 PCG) and warm starts
from the resident previous solution run on the same chain beside it.

Beside the headline: the campaign served concurrently (directions' grids
admitted under an HBM budget that holds some of them at once, the others
queued and built while the admitted ones solve; each direction a warm-started
mixed-precision chain, the last steps against the oracle), the prism city, the urban block, the lid cavity, the
served mix of uneven grids and tolerances, and the published fixtures.

roofline: the kernel class with the largest share of a solve's device time,
measured inside solves of the same chain (every launch of the class bracketed
by a pair of HIP events on the grid's stream): algorithmic bytes over summed
launch durations. traffic: HBM bytes per launch of that kernel from the
committed rocprofv3 --pmc summary. cpu_baseline: the oracle (the reference's
serial FP64 multigrid PCG restated in C++), one core of this host, on the
urban block: one cold solve and the median of three reused ones.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extra]
                      [--campaign-directions D] [--campaign-steps S] [--campaign-resident R]
(the full campaign: --campaign-directions 36 --campaign-steps 200 --campaign-resident 0)
"""
import argparse
import json
import math
import multiprocessing
import os
import statistics
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "cfd-craft_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
HEADLINE_DIRECTION = 200.0
HEADLINE_TOL = 1e-8
METRIC = "Poisson solves/s at rel-res 1e-8 + HBM GB/s %peak; p50/p99 latency served vs plain"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_k_rb_gs2_lin_street_canyon_384x384x160.txt")
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
CUBE48_IC0_MS_PER_ITERATION = 390.576 / 113
CANYON93K_IC0_CPU_MS = 393.146


_T0 = time.time()


def log(msg):
    print(f"[bench {time.time() - _T0:6.1f} s] {msg}", file=sys.stderr, flush=True)


def pct(v, q):
    v = sorted(v)
    return v[min(len(v) - 1, int(math.ceil(q * len(v))) - 1)]


# ---------------------------------------------------------------- CPU side (no GPU touched)
def host_cpu():
    model = None
    try:
        import subprocess

        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except Exception:
        pass
    if model is None:
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        model = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
    return {"model": model, "logical_cpus": os.cpu_count(), "usable_cpus": len(os.sched_getaffinity(0))}


def _timed_protocol(make_grid, solve, reused=3):
    """performance_probe.rs:110-136: build the grid, one cold solve (factor or
    hierarchy built inside it), then `reused` solves; ms of the cold one and the
    median of the reused ones."""
    t0 = time.perf_counter()
    g = make_grid()
    info = solve(g)
    cold = time.perf_counter() - t0
    times = []
    for _ in range(reused):
        t0 = time.perf_counter()
        info = solve(g)
        times.append(time.perf_counter() - t0)
    return cold, statistics.median(times), info


def cpu_baseline():
    """The oracle pinned to one core (the reference's pressure path is serial).
    value: reused multigrid-PCG solves/s of the urban block at rel-res 1e-8; beside
    it the lid cavity (IC(0), 1e-10) and the cube48 IC(0) time per iteration
    against the reference's published 3.46 ms on an EPYC 9V74."""
    import oracle
    from cfdgpu import workloads as W
    from cfdgpu.case import IC0

    old = os.sched_getaffinity(0)
    os.sched_setaffinity(0, {min(old)})
    rows = {}
    try:
        case, b, t = W.urban_block()
        cold, warm, info = _timed_protocol(lambda: oracle.Grid(case), lambda g: g.solve_scaled(b, t, 1e-8, 3000)[1])
        rows["urban_block_128x128x64_mg_1e-8"] = {"cold_s": round(cold, 3), "reused_median_s": round(warm, 3),
                                                  "iterations": info["iterations"]}
        urban = (warm, info["iterations"])
        c1, b1, t1 = W.lid_cavity_48()
        cold, warm, info = _timed_protocol(lambda: oracle.Grid(c1), lambda g: g.solve_scaled(b1, t1, 1e-10, 1200)[1])
        rows["lid_cavity_48_ic0_1e-10"] = {"cold_s": round(cold, 3), "reused_median_s": round(warm, 3),
                                           "iterations": info["iterations"]}
        cube = W.cavity([48, 48, 48], IC0)
        g = oracle.Grid(cube)
        ex = g.center_nullspaces(W.manufactured(cube, g.fluid()))
        bc = g.apply(ex)
        cold, warm, info = _timed_protocol(lambda: oracle.Grid(cube), lambda g: g.solve(bc, 1e-11, 10000)[1])
        rows["cube48_ic0_1e-11"] = {"cold_s": round(cold, 3), "reused_median_s": round(warm, 3),
                                    "iterations": info["iterations"],
                                    "ms_per_iteration": round(1e3 * warm / info["iterations"], 3),
                                    "published_ms_per_iteration_epyc_9v74": round(CUBE48_IC0_MS_PER_ITERATION, 3)}
        del case, b, t
    finally:
        os.sched_setaffinity(0, old)
    warm, its = urban
    return {
        "value": round(1.0 / warm, 4),
        "unit": "solves/s",
        "cores": 1,
        "kind": "port",
        "sample": f"urban block 128x128x64 (24 boxes, outlet x), the oracle's multigrid PCG (the reference's masked "
                  f"aggregation cycle) at rel-res 1e-8, {its} iterations: median of 3 reused solves after a cold one "
                  f"(the 384x384x160 canyon grid needs about 7 GB of host face records and minutes per solve on one "
                  f"core)",
        "host": host_cpu(),
        "rows": rows,
    }


def _c3_oracle_child(path, core):
    """Background process: the oracle's FP64 field of the prism city at rel-res
    1e-10 (one core), saved for the GPU fields' error."""
    try:
        os.sched_setaffinity(0, {core})
    except OSError:
        pass
    import oracle
    from cfdgpu import workloads as W

    case, b, t = W.urban_prisms()
    t0 = time.perf_counter()
    x, info = oracle.Grid(case).solve_scaled(b, t, 1e-10, 3000)
    np.save(path, x)
    with open(path + ".json", "w") as f:
        json.dump({"iterations": info["iterations"], "seconds": time.perf_counter() - t0,
                   "true_rel_residual": info["relative_residual"]}, f)


# ---------------------------------------------------------------- GPU legs
def chain(b, fluid, count, seed=5):
    """b_0 = b, b_{k+1} = b_k + 0.02 rms(b_k) xi on the fluid cells (solid cells
    have no equation; they stay zero)."""
    rng = np.random.default_rng(seed)
    out = [b]
    for _ in range(count - 1):
        prev = out[-1]
        rms = math.sqrt(float(np.mean(prev[fluid] ** 2)))
        out.append(prev + np.where(fluid, 0.02 * rms * rng.standard_normal(prev.size), 0.0))
    return out


def timed_chain(cfdgpu, g, dbs, dx, tol, precision, warm, warmup):
    """Solves of dbs[:warmup] untimed, then of dbs[warmup:], each timed on the host;
    the device is synchronized before and after the timed region."""
    g.set_options(precision, warm_start=warm)
    for db in dbs[:warmup]:
        g.solve_device(db, dx, tol, 3000)
    cfdgpu.synchronize()
    lat, its = [], []
    t0 = time.perf_counter()
    for db in dbs[warmup:]:
        s0 = time.perf_counter()
        its.append(g.solve_device(db, dx, tol, 3000)["iterations"])
        lat.append(time.perf_counter() - s0)
    cfdgpu.synchronize()
    return time.perf_counter() - t0, lat, its


def probe_table(g, db, dx, tol, precision=0):
    """Every kernel class of one solve, timed inside the solve: launches, device
    seconds, algorithmic bytes and GB/s; the classes sorted by time."""
    import cfdgpu

    g.set_options(precision)
    g.solve_device(db, dx, tol, 3000)  # the iteration count is learned: nothing queued past convergence
    rows = {}
    for name in cfdgpu.PROBE_KERNELS:
        r = g.profile_solve(db, dx, tol, name)
        if r["launches"] == 0:
            continue
        rows[name] = {"launches": r["launches"], "device_ms": round(1e3 * r["seconds"], 4),
                      "avg_us": round(1e6 * r["seconds"] / r["launches"], 3),
                      "bytes_per_launch": round(r["bytes"] / r["launches"]),
                      "GBps": round(r["bytes"] / r["seconds"] / 1e9, 1) if r["bytes"] else None,
                      "iterations": r["info"]["iterations"]}
    return dict(sorted(rows.items(), key=lambda kv: -kv[1]["device_ms"]))


def measured_traffic():
    """HBM bytes per launch of the roofline kernel from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE, each scaled by its copy-kernel factor)."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":")[1])
    except OSError:
        pass
    return None


def headline(steps, warmup, extra):
    import cfdgpu
    from cfdgpu import workloads as W

    log("street canyon 384x384x160: workload")
    case, b, t = W.street_canyon(HEADLINE_DIRECTION)
    fluid = W.fluid_mask(case)
    t0 = time.perf_counter()
    g = cfdgpu.Grid(case)
    cfdgpu.synchronize()
    setup = time.perf_counter() - t0
    log(f"set-up {setup:.2f} s; chain of {warmup + steps} right-hand sides")
    rhs = chain(b, fluid, warmup + steps)
    dbs = [cfdgpu.DeviceArray.from_host(v) for v in rhs]
    del rhs
    dx = cfdgpu.DeviceArray(g.count)
    el, lat, its = timed_chain(cfdgpu, g, dbs, dx, HEADLINE_TOL, cfdgpu.PRECISION_FP64, False, warmup)
    x_plain = dx.numpy()
    out = {"setup_s": setup, "elapsed": el, "lat": lat, "its": its, "cells": g.count,
           "solid_fraction": 1.0 - float(fluid.mean())}
    log(f"headline {steps / el:.2f} solves/s")
    modes = {}
    if extra:
        for name, prec, warm in (("mixed_cold", cfdgpu.PRECISION_MIXED, False),
                                 ("fp64_warm", cfdgpu.PRECISION_FP64, True),
                                 ("mixed_warm", cfdgpu.PRECISION_MIXED, True)):
            e2, l2, i2 = timed_chain(cfdgpu, g, dbs, dx, HEADLINE_TOL, prec, warm, warmup)
            modes[name] = {"solves_per_s": round(steps / e2, 2), "mean_iterations": round(float(np.mean(i2)), 2),
                           "p50_ms": round(1e3 * pct(l2, 0.5), 3), "p99_ms": round(1e3 * pct(l2, 0.99), 3)}
            if name == "mixed_warm":
                x_mixed_warm = dx.numpy()
        # errors on the last right-hand side against a tight FP64 solve of it
        g.set_options(cfdgpu.PRECISION_FP64)
        ref = cfdgpu.DeviceArray(g.count)
        g.solve_device(dbs[-1], ref, 1e-12, 3000)
        xr = ref.numpy()
        rel = lambda a: float(np.linalg.norm(a - xr) / np.linalg.norm(xr))  # noqa: E731
        modes["field_rel_error_vs_fp64_1e-12"] = {"fp64_cold_1e-8": rel(x_plain), "mixed_warm_1e-8": rel(x_mixed_warm)}
        del ref
    log("kernel classes inside the solve")
    out["probe"] = probe_table(g, dbs[-1], dx, HEADLINE_TOL)
    out["modes"] = modes
    return out, g, dbs, dx


def _campaign_oracle_child(direction, path, core, timeout=1200.0):
    """Background process (forked before the GPU is first used): the oracle's
    grid of one campaign direction is built at once; when the direction's last
    right-hand side appears at `path` (with `path`.ready), the oracle's FP64
    field of it (rel-res 1e-10, one core) is written beside it."""
    try:
        os.sched_setaffinity(0, {core})
    except OSError:
        pass
    import oracle
    from cfdgpu import workloads as W

    case = W.street_canyon_case(direction)
    grid = oracle.Grid(case)
    del case
    end = time.time() + timeout
    while not os.path.exists(path + ".ready"):
        if time.time() > end:
            return
        time.sleep(0.5)
    b = np.load(path)
    t0 = time.perf_counter()
    x, info = grid.solve(b, 1e-10, 3000)
    np.save(path + ".x.npy", x)
    with open(path + ".json", "w") as f:
        json.dump({"iterations": info["iterations"], "seconds": time.perf_counter() - t0,
                   "true_rel_residual": info["relative_residual"]}, f)


def _campaign_rhs_child(jobs, core):
    """Background process (forked before the GPU is first used): the right-hand
    sides of campaign directions, -div(u) of each direction's seeded log-law
    field, written to their paths (with `path`.ready) for the serving loop."""
    try:
        os.sched_setaffinity(0, {core})
    except OSError:
        pass
    from cfdgpu import workloads as W

    for d, path in jobs:
        _, b, _ = W.street_canyon(d)
        np.save(path, b)
        with open(path + ".ready", "w") as f:
            f.write("1")
        del b


def _wait_file(path, timeout=600.0):
    end = time.time() + timeout
    while not os.path.exists(path + ".ready"):
        if time.time() > end:
            raise RuntimeError(f"{path} was not written")
        time.sleep(0.01)
    return np.load(path)


def campaign(directions, steps, rhs_paths, resident, footprint, lanes=3, seed=11, oracle_paths=()):
    """The wind-direction campaign served concurrently: one street-canyon grid
    (384x384x160, h = 1 m) per inflow direction, admitted by the serving
    library's queued admission under an HBM budget that holds `resident` such
    grids (footprint: one canyon grid's measured bytes), so that a direction's
    grid is built on the server's builder thread while the admitted directions
    solve, and admitted when a finished direction's grid is released. Each
    direction is a chain of `steps` projection solves whose right-hand side is
    the previous one plus a seeded 2 % perturbation (formed in HBM between
    solves, cfd_grid_perturb), each solve warm-started from the direction's
    resident previous pressure, FP32 multigrid inside FP64 PCG at rel-res
    1e-8, all admitted directions at once through the server's batched lanes.
    The first right-hand side of each direction comes from rhs_paths (written
    by background processes). The last right-hand side of every direction is
    solved again in FP64 at 1e-12 on the GPU before its grid is released, and
    written to oracle_paths[i] for the oracle's processes."""
    import cfdgpu
    from cfdgpu import workloads as W

    n = len(directions)
    resident = n if resident <= 0 else min(resident, n)
    budget = int((resident + 0.5) * footprint)
    server = cfdgpu.Server(max_active=n, hbm_budget_bytes=budget, lanes=lanes)
    cases = [W.street_canyon_case(d) for d in directions]
    t0 = time.perf_counter()
    tickets = [server.admit_async(c) for c in cases]
    del cases
    runs = [{"i": i, "d": d, "its": [], "lat": [], "k": 0} for i, d in enumerate(directions)]
    waiting = list(runs)
    live, refs, last, jobs = [], [], [], []
    max_used = 0
    first_submit = None

    def submit(r):
        r["t"] = time.perf_counter()
        r["tk"] = server.submit(r["g"], r["b"][r["k"] % 2], r["x"], HEADLINE_TOL, 3000, cfdgpu.PRECISION_MIXED)

    def finish(r):
        # the last right-hand side solved again in FP64 at 1e-12 (a request of its own)
        r["ref"] = cfdgpu.DeviceArray(r["g"].count)
        r["tk"] = server.submit(r["g"], r["b"][(steps - 1) % 2], r["ref"], 1e-12, 3000, cfdgpu.PRECISION_FP64)
        refs.append(r)

    def release(r):
        # the fields go to the host (compared after the loop), then the grid makes room
        r["fields"] = (r["b"][(steps - 1) % 2].numpy(), r["x"].numpy(), r["ref"].numpy())
        for v in r["b"] + [r["x"], r["ref"]]:
            v.free()
        server.release(r["g"])
        r["released_s"] = time.perf_counter() - t0

    while waiting or live or refs:
        for r in list(waiting):
            g = server.admitted(tickets[r["i"]])
            if g is None:
                continue
            waiting.remove(r)
            r["admitted_s"] = time.perf_counter() - t0
            max_used = max(max_used, server.usage()[0])
            g.set_options(cfdgpu.PRECISION_MIXED, warm_start=True)
            r["g"] = g
            r["b"] = [cfdgpu.DeviceArray.from_host(_wait_file(rhs_paths[r["i"]])), cfdgpu.DeviceArray(g.count)]
            r["x"] = cfdgpu.DeviceArray(g.count)
            if first_submit is None:
                first_submit = time.perf_counter()
            submit(r)
            live.append(r)
        server.progress()
        for r in list(refs):
            if server.finished(r["tk"]):
                server.result(r["tk"])
                refs.remove(r)
                release(r)
        for r in list(live):
            if not server.finished(r["tk"]):
                continue
            info, _, _ = server.result(r["tk"])
            r["lat"].append(time.perf_counter() - r["t"])
            r["its"].append(info["iterations"])
            r["k"] += 1
            if r["k"] == steps:
                live.remove(r)
                finish(r)
                continue
            # b_{k+1} = b_k + 0.02 rms(b_k) xi on the fluid cells
            r["g"].perturb(r["b"][(r["k"] - 1) % 2], r["b"][r["k"] % 2], 0.02,
                           (seed * 1000 + int(r["d"])) * 100000 + r["k"])
            submit(r)
    elapsed = time.perf_counter() - t0
    for r in runs:
        bl, x, xr = r.pop("fields")
        row = {"direction": r["d"], "iterations": r["its"][-1],
               "field_rel_error_vs_gpu_fp64_1e-12": float(np.linalg.norm(x - xr) / np.linalg.norm(xr))}
        if r["i"] < len(oracle_paths):
            path = oracle_paths[r["i"]]
            np.save(path, bl)
            np.save(path + ".gpu.npy", x)
            with open(path + ".ready", "w") as f:
                f.write("1")
            jobs.append((row, path))
        last.append(row)
        del bl, x, xr
    lat_ms = [1e3 * v for r in runs for v in r["lat"]]
    out = {"directions": [float(d) for d in directions], "steps": steps, "solves": n * steps, "lanes": lanes,
           "resident_grids": resident, "hbm_budget_gb": round(budget / 1e9, 2),
           "hbm_max_admitted_gb": round(max_used / 1e9, 2),
           "solves_per_s": round(n * steps / elapsed, 2), "elapsed_s": round(elapsed, 3),
           "first_solve_s": round(first_submit - t0, 3),
           "admitted_s": [round(r["admitted_s"], 3) for r in runs],
           "released_s": [round(r["released_s"], 3) for r in runs],
           "iterations_mean": round(float(np.mean([i for r in runs for i in r["its"]])), 2),
           "latency_ms": {"p50": round(pct(lat_ms, 0.5), 3), "p99": round(pct(lat_ms, 0.99), 3)},
           "last_step": last}
    server.close()
    return out, jobs


def campaign_oracle_errors(jobs, procs, timeout):
    """Joins the campaign's oracle processes; each direction's last-step field
    error against the oracle's FP64 field."""
    end = time.time() + timeout
    for p in procs:
        p.join(timeout=max(1.0, end - time.time()))
    for row, path in jobs:
        if os.path.exists(path + ".json"):
            xo = np.load(path + ".x.npy")
            x = np.load(path + ".gpu.npy")
            with open(path + ".json") as f:
                row["oracle"] = json.load(f)
            row["field_rel_error_vs_oracle_1e-10"] = float(np.linalg.norm(x - xo) / np.linalg.norm(xo))
        else:
            row["field_rel_error_vs_oracle_1e-10"] = None


def prisms(c3_path):
    """The 256x256x128 prism city (8.4M cells, 200 convex 4-8-vertex prisms, about
    20 % solid): plain FP64 and mixed precision at 1e-8 and 1e-10 (the median of
    three rounds of 10 solves after two untimed ones), each field's
    error against the oracle's FP64 solve at 1e-10 (computed on one core beside
    the GPU legs) and against the GPU's own FP64 solve at 1e-12."""
    import cfdgpu
    from cfdgpu import workloads as W

    case, b, t = W.urban_prisms()
    t0 = time.perf_counter()
    g = cfdgpu.Grid(case)
    setup = time.perf_counter() - t0
    db = cfdgpu.DeviceArray.from_host(b)
    ds = cfdgpu.DeviceArray.from_host(t)
    dx = cfdgpu.DeviceArray(g.count)
    g.set_options(cfdgpu.PRECISION_FP64)
    g.solve_device(db, dx, 1e-12, 3000, d_scale=ds)
    ref = dx.numpy()
    out = {"cells": g.count, "solid_fraction": round(1 - g.fluid_count / g.count, 4), "cold_setup_s": round(setup, 3)}
    fields = {}
    for tol in (1e-8, 1e-10):
        for name, prec in (("plain", cfdgpu.PRECISION_FP64), ("mixed", cfdgpu.PRECISION_MIXED)):
            g.set_options(prec)
            for _ in range(2):
                g.solve_device(db, dx, tol, 3000, d_scale=ds)
            cfdgpu.synchronize()
            rates = []
            for _ in range(3):
                t0 = time.perf_counter()
                for _ in range(10):
                    info = g.solve_device(db, dx, tol, 3000, d_scale=ds)
                cfdgpu.synchronize()
                rates.append(10 / (time.perf_counter() - t0))
            x = dx.numpy()
            key = f"{name}_tol_{tol:g}"
            fields[key] = x
            out[key] = {"solves_per_s": round(statistics.median(rates), 2),
                        "rounds_of_10": [round(v, 1) for v in rates], "iterations": info["iterations"],
                        "true_rel_residual": info["relative_residual"],
                        "field_rel_error_vs_gpu_fp64_1e-12": float(np.linalg.norm(x - ref) / np.linalg.norm(ref))}
    g.set_options(cfdgpu.PRECISION_FP64)
    return out, fields


def urban(steps):
    """The urban block (128x128x64, 24 boxes): red-black multigrid at the
    configuration's 1e-10 and at 1e-8, plain and mixed; IC(0) and the reference's
    own cycle on the same system."""
    import cfdgpu
    from cfdgpu import workloads as W
    from cfdgpu.case import IC0

    case, b, t = W.urban_block()
    g = cfdgpu.Grid(case)
    db = cfdgpu.DeviceArray.from_host(b)
    ds = cfdgpu.DeviceArray.from_host(t)
    dx = cfdgpu.DeviceArray(g.count)
    out = {}
    for tol in (1e-10, 1e-8):
        for name, prec in (("plain", cfdgpu.PRECISION_FP64), ("mixed", cfdgpu.PRECISION_MIXED)):
            g.set_options(prec)
            for _ in range(2):
                g.solve_device(db, dx, tol, 3000, d_scale=ds)
            cfdgpu.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                info = g.solve_device(db, dx, tol, 3000, d_scale=ds)
            cfdgpu.synchronize()
            out[f"{name}_tol_{tol:g}"] = {"solves_per_s": round(steps / (time.perf_counter() - t0), 1),
                                          "iterations": info["iterations"]}
    g.set_options(cfdgpu.PRECISION_FP64, cycle=cfdgpu.CYCLE_REFERENCE)
    g.solve_device(db, dx, 1e-8, 3000, d_scale=ds)
    cfdgpu.synchronize()
    t0 = time.perf_counter()
    info = g.solve_device(db, dx, 1e-8, 3000, d_scale=ds)
    out["reference_cycle_tol_1e-8"] = {"solves_per_s": round(1.0 / (time.perf_counter() - t0), 3),
                                       "iterations": info["iterations"]}
    case.preconditioner = IC0
    gi = cfdgpu.Grid(case)
    for prec, name in ((cfdgpu.PRECISION_FP64, "ic0_tol_1e-10"), (cfdgpu.PRECISION_MIXED, "ic0_mixed_tol_1e-10")):
        gi.set_options(prec)
        gi.solve_device(db, dx, 1e-10, 3000, d_scale=ds)
        cfdgpu.synchronize()
        t0 = time.perf_counter()
        info = gi.solve_device(db, dx, 1e-10, 3000, d_scale=ds)
        out[name] = {"solves_per_s": round(1.0 / (time.perf_counter() - t0), 3), "iterations": info["iterations"]}
    return out


def lid_cavity(steps=5):
    """The closed 48^3 lid cavity (the reference's CPU-runnable case): IC(0)-PCG at
    rel-res 1e-10, FP64 and with IC(0) applied in FP32."""
    import cfdgpu
    from cfdgpu import workloads as W

    case, b, t = W.lid_cavity_48()
    g = cfdgpu.Grid(case)
    db = cfdgpu.DeviceArray.from_host(b)
    ds = cfdgpu.DeviceArray.from_host(t)
    dx = cfdgpu.DeviceArray(g.count)
    out = {}
    for name, prec in (("ic0_fp64", cfdgpu.PRECISION_FP64), ("ic0_mixed", cfdgpu.PRECISION_MIXED)):
        g.set_options(prec)
        g.solve_device(db, dx, 1e-10, 1200, d_scale=ds)
        cfdgpu.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            info = g.solve_device(db, dx, 1e-10, 1200, d_scale=ds)
        cfdgpu.synchronize()
        el = (time.perf_counter() - t0) / steps
        out[name] = {"solves_per_s": round(1.0 / el, 2), "ms_per_solve": round(el * 1e3, 2),
                     "iterations": info["iterations"]}
    return out


def served(requests=512, n_grids=16, lanes=3, seed=4):
    """Served mix: resident grids of 64^3 .. 192x192x96 cells (seeded boxes, 5-25 %
    solid), tolerances 1e-6/1e-8/1e-10. The same request list runs (1) plain:
    one FP64 solve after another on the caller's thread, on grids built alone
    (cfd_grid_create); (2) served, batched: FP64 and mixed precision through the
    continuous-batching server, `lanes` concurrent lanes of up to 8 solves per
    precision (one launch per iteration stage for a lane's batch), closed loop
    (all requests queued at once: capacity); (3) the same server without
    batching (every solve on its grid's own stream); (4) open loop: Poisson
    arrivals at 80 % of the batched mixed capacity, per-request latency p50/p99
    from arrival to completion. Grid set-up is excluded (resident grids)."""
    import cfdgpu
    from cfdgpu import workloads as W

    pool = [W.served_request(seed, k) for k in range(n_grids)]
    rng = np.random.default_rng(seed)
    reqs = [(int(rng.integers(0, n_grids)), W.SERVED_TOLERANCES[int(rng.integers(0, 3))]) for _ in range(requests)]
    plain = [cfdgpu.Grid(c) for c, _, _, _ in pool]
    bufs = [(cfdgpu.DeviceArray.from_host(b), cfdgpu.DeviceArray.from_host(t), cfdgpu.DeviceArray(g.count))
            for (_, b, t, _), g in zip(pool, plain)]
    cells = sum(plain[k].count for k, _ in reqs)
    out = {"requests": requests, "grids": n_grids, "lanes": lanes, "batch_max": 8,
           "shapes": [list(s) for s in W.SERVED_SHAPES], "tolerances": W.SERVED_TOLERANCES,
           "mean_cells": round(cells / requests)}
    for k, g in enumerate(plain):
        g.set_options(cfdgpu.PRECISION_FP64)
        g.solve_device(bufs[k][0], bufs[k][2], 1e-8, 3000, d_scale=bufs[k][1])
    cfdgpu.synchronize()
    lat = []
    t0 = time.perf_counter()
    for k, tol in reqs:
        s0 = time.perf_counter()
        plain[k].solve_device(bufs[k][0], bufs[k][2], tol, 3000, d_scale=bufs[k][1])
        lat.append(time.perf_counter() - s0)
    el = time.perf_counter() - t0
    out["plain_fp64"] = {"solves_per_s": round(requests / el, 1), "p50_ms": round(1e3 * pct(lat, 0.5), 3),
                         "p99_ms": round(1e3 * pct(lat, 0.99), 3)}
    del plain

    def closed_loop(server, grids, prec):
        tks = [server.submit(grids[k], bufs[k][0], bufs[k][2], 1e-8, 3000, prec, d_scale=bufs[k][1])
               for k in range(n_grids)]
        server.drain()
        for tk in tks:
            server.result(tk)
        t0 = time.perf_counter()
        tks = [server.submit(grids[k], bufs[k][0], bufs[k][2], tol, 3000, prec, d_scale=bufs[k][1]) for k, tol in reqs]
        server.drain()
        el = time.perf_counter() - t0
        lat = [server.result(tk)[1] for tk in tks]
        return requests / el, lat

    cap = {}
    for name, nl in (("served", lanes), ("streams", 0)):
        server = cfdgpu.Server(max_active=n_grids, lanes=nl)
        grids = [server.admit(c) for c, _, _, _ in pool]
        for pname, prec in (("fp64", cfdgpu.PRECISION_FP64), ("mixed", cfdgpu.PRECISION_MIXED)):
            rate, lat = closed_loop(server, grids, prec)
            cap[f"{name}_{pname}"] = rate
            out[f"{name}_{pname}"] = {"solves_per_s": round(rate, 1), "p50_ms": round(1e3 * pct(lat, 0.5), 3),
                                      "p99_ms": round(1e3 * pct(lat, 0.99), 3)}
        if name == "served":
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
        server.close()
    out["served_mixed_over_plain"] = round(cap["served_mixed"] / out["plain_fp64"]["solves_per_s"], 3)
    out["served_fp64_over_plain"] = round(cap["served_fp64"] / out["plain_fp64"]["solves_per_s"], 3)
    return out


def served_probe(lanes, requests=512, n_grids=16, seed=4):
    """The served mix's batched FP64 closed loop alone (the request list of
    `served`), for kernel traces of the batched launches: with one lane no two
    batched kernels overlap on the device, so each traced duration is the
    kernel's own."""
    import cfdgpu
    from cfdgpu import workloads as W

    pool = [W.served_request(seed, k) for k in range(n_grids)]
    rng = np.random.default_rng(seed)
    reqs = [(int(rng.integers(0, n_grids)), W.SERVED_TOLERANCES[int(rng.integers(0, 3))]) for _ in range(requests)]
    server = cfdgpu.Server(max_active=n_grids, lanes=lanes)
    grids = [server.admit(c) for c, _, _, _ in pool]
    bufs = [(cfdgpu.DeviceArray.from_host(b), cfdgpu.DeviceArray.from_host(t), cfdgpu.DeviceArray(g.count))
            for (_, b, t, _), g in zip(pool, grids)]
    prec = cfdgpu.PRECISION_FP64
    for tk in [server.submit(grids[k], bufs[k][0], bufs[k][2], 1e-8, 3000, prec, d_scale=bufs[k][1])
               for k in range(n_grids)]:
        server.wait(tk)
        server.result(tk)
    t0 = time.perf_counter()
    tks = [server.submit(grids[k], bufs[k][0], bufs[k][2], tol, 3000, prec, d_scale=bufs[k][1]) for k, tol in reqs]
    server.drain()
    el = time.perf_counter() - t0
    its = [server.result(tk)[0]["iterations"] for tk in tks]
    server.close()
    print(json.dumps({"served_probe_lanes": lanes, "requests": requests, "solves_per_s": round(requests / el, 1),
                      "iterations": int(sum(its))}), flush=True)


def published_fixtures(reps=3):
    """The fixtures whose reference CPU times are published, on the GPU with the
    reference's protocol: the manufactured field sin(2 pi x/Lx) + 0.3 cos(4 pi z/Lz)
    (centered), b = A x built on the device, tol 1e-11, one cold solve, then the
    median of `reps` reused solves (inputs resident in HBM). IC(0), the reference's
    multigrid cycle (its iteration counts) and the red-black cycle; canyon-93k
    with IC(0) (its published 166 iterations)."""
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
        g.solve_device(db, dx, 1e-11, 10000)
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
        case = Case.from_json(json.load(f))
    case.preconditioner = IC0
    r = timed(case)
    r.update(reference_cpu_ms=CANYON93K_IC0_CPU_MS, speedup=round(CANYON93K_IC0_CPU_MS / r["gpu_ms"], 1))
    out["canyon93k_ic0"] = r
    return out


def pmc_probe():
    """The kernels the --pmc passes count: 20 fine half-sweeps of the street canyon
    (FP64, 384x384x160) and 5 elementwise products c = 2 a over 64 Mi floats
    (256 MiB read, 256 MiB written), which calibrate FETCH_SIZE and WRITE_SIZE."""
    import torch

    import cfdgpu
    from cfdgpu import workloads as W

    case, b, t = W.street_canyon(HEADLINE_DIRECTION)
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
    ap.add_argument("--no-cpu", action="store_true", help="no CPU baseline and no oracle field of the prism city")
    ap.add_argument("--no-extra", action="store_true", help="only the headline and its kernel table")
    ap.add_argument("--pmc-probe", action="store_true")
    ap.add_argument("--served-probe", type=int, default=None, metavar="LANES",
                    help="only the served mix's batched FP64 closed loop, on LANES lanes (kernel traces)")
    ap.add_argument("--campaign-directions", type=int, default=8,
                    help="wind directions of the served campaign (36: every 10 degrees)")
    ap.add_argument("--campaign-steps", type=int, default=100, help="chained solves per campaign direction")
    ap.add_argument("--campaign-resident", type=int, default=4,
                    help="campaign grids the HBM budget holds at once (0: all); the others queue for admission")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the pressure solve of one grid is measured on one MI355X; "
              "several GPUs would each hold their own grids (one wind direction or served grid per GPU)",
              file=sys.stderr)
        sys.exit(2)
    if args.pmc_probe:
        pmc_probe()
        return
    if args.served_probe is not None:
        served_probe(args.served_probe)
        return

    base = None
    c3_job = None
    campaign_procs, campaign_paths = [], []
    # the campaign's directions: every 360 / n degrees from 20
    directions = [(20.0 + k * 360.0 / args.campaign_directions) % 360.0 for k in range(args.campaign_directions)]
    tmp = tempfile.mkdtemp(prefix="cfdbench")
    c3_path = os.path.join(tmp, "c3_oracle.npy")
    cores = sorted(os.sched_getaffinity(0))
    ctx = multiprocessing.get_context("fork")
    rhs_procs, rhs_paths = [], []
    if not args.no_extra:
        # the campaign's first right-hand sides, formed on the host beside the GPU legs
        rhs_paths = [os.path.join(tmp, f"rhs_{i}.npy") for i in range(len(directions))]
        n_rhs = max(1, min(4, len(directions), len(cores) // 4))
        for w in range(n_rhs):
            jobs = [(d, rhs_paths[i]) for i, d in enumerate(directions) if i % n_rhs == w]
            job = ctx.Process(target=_campaign_rhs_child, args=(jobs, cores[w % len(cores)]), daemon=True)
            job.start()
            rhs_procs.append(job)
    if not args.no_cpu:
        log("cpu baseline (oracle, one core)")
        base = cpu_baseline()
        if not args.no_extra:
            # the oracle's prism-city field, on one core beside the GPU legs (no GPU in the child)
            c3_job = ctx.Process(target=_c3_oracle_child, args=(c3_path, cores[-1]), daemon=True)
            c3_job.start()
            # the oracle's fields of the campaign's last steps, one core each
            for i, d in enumerate(directions[:max(0, min(4, len(directions), len(cores) - 8))]):
                path = os.path.join(tmp, f"campaign_{i}.npy")
                job = ctx.Process(target=_campaign_oracle_child, args=(d, path, cores[-2 - i]), daemon=True)
                job.start()
                campaign_procs.append(job)
                campaign_paths.append(path)

    import cfdgpu  # noqa: F401  (the first GPU use comes after the fork)

    h, g, dbs, dx = headline(args.steps, args.warmup, not args.no_extra)
    value = args.steps / h["elapsed"]
    probe = h["probe"]
    top, row = next(iter(probe.items()))
    kernel_names = {"fine_half_sweep": "k_rb_gs2_lin<double> (fine red-black half-sweep, a pair of cells per thread)",
                    "coarse_half_sweep": "k_agg_gs<double> (coarse half-sweep)",
                    "operator": "k_pcg_op_march (operator marching through z)", "axpy": "k_pcg_axpy"}
    prow = probe[top]
    achieved = prow["GBps"]
    solve_s = h["elapsed"] / args.steps
    total_bytes = sum(r["bytes_per_launch"] * r["launches"] for r in probe.values())
    device_ms = sum(r["device_ms"] for r in probe.values())
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None, "traffic": measured_traffic(),
                "kernel": kernel_names.get(top, top), "grid": "street canyon 384x384x160",
                "avg_launch_us": prow["avg_us"], "launches_per_solve": prow["launches"],
                "algorithmic_bytes_per_launch": prow["bytes_per_launch"],
                "share_of_solve_device_time": round(prow["device_ms"] / device_ms, 4),
                "solve_algorithmic_GBps": round(total_bytes / solve_s / 1e9, 1),
                "solve_frac": round(total_bytes / solve_s / 1e9 / HBM_PEAK_GBS, 4)}
    extra = {"street_canyon_modes": h["modes"], "kernel_classes_in_solve": probe,
             "headline_setup_s": round(h["setup_s"], 3)}
    footprint = cfdgpu.lib().cfd_grid_footprint(g.h)
    del g, dbs, dx
    if not args.no_extra:
        log(f"campaign: {len(directions)} directions x {args.campaign_steps} chained solves, served, "
            f"{args.campaign_resident or len(directions)} grids resident at once")
        extra["campaign"], campaign_jobs = campaign(directions, args.campaign_steps, rhs_paths, args.campaign_resident,
                                                    footprint, oracle_paths=campaign_paths)
        for job in rhs_procs:
            job.join(timeout=60)
        for path in rhs_paths:
            for f in (path, path + ".ready"):
                if os.path.exists(f):
                    os.remove(f)
        log("prism city")
        extra["prisms_256x256x128"], fields = prisms(c3_path)
        log("urban block")
        extra["urban_block_128x128x64"] = urban(max(3, args.steps // 2))
        log("lid cavity")
        extra["lid_cavity_48"] = lid_cavity()
        log("served mix")
        extra["served"] = served()
        log("published fixtures")
        extra["published_fixtures"] = published_fixtures()
        if c3_job is not None:
            log("waiting for the oracle's prism-city field")
            c3_job.join(timeout=900)
            if os.path.exists(c3_path + ".json"):
                xo = np.load(c3_path)
                with open(c3_path + ".json") as f:
                    meta = json.load(f)
                errs = {k: float(np.linalg.norm(v - xo) / np.linalg.norm(xo)) for k, v in fields.items()}
                extra["prisms_256x256x128"]["field_rel_error_vs_oracle_1e-10"] = errs
                extra["prisms_256x256x128"]["oracle"] = meta
        if campaign_procs:
            log("waiting for the oracle's campaign fields")
            campaign_oracle_errors(campaign_jobs, campaign_procs, 900)
    lat_ms = [1e3 * v for v in h["lat"]]
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * solve_s, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded street-canyon city, log-law velocity field, chained perturbed right-hand sides)",
        "config": {
            "workload": "street-canyon wind campaign, one direction (200 deg): 384x384x160 cells h=1 m, 400 convex "
                        "4-8-vertex prisms, x/y open, plain FP64 PCG + red-black multigrid V(2,2) on the split "
                        "aggregation hierarchy from x=0, rel-res 1e-8, a distinct chained right-hand side per solve",
            "cells": h["cells"],
            "solid_fraction": round(h["solid_fraction"], 4),
            "iterations_mean": round(float(np.mean(h["its"])), 2),
            "iterations_range": [min(h["its"]), max(h["its"])],
            "latency_ms": {"p50": round(pct(lat_ms, 0.5), 3), "p99": round(pct(lat_ms, 0.99), 3)},
        },
        "roofline": roofline,
        "cpu_baseline": base,
        "extra": extra,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
