"""The phase profile of pressure work (profiling.rs, multigrid.rs:74-130) and its
consistency rules (tests/profile_rules.py, after validation/check_profile.py).

CPU: the collector's nesting arithmetic; the oracle, which counts MultigridWork
where the reference's cycles increment it (multigrid.rs:672-725, masked.rs:301-341),
yields profiles that keep the rules, and the reference's own corruptions of them
(validation/test_profiling.py:46-185) break the rule they should.

GPU: the device library's work counters equal the oracle's for the reference cycle
(the same hierarchy and iteration count), its profiles pass the checker for every
cycle, and the setup and reuse rules of tests/multigrid.rs:175-208, 297-331 hold.
"""
import copy
import time

import numpy as np
import pytest

import oracle
from cfdgpu import profiling
from cfdgpu.case import IC0, MULTIGRID, OPEN, PERIODIC, WALL, Case

from profile_rules import broken, check
from test_oracle_goldens import aniso_case, aniso_rhs

ANISO = {"x": "wall", "cells": [12, 4, 12], "size_m": [1.0, 1.0, 0.1]}
WORK = profiling.MULTIGRID_FIELDS + profiling.BLOCK_FIELDS


def lid(n, bnd, pc=MULTIGRID):
    """tests/multigrid.rs:12-27: lid-cavity.json, x boundary replaced, initial velocity [1,0,0]."""
    return Case(cells=list(n), size_m=[1.0, 1.0, 1.0], boundaries=[bnd, PERIODIC, WALL], velocity_m_s=[1.0, 0.0, 0.0],
                preconditioner=pc)


def oracle_profile(case, solves=2, tol=1e-12):
    """A profile of ``solves`` oracle solves of one grid, each in a pressure_solve span;
    the hierarchy dimensions come from the oracle's MultigridStats."""
    o = oracle.Grid(case)
    c = profiling.Collector()
    mg = case.preconditioner == MULTIGRID
    for k in range(solves):
        i = np.arange(o.count, dtype=float)
        ex = o.center_nullspaces(np.sin(0.37 * i + k))
        b = o.apply(ex)
        with c.span("pressure_solve") as s:
            began = time.perf_counter()
            x, info = o.solve(b, tol, 2000)
            if mg:
                stats = o.mg_stats()  # after the solve: the first one builds the hierarchy
                info = dict(info)
                info.update(hierarchy_levels=len(stats["level_cells"]), hierarchy_cells=sum(stats["level_cells"]),
                            hierarchy_nonzeros=sum(stats["level_nonzeros"]),
                            hierarchy_bytes=12 * sum(stats["level_nonzeros"]),
                            hierarchy_build_seconds=1e-4)
            profiling.record_solve(s, info, mg, began)
    return c.snapshot()


# ------------------------------------------------------------------ collector
def test_nested_spans_account_exclusive_time_once():
    c = profiling.Collector()
    with c.span("pressure_projection"):
        time.sleep(0.002)
        with c.span("pressure_solve") as s:
            time.sleep(0.003)
            s.work_units = 4
            s.operator_applications = 5
            s.true_residual_checks = 1
    p = c.snapshot()
    outer, inner = p["phases"]["pressure_projection"], p["phases"]["pressure_solve"]
    assert outer["inclusive_s"] >= inner["inclusive_s"] + outer["exclusive_s"] - 1e-12
    assert abs(outer["inclusive_s"] - outer["exclusive_s"] - inner["inclusive_s"]) < 1e-12
    assert abs(p["accounted_exclusive_s"] - outer["inclusive_s"]) < 1e-12
    assert abs(p["unattributed_s"] + p["accounted_exclusive_s"] - p["elapsed_s"]) < 1e-12
    assert inner["work_units"] == 4 and "multigrid" not in inner
    check(p, "ic0")


def test_setup_reported_from_before_a_span_extends_it():
    """A build that ran before the solve span began (grid creation) is counted inside
    the span and, when it ran before the collector existed, inside elapsed."""
    began = time.perf_counter() - 0.5
    c = profiling.Collector()
    with c.span("pressure_solve") as s:
        s.nested("pressure_multigrid_setup", 0.5, began)
    p = c.snapshot()
    solve, setup = p["phases"]["pressure_solve"], p["phases"]["pressure_multigrid_setup"]
    assert setup["calls"] == 1 and setup["inclusive_s"] == 0.5
    assert solve["inclusive_s"] >= 0.5 and abs(solve["inclusive_s"] - solve["exclusive_s"] - 0.5) < 1e-12
    assert p["elapsed_s"] >= 0.5 and p["unattributed_s"] >= 0.0


def test_bad_nesting_and_open_spans_are_coded_errors():
    c = profiling.Collector()
    a = c.span("pressure_projection").__enter__()
    b = c.span("pressure_solve").__enter__()
    with pytest.raises(ValueError, match="CFD-B003 profile.active"):
        c.snapshot()
    a.__exit__(None, None, None)
    b.__exit__(None, None, None)
    with pytest.raises(ValueError, match="CFD-B003 profile.nesting"):
        c.snapshot()


def test_merge_keeps_hierarchy_maxima_and_sums_work():
    a = {f: 1 for f in WORK}
    profiling.merge_multigrid(a, {f: 3 for f in WORK})
    for f in WORK:
        assert a[f] == (3 if f in ("hierarchy_levels", "hierarchy_cells", "hierarchy_nonzeros", "hierarchy_bytes") else 4)


# ------------------------------------------------------------------ the checker on oracle work
def test_oracle_point_work_passes_and_corruptions_fail():
    """validation/test_profiling.py:46-87 on the 8x2x8 lid cavity with multigrid."""
    p = oracle_profile(lid([8, 2, 8], WALL))
    check(p, "multigrid")
    work = p["phases"]["pressure_solve"]["multigrid"]
    assert work["hierarchy_builds"] == 1 and work["cycles"] > 0
    assert not any(f in work for f in profiling.BLOCK_FIELDS)
    for field, rule in (("cycles", "mg_cycles"), ("hierarchy_builds", "mg_setup"),
                        ("smoothing_cell_updates", "mg_updates"), ("coarse_solves", "mg_cycles"),
                        ("restriction_cell_updates", "mg_updates")):
        changed = copy.deepcopy(p)
        changed["phases"]["pressure_solve"]["multigrid"][field] += 1
        assert broken(changed, "multigrid") == {rule}, field
    changed = copy.deepcopy(p)
    del changed["phases"]["pressure_solve"]["multigrid"]
    assert broken(changed, "multigrid") == {"choice"}
    assert broken(p, "ic0") == {"choice"}
    with pytest.raises(AssertionError, match="choice"):
        check(p, "ic0")


def test_oracle_block_work_passes_and_corruptions_fail():
    """validation/test_profiling.py:152-184 and tests/multigrid.rs:297-331 on the
    12x4x12 cavity of size [1, 1, 0.1]: every smoothed level takes line blocks."""
    p = oracle_profile(aniso_case(ANISO, "mg"), tol=1e-14)
    check(p, "multigrid")
    work = p["phases"]["pressure_solve"]["multigrid"]
    assert work["block_solves"] > 0
    assert work["block_smoothing_cell_updates"] == 2 * work["block_level_cell_visits"]
    assert work["smoothing_cell_updates"] == work["block_smoothing_cell_updates"]
    for field, value, rule in (("block_solves", 0, "mg_blocks"), ("block_smoothing_cell_updates", True, "mg_fields"),
                               ("block_smoothing_cell_updates", work["smoothing_cell_updates"] + 1, "mg_blocks")):
        changed = copy.deepcopy(p)
        changed["phases"]["pressure_solve"]["multigrid"][field] = value
        assert rule in broken(changed, "multigrid"), (field, value)
    changed = copy.deepcopy(p)
    del changed["phases"]["pressure_solve"]["multigrid"]["block_solves"]
    assert broken(changed, "multigrid") == {"mg_fields"}


def test_oracle_ic0_work_and_operator_count():
    p = oracle_profile(lid([9, 2, 7], OPEN, IC0))
    check(p, "ic0")
    changed = copy.deepcopy(p)
    changed["phases"]["pressure_solve"]["operator_applications"] += 1
    assert broken(changed, "ic0") == {"operators"}
    changed = copy.deepcopy(p)
    changed["phases"]["pressure_solve"]["exclusive_s"] = changed["phases"]["pressure_solve"]["inclusive_s"] * 2 + 1
    assert broken(changed, "ic0") == {"format"}
    changed = copy.deepcopy(p)
    changed["unattributed_s"] += 1.0
    assert broken(changed, "ic0") == {"time"}
    assert broken(p, "multigrid") == {"choice"}


def test_oracle_work_follows_the_hierarchy():
    """tests/multigrid.rs:199-206: smoothing = 4 x smoothed cells x cycles, residual =
    restriction = prolongation = smoothed cells x cycles (point smoothers)."""
    o = oracle.Grid(lid([9, 2, 7], OPEN))
    i = np.arange(o.count, dtype=float)
    b = o.apply(o.center_nullspaces(np.sin(0.37 * i)))
    x, info = o.solve(b, 1e-12, 2000)
    smoothed = sum(o.mg_stats()["level_cells"][:-1])
    assert info["smoothing_cell_updates"] == 4 * smoothed * info["cycles"]
    for f in ("residual_cell_updates", "restriction_cell_updates", "prolongation_cell_updates"):
        assert info[f] == smoothed * info["cycles"]
    assert info["cycles"] == info["coarse_solves"] == info["preconditioner_applications"]


# ------------------------------------------------------------------ device
def _device(case, cycle=None):
    import cfdgpu

    g = cfdgpu.Grid(case)
    if cycle is not None:
        g.set_options(cycle=cycle)
    return g


@pytest.mark.gpu
@pytest.mark.parametrize("spec_name", ["lid_9x2x7_open", "aniso_12x4x12"])
def test_device_reference_cycle_work_equals_oracle(gpu_available, spec_name):
    import cfdgpu

    case = lid([9, 2, 7], OPEN) if spec_name == "lid_9x2x7_open" else aniso_case(ANISO, "mg")
    tol = 1e-12 if spec_name == "lid_9x2x7_open" else 1e-14
    print(f"[profile] reference-cycle work {spec_name}", flush=True)
    o = oracle.Grid(case)
    ex, b = aniso_rhs({"x": "open" if case.boundaries[0] == OPEN else "wall"}, o.count, o.apply)
    _, want = o.solve(b, tol, 1200)
    g = _device(case, cfdgpu.CYCLE_REFERENCE)
    ctl = cfdgpu.Control(profile=True)
    _, got = g.solve(b, tol, 1200, control=ctl)
    assert got["iterations"] == want["iterations"]
    for f in ("cycles", "smoothing_cell_updates", "residual_cell_updates", "restriction_cell_updates",
              "prolongation_cell_updates", "coarse_solves", "block_smoothing_cell_updates", "block_solves",
              "block_level_cell_visits", "hierarchy_builds", "preconditioner_applications", "operator_applications",
              "true_residual_checks"):
        assert got[f] == want[f], f
    stats = o.mg_stats()
    assert got["hierarchy_levels"] == len(stats["level_cells"])
    assert got["hierarchy_cells"] == sum(stats["level_cells"])
    assert got["hierarchy_nonzeros"] == sum(stats["level_nonzeros"])
    shape = o.mg_shape()
    dev = g.multigrid_stats()
    assert dev.get("level_dimensions", []) == shape["level_dimensions"]
    assert [{k: v for k, v in b.items() if k != "allocated_bytes"} for b in dev.get("block_smoothers", [])] == \
        shape["block_smoothers"]
    assert dev["allocated_bytes"] == got["hierarchy_bytes"] > 0
    check(ctl.profile(), "multigrid")


@pytest.mark.gpu
@pytest.mark.parametrize("cycle", [0, 1], ids=["rb", "reference"])
def test_device_profile_reuse_and_projection(gpu_available, cycle):
    """tests/multigrid.rs:175-208 on an unchanged grid: one build over several solves,
    setup calls equal builds, work follows the hierarchy; a projection nests its solve."""
    import cfdgpu

    print(f"[profile] reuse cycle={cycle}", flush=True)
    case = lid([9, 2, 7], OPEN)
    ctl = cfdgpu.Control(profile=True)
    g = cfdgpu.Grid(case)
    g.set_options(cycle=cycle)
    i = np.arange(g.count, dtype=float)
    exact = np.sin(0.37 * i)
    with ctl.span("solver_run"):
        for _ in range(3):
            x, info = g.solve(g.apply(exact), 1e-12, 2000, control=ctl)
            assert info["relative_residual"] < 1.1e-11
            assert np.abs(x - exact).max() < 1e-9
        u = np.cos(0.21 * np.arange(g.face_count))
        g.project(u, np.zeros(g.count), 0.125, 1e-12, 2000, control=ctl)
    p = ctl.profile()
    check(p, "multigrid", roots=["solver_run"])
    assert p["phases"]["pressure_multigrid_setup"]["calls"] == 1
    assert p["phases"]["pressure_projection"]["calls"] == 1
    solve = p["phases"]["pressure_solve"]
    w = solve["multigrid"]
    assert solve["calls"] == 4 and w["hierarchy_builds"] == 1
    assert w["cycles"] == solve["preconditioner_applications"] == w["coarse_solves"]
    smoothed = sum(g.multigrid_stats()["level_cells"][:-1])
    assert w["smoothing_cell_updates"] == 4 * smoothed * w["cycles"]
    assert w["residual_cell_updates"] == w["restriction_cell_updates"] == smoothed * w["cycles"]
    assert w["hierarchy_bytes"] > 0


@pytest.mark.gpu
def test_device_ic0_profile_has_no_multigrid_work(gpu_available):
    import cfdgpu

    print("[profile] ic0", flush=True)
    case = lid([9, 2, 7], OPEN, IC0)
    g = cfdgpu.Grid(case)
    ctl = cfdgpu.Control(profile=True)
    i = np.arange(g.count, dtype=float)
    g.solve(g.apply(np.sin(0.37 * i)), 1e-12, 2000, control=ctl)
    p = ctl.profile()
    check(p, "ic0")
    assert "multigrid" not in p["phases"]["pressure_solve"]
    assert "pressure_multigrid_setup" not in p["phases"]
