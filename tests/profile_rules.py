"""Consistency rules for a phase profile of pressure work (test helper).

Written from the rules the reference's profile validation applies to pressure
work (validation/check_profile.py:127-180 timing, 213-323 multigrid work,
387-393 operator count); the profile format itself is the one of
crates/cfd-craft/src/profiling.rs as produced by cfdgpu.profiling.Collector.

``problems(profile, selected)`` returns every broken rule as (rule, detail)
pairs instead of stopping at the first; ``check(...)`` raises an
AssertionError naming them. Rule names:

  format       schema, phase rows, number types and ranges
  time         exclusive times add up to the outermost phases and to elapsed
  operators    operator applications = iterations + true-residual checks
  choice       multigrid work exists exactly when the case selects multigrid
  mg_fields    the multigrid work record has its fields (block counters all or none)
  mg_cycles    cycles = preconditioner applications = coarse solves
  mg_setup     hierarchy builds = multigrid set-up spans, no IC(0) set-up
  mg_updates   residual = restriction = prolongation updates; smoothing follows
               the fixed V(2,2) shape (two sweeps a side, one block sweep a side
               on block-smoothed levels)
  mg_blocks    block work is consistent and bounded by the smoothing work
  mg_size      the hierarchy has levels, cells, entries and storage
"""
import math

from cfdgpu.profiling import BLOCK_FIELDS, COUNTERS, MULTIGRID_FIELDS, SCHEMA


def _count(v):
    return type(v) is int and v >= 0


def _seconds(v):
    return type(v) in (int, float) and math.isfinite(v) and v >= 0


def problems(profile, selected, stopped=False, roots=None):
    """Broken rules of `profile` for a case whose pressure_preconditioner is
    `selected` ("ic0" or "multigrid"); `stopped`: the run ended early
    (cancelled, wall-time limit, diverged); `roots`: the outermost phases (by
    default those of the pressure path that are present)."""
    out = []
    if profile.get("schema_version") != SCHEMA:
        return [("format", f"schema {profile.get('schema_version')!r} is not {SCHEMA!r}")]
    phases = profile.get("phases")
    if not isinstance(phases, dict):
        return [("format", "phases is not a mapping")]
    for name, row in phases.items():
        bad = [f for f in ("calls",) + COUNTERS if not _count(row.get(f))]
        bad += [f for f in ("inclusive_s", "exclusive_s") if not _seconds(row.get(f))]
        if bad:
            out.append(("format", f"phase {name}: bad {', '.join(bad)}"))
        elif row["exclusive_s"] > row["inclusive_s"]:
            out.append(("format", f"phase {name}: more exclusive than inclusive time"))
    if out:
        return out

    # time: every second counted once
    elapsed = profile.get("elapsed_s")
    spare = profile.get("unattributed_s")
    if not (_seconds(elapsed) and _seconds(spare)):
        return [("format", "elapsed or unattributed time is not a finite nonnegative number")]
    slack = max(1e-9, 1e-9 * elapsed)
    exclusive = sum(r["exclusive_s"] for r in phases.values())
    if roots is None:
        roots = ["pressure_projection"] if "pressure_projection" in phases else ["pressure_solve"]
    outermost = sum(phases[r]["inclusive_s"] for r in roots if r in phases)
    if abs(exclusive - outermost) > slack:
        out.append(("time", f"exclusive times sum to {exclusive} s, the outermost phases span {outermost} s"))
    if abs(exclusive - profile.get("accounted_exclusive_s", -1.0)) > slack:
        out.append(("time", "accounted_exclusive_s is not the sum of the exclusive times"))
    if abs(exclusive + spare - elapsed) > slack:
        out.append(("time", "accounted plus unattributed time differs from elapsed"))

    solve = phases.get("pressure_solve", {})
    if solve.get("operator_applications", 0) != solve.get("work_units", 0) + solve.get("true_residual_checks", 0):
        out.append(("operators", "operator applications are not iterations plus true-residual checks"))

    work = solve.get("multigrid")
    calls = lambda name: phases.get(name, {}).get("calls", 0)  # noqa: E731
    if work is None:
        if selected == "multigrid" and calls("pressure_solve"):
            out.append(("choice", "multigrid is selected but the solves report no multigrid work"))
        return out
    if selected != "multigrid":
        out.append(("choice", f"multigrid work in a profile of a case that selects {selected}"))
    if not isinstance(work, dict):
        return out + [("mg_fields", "multigrid work is not a mapping")]
    present = set(work)
    missing = set(MULTIGRID_FIELDS) - present
    extra = present - set(MULTIGRID_FIELDS) - set(BLOCK_FIELDS)
    blocks = present & set(BLOCK_FIELDS)
    if missing or extra:
        out.append(("mg_fields", f"missing {sorted(missing)}, unknown {sorted(extra)}"))
    if blocks and blocks != set(BLOCK_FIELDS):
        out.append(("mg_fields", f"only some block counters: {sorted(blocks)}"))
    if any(not _count(v) for v in work.values()):
        out.append(("mg_fields", "a work counter is not a nonnegative integer"))
    if out and out[-1][0] == "mg_fields":
        return out

    cycles, updates = work["cycles"], work["smoothing_cell_updates"]
    residual = work["residual_cell_updates"]
    b_updates = work.get("block_smoothing_cell_updates", 0)
    b_solves = work.get("block_solves", 0)
    b_visits = work.get("block_level_cell_visits", 0)
    if not cycles == solve.get("preconditioner_applications", 0) == work["coarse_solves"]:
        out.append(("mg_cycles", f"{cycles} cycles, {solve.get('preconditioner_applications', 0)} preconditioner "
                                 f"applications, {work['coarse_solves']} coarse solves"))
    if work["hierarchy_builds"] != calls("pressure_multigrid_setup") or calls("pressure_factor_setup"):
        out.append(("mg_setup", f"{work['hierarchy_builds']} builds against {calls('pressure_multigrid_setup')} "
                                f"multigrid set-up spans and {calls('pressure_factor_setup')} IC(0) set-ups"))
    if not residual == work["restriction_cell_updates"] == work["prolongation_cell_updates"]:
        out.append(("mg_updates", "residual, restriction and prolongation updates differ"))
    if updates != 4 * residual - 2 * b_visits:
        out.append(("mg_updates", f"{updates} smoothing updates for {residual} residual updates and {b_visits} "
                                  "block-level visits"))
    if b_updates != 2 * b_visits or b_visits > residual:
        out.append(("mg_blocks", "block updates are not one sweep a side over the block-smoothed levels"))
    if not (b_updates == b_solves == 0 or 0 < 2 * b_solves <= b_updates <= updates):
        out.append(("mg_blocks", "block smoothing without solves, or more of it than smoothing overall"))
    sized = work["hierarchy_levels"] > 0 and work["hierarchy_nonzeros"] >= work["hierarchy_cells"] > 0 \
        and work["hierarchy_bytes"] > 0
    if not sized and not (stopped and cycles == 0):
        out.append(("mg_size", "the multigrid work names no hierarchy"))
    return out


def check(profile, selected, stopped=False, roots=None):
    found = problems(profile, selected, stopped, roots)
    assert not found, "profile rules broken: " + "; ".join(f"{r}: {d}" for r, d in found)


def broken(profile, selected, stopped=False, roots=None):
    """The names of the broken rules."""
    return {r for r, _ in problems(profile, selected, stopped, roots)}
