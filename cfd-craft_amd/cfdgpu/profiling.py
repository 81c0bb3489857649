"""Phase profile of the pressure path: the reference's ``cfd-craft.phase-profile/0.1``.

Mirrors crates/cfd-craft/src/profiling.rs: a :class:`Collector` hands out nested
spans (profiling.rs:170-191), each finished span adds its wall time and work to its
phase's ``PhaseTiming`` (profiling.rs:112-135, 240-281), and :meth:`Collector.snapshot`
returns the ``Profile`` (profiling.rs:140-153, 192-210) as a JSON-ready ``dict``.

The pressure solve records the reference's work (pressure.rs:350-410): attempted PCG
iterations as ``work_units``, operator products, preconditioner applications and
true-residual checks, and with the multigrid preconditioner the ``MultigridWork``
counters (multigrid.rs:74-130) that the device library counts per solve.

The device library builds a hierarchy when the grid is created or a cycle is first
selected, not lazily inside the first solve as the reference does (pressure.rs:71-93).
The first solve after a build reports it (``hierarchy_builds`` = 1) and its span
receives a nested ``pressure_multigrid_setup`` phase with the build's measured host
seconds, so the setup calls still equal the reported builds.
"""
from __future__ import annotations

import threading
import time

SCHEMA = "cfd-craft.phase-profile/0.1"

# MultigridWork fields (multigrid.rs:78-114); the block counters are omitted when zero
MULTIGRID_FIELDS = (
    "hierarchy_builds",
    "hierarchy_levels",
    "hierarchy_cells",
    "hierarchy_nonzeros",
    "hierarchy_bytes",
    "cycles",
    "smoothing_cell_updates",
    "residual_cell_updates",
    "restriction_cell_updates",
    "prolongation_cell_updates",
    "coarse_solves",
)
BLOCK_FIELDS = ("block_smoothing_cell_updates", "block_solves", "block_level_cell_visits")
# hierarchy dimensions merge as maxima, everything else as sums (multigrid.rs:115-130)
_MAXIMA = ("hierarchy_levels", "hierarchy_cells", "hierarchy_nonzeros", "hierarchy_bytes")

COUNTERS = ("work_units", "operator_applications", "preconditioner_applications", "true_residual_checks")


def merge_multigrid(into: dict, work: dict) -> None:
    """MultigridWork::merge (multigrid.rs:115-130)."""
    for f in MULTIGRID_FIELDS + BLOCK_FIELDS:
        a, b = into.get(f, 0), work.get(f, 0)
        into[f] = max(a, b) if f in _MAXIMA else a + b


def multigrid_work(info: dict) -> dict:
    """The MultigridWork of one solve from the library's linear info."""
    return {f: int(info[f]) for f in MULTIGRID_FIELDS + BLOCK_FIELDS}


class Span:
    """A scoped phase timer (profiling.rs:213-281). Work fields are set by the code
    that runs inside the span and added to the phase when the span ends."""

    def __init__(self, collector: "Collector | None", phase: str):
        self._collector = collector
        self.phase = phase
        self.work_units = 0
        self.operator_applications = 0
        self.preconditioner_applications = 0
        self.true_residual_checks = 0
        self.multigrid = None
        self._frame = None

    def __enter__(self):
        if self._collector is not None:
            self._frame = self._collector._push()
        return self

    def __exit__(self, *exc):
        if self._collector is not None:
            self._collector._pop(self)
        return False

    def nested(self, phase: str, seconds: float, began: float) -> None:
        """Record a finished child phase that ran at ``began`` (a perf_counter value),
        possibly before this span started, for ``seconds``."""
        if self._collector is not None:
            self._collector._nested(phase, seconds, began)


class _Frame:
    __slots__ = ("id", "began", "children", "extra")

    def __init__(self, ident: int, began: float):
        self.id = ident
        self.began = began
        self.children = 0.0
        self.extra = 0.0  # time of children that ran before this frame began


class Collector:
    """Profile collector (profiling.rs:160-211). ``span(phase)`` is a context manager;
    spans of one thread must finish in nested order."""

    def __init__(self):
        self._lock = threading.Lock()
        self._t0 = time.perf_counter()
        self._extra = 0.0
        self._phases: dict[str, dict] = {}
        self._stacks: dict[int, list[_Frame]] = {}
        self._next = 0
        self._invalid = False

    def span(self, phase: str) -> Span:
        return Span(self, phase)

    def _push(self) -> _Frame:
        with self._lock:
            f = _Frame(self._next, time.perf_counter())
            self._next += 1
            self._stacks.setdefault(threading.get_ident(), []).append(f)
            return f

    def _timing(self, phase: str) -> dict:
        return self._phases.setdefault(
            phase, dict(calls=0, inclusive_s=0.0, exclusive_s=0.0, **{c: 0 for c in COUNTERS})
        )

    def _nested(self, phase: str, seconds: float, began: float) -> None:
        with self._lock:
            stack = self._stacks.get(threading.get_ident(), [])
            t = self._timing(phase)
            t["calls"] += 1
            t["inclusive_s"] += seconds
            t["exclusive_s"] += seconds
            if stack:
                stack[-1].children += seconds
            # time that ran before an open span began extends that span
            for f in stack:
                if began < f.began:
                    f.extra += seconds
            if began < self._t0:
                self._extra += seconds

    def _pop(self, span: Span) -> None:
        now = time.perf_counter()
        with self._lock:
            stack = self._stacks.get(threading.get_ident(), [])
            frame = span._frame
            if not stack or stack[-1] is not frame:
                self._invalid = True
                if frame in stack:
                    stack.remove(frame)
                return
            stack.pop()
            elapsed = now - frame.began + frame.extra
            if stack:
                stack[-1].children += elapsed
            else:
                self._stacks.pop(threading.get_ident(), None)
            t = self._timing(span.phase)
            t["calls"] += 1
            t["inclusive_s"] += elapsed
            t["exclusive_s"] += max(0.0, elapsed - frame.children)
            for c in COUNTERS:
                t[c] += getattr(span, c)
            if span.multigrid is not None:
                merge_multigrid(t.setdefault("multigrid", {}), span.multigrid)

    def snapshot(self) -> dict:
        """Profile (profiling.rs:192-210): phases keyed by name in sorted order, block
        counters omitted when zero. Raises ``ValueError`` for unfinished or badly
        nested spans, as the reference's coded CFD-B003 errors."""
        with self._lock:
            if self._invalid:
                raise ValueError("CFD-B003 profile.nesting: spans must finish in nested order")
            if any(self._stacks.values()):
                raise ValueError("CFD-B003 profile.active: finish all phase spans before snapshot")
            elapsed = time.perf_counter() - self._t0 + self._extra
            phases = {}
            for name in sorted(self._phases):
                row = {k: v for k, v in self._phases[name].items() if k != "multigrid"}
                if "multigrid" in self._phases[name]:
                    mg = dict(self._phases[name]["multigrid"])
                    for f in BLOCK_FIELDS:
                        if mg.get(f, 0) == 0:
                            mg.pop(f, None)
                    row["multigrid"] = mg
                phases[name] = row
        accounted = sum(p["exclusive_s"] for p in phases.values())
        return dict(
            schema_version=SCHEMA,
            elapsed_s=elapsed,
            accounted_exclusive_s=accounted,
            unattributed_s=elapsed - accounted,
            phases=phases,
        )


def record_solve(span: Span, info: dict, multigrid: bool, build_began: float | None) -> None:
    """Set a pressure_solve span's work from the library's linear info
    (pressure.rs:350-410, 71-93)."""
    span.work_units += int(info["iterations"])
    span.operator_applications += int(info["operator_applications"])
    span.preconditioner_applications += int(info["preconditioner_applications"])
    span.true_residual_checks += int(info["true_residual_checks"])
    if not multigrid:
        return
    work = multigrid_work(info)
    if span.multigrid is None:
        span.multigrid = work
    else:
        merge_multigrid(span.multigrid, work)
    if work["hierarchy_builds"]:
        began = build_began if build_began is not None else time.perf_counter()
        span.nested("pressure_multigrid_setup", float(info["hierarchy_build_seconds"]), began)
