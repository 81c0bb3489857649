"""Python host of the cfdgpu C-ABI: the reference's pressure API on the MI355X.

Mirrors crates/cfd-craft/src/pressure.rs and the parts of grid.rs it reads:

=============================  =====================================================
reference                      here
=============================  =====================================================
``Grid::build`` (grid.rs:246)  ``Grid(case)``
``pressure::apply`` (:293)     ``Grid.apply(x)``
``center_nullspaces`` (:316)   ``Grid.center_nullspaces(x)``
``solve`` (:329)               ``Grid.solve(b, tol, max_iterations, control)``
``solve_scaled`` (:342)        ``Grid.solve_scaled(b, scale, tol, max_iterations, control)``
``project`` (:469)             ``Grid.project(u, p, dt, tol, max_iterations, control)``
``multigrid_stats`` (:100)     ``Grid.multigrid_stats()``
``LinearInfo`` (:29-36)        the ``dict`` returned beside the solution
``Control`` (solver.rs:75)     ``Control(deadline_s, cancel=callable, profile=True)``
``Control::profile``            ``Control.profile()`` (phase profile, ``profiling.py``)
coded ``Err(String)``          ``CfdError`` (``.code`` 31..36, ``"cancelled"``, ...)
=============================  =====================================================

All arithmetic runs in hand-written HIP kernels (libcfdgpu.so, built for gfx950).
There is no CPU fallback: importing this module without the built library raises.
"""
from __future__ import annotations

import ctypes
import os
import time

import numpy as np

from . import profiling
from .case import (  # noqa: F401  (re-exported)
    BLOCKED,
    FLOW_NEUTRAL_RANS,
    FLOW_VECTOR,
    GROUND_OPEN_TOP,
    IC0,
    INLET,
    INTERIOR,
    MULTIGRID,
    OPEN,
    OUTLET,
    PERIODIC,
    SLIP,
    WALL,
    Block,
    Case,
)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(_HERE, "libcfdgpu.so")

PRECISION_FP64, PRECISION_MIXED = 0, 1
CYCLE_RB, CYCLE_REFERENCE, CYCLE_ZEBRA = 0, 1, 3
CANCELLED, WALL_TIME = 100, 101


class CfdError(RuntimeError):
    """A coded failure of the pressure solve (CFD-E031..E036, cancelled, wall_time_limit)."""

    def __init__(self, code: int, text: str):
        super().__init__(text)
        self.code = code


class LinearInfo(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int64),
        ("relative_residual", ctypes.c_double),
        ("recurrent_residual", ctypes.c_double),
        ("operator_applications", ctypes.c_int64),
        ("preconditioner_applications", ctypes.c_int64),
        ("true_residual_checks", ctypes.c_int64),
        ("cycles", ctypes.c_int64),
        ("smoothing_cell_updates", ctypes.c_int64),
        ("residual_cell_updates", ctypes.c_int64),
        ("coarse_solves", ctypes.c_int64),
        ("hierarchy_builds", ctypes.c_int64),
        ("seconds", ctypes.c_double),
        ("restriction_cell_updates", ctypes.c_int64),
        ("prolongation_cell_updates", ctypes.c_int64),
        ("block_smoothing_cell_updates", ctypes.c_int64),
        ("block_solves", ctypes.c_int64),
        ("block_level_cell_visits", ctypes.c_int64),
        ("hierarchy_levels", ctypes.c_int64),
        ("hierarchy_cells", ctypes.c_int64),
        ("hierarchy_nonzeros", ctypes.c_int64),
        ("hierarchy_bytes", ctypes.c_int64),
        ("hierarchy_build_seconds", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BlockSmootherStats(ctypes.Structure):
    _fields_ = [
        ("level", ctypes.c_int64),
        ("n_axes", ctypes.c_int64),
        ("axes", ctypes.c_int64 * 3),
        ("block_unknowns", ctypes.c_int64),
        ("block_count", ctypes.c_int64),
        ("half_bandwidth", ctypes.c_int64),
        ("allocated_bytes", ctypes.c_int64),
    ]


class MgStats(ctypes.Structure):
    _fields_ = [
        ("levels", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("level_cells", ctypes.c_int64 * 32),
        ("level_nonzeros", ctypes.c_int64 * 32),
        ("level_dims", (ctypes.c_int64 * 3) * 32),
        ("allocated_bytes", ctypes.c_int64),
        ("n_block_smoothers", ctypes.c_int64),
        ("block_smoothers", BlockSmootherStats * 32),
        ("digest", ctypes.c_uint64),
    ]

    def as_dict(self):
        """MultigridStats (multigrid.rs:55-72) as the reference serializes it: the
        dimensions and block smoothers only where the hierarchy has them."""
        n = self.levels
        d = dict(level_cells=list(self.level_cells[:n]), level_nonzeros=list(self.level_nonzeros[:n]),
                 allocated_bytes=int(self.allocated_bytes))
        dims = [list(self.level_dims[i]) for i in range(n)]
        if any(any(v) for v in dims):
            d["level_dimensions"] = dims
        blocks = []
        for k in range(self.n_block_smoothers):
            b = self.block_smoothers[k]
            blocks.append(dict(level=int(b.level), axes=list(b.axes[: b.n_axes]), block_unknowns=int(b.block_unknowns),
                               block_count=int(b.block_count), half_bandwidth=int(b.half_bandwidth),
                               allocated_bytes=int(b.allocated_bytes)))
        if blocks:
            d["block_smoothers"] = blocks
        return d


class KernelProbe(ctypes.Structure):
    _fields_ = [
        ("launches", ctypes.c_int64),
        ("seconds", ctypes.c_double),
        ("bytes", ctypes.c_double),
        ("dropped", ctypes.c_int64),
        ("truncated", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


# kernel classes of Grid.profile_solve (include/cfdgpu.h CFD_PROBE_*)
PROBE_KERNELS = {
    "fine_half_sweep": 1, "coarse_half_sweep": 2, "operator": 3, "axpy": 4, "fine_restrict": 5,
    "fine_rz_half_sweep": 6, "fine_first_half_sweep": 7, "coarse_restrict": 8, "coarsest_solve": 9,
    "fine_prolong": 10, "ic0_apply": 11, "true_residual": 12,
}


class CFace(ctypes.Structure):
    """cfd_face: one entry of Grid::faces (grid.rs:85-108)."""
    _fields_ = [("axis", ctypes.c_int32), ("kind", ctypes.c_int32), ("lo", ctypes.c_int64), ("hi", ctypes.c_int64),
                ("area", ctypes.c_double), ("distance", ctypes.c_double)]


class GridDesc(ctypes.Structure):
    """cfd_grid_desc: the built Grid as the pressure path reads it (grid.rs:114-141)."""
    _fields_ = [("n", ctypes.c_int64 * 3), ("dx", ctypes.c_double * 3), ("volume", ctypes.c_double),
                ("periodic", ctypes.c_int32 * 3), ("fluid", ctypes.c_void_p), ("face_count", ctypes.c_int64),
                ("faces", ctypes.c_void_p), ("component", ctypes.c_void_p), ("component_count", ctypes.c_int64),
                ("anchored", ctypes.c_void_p), ("preconditioner", ctypes.c_int32)]


FACE_INTERIOR, FACE_WALL, FACE_SLIP, FACE_INLET, FACE_OUTLET, FACE_BLOCKED = range(6)
FACE_DTYPE = np.dtype([("axis", np.int32), ("kind", np.int32), ("lo", np.int64), ("hi", np.int64),
                       ("area", np.float64), ("distance", np.float64)], align=True)


class GridDescriptor:
    """A caller's built Grid for cfd_grid_create_desc: cell counts n, cell sizes dx,
    volume, periodic flags, the fluid mask (n0 n1 n2 flags) and the faces as a
    FACE_DTYPE record array in Grid::faces order; optionally component (-1 for
    solid) and anchored, which the library checks against the faces."""

    def __init__(self, n, dx, volume, periodic, fluid, faces, preconditioner=MULTIGRID, component=None,
                 anchored=None, tolerance=1e-10, max_iterations=1200):
        self.n = [int(v) for v in n]
        self.dx = [float(v) for v in dx]
        self.volume = float(volume)
        self.periodic = [1 if p else 0 for p in periodic]
        self.fluid = np.ascontiguousarray(fluid, dtype=np.uint8)
        self.faces = np.ascontiguousarray(faces, dtype=FACE_DTYPE)
        self.preconditioner = int(preconditioner)
        self.component = None if component is None else np.ascontiguousarray(component, dtype=np.int64)
        self.anchored = None if anchored is None else np.ascontiguousarray(anchored, dtype=np.uint8)
        self.tolerance = tolerance
        self.max_iterations = max_iterations
        if self.fluid.size != self.n[0] * self.n[1] * self.n[2]:
            raise ValueError("fluid: one flag per cell")

    def to_c(self) -> GridDesc:
        d = GridDesc()
        for a in range(3):
            d.n[a], d.dx[a], d.periodic[a] = self.n[a], self.dx[a], self.periodic[a]
        d.volume = self.volume
        d.fluid = self.fluid.ctypes.data
        d.face_count = self.faces.size
        d.faces = self.faces.ctypes.data
        if self.component is not None:
            if self.component.size != self.fluid.size:
                raise ValueError("component: one entry per cell")
            d.component = self.component.ctypes.data
            d.component_count = 0 if self.anchored is None else self.anchored.size
            if self.anchored is not None:
                d.anchored = self.anchored.ctypes.data
        d.preconditioner = self.preconditioner
        return d


class Options(ctypes.Structure):
    _fields_ = [
        ("precision", ctypes.c_int32),
        ("warm_start", ctypes.c_int32),
        ("poll_every", ctypes.c_int32),
        ("cycle", ctypes.c_int32),
    ]


POLL_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p)

_LIB = None


def lib():
    """The loaded libcfdgpu.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBRARY):
            raise ImportError(f"{LIBRARY} is not built; run `make -C cfd-craft_amd` (hipcc, gfx950)")
        L = ctypes.CDLL(LIBRARY)
        P = ctypes.c_void_p
        I = ctypes.c_int
        D = ctypes.c_double
        I64 = ctypes.c_int64
        sig = {
            "cfd_grid_create": (I, [P, ctypes.POINTER(ctypes.c_void_p)]),
            "cfd_grid_destroy": (None, [P]),
            "cfd_grid_matches": (I, [P, P, P]),
            "cfd_grid_create_desc": (I, [P, ctypes.POINTER(ctypes.c_void_p)]),
            "cfd_grid_matches_desc": (I, [P, P, P]),
            "cfd_grid_counts": (I, [P, P]),
            "cfd_grid_codes": (I, [P, P]),
            "cfd_grid_components": (I, [P, P, P]),
            "cfd_grid_set_options": (I, [P, P]),
            "cfd_ic0_diagonal": (I, [P, P]),
            "cfd_multigrid_stats": (I, [P, P]),
            "cfd_apply": (I, [P, P, P]),
            "cfd_center_nullspaces": (I, [P, P]),
            "cfd_solve_scaled": (I, [P, P, P, D, I64, POLL_FN, P, P, P]),
            "cfd_project": (I, [P, P, P, D, D, I64, POLL_FN, P, P]),
            "cfd_project_static": (I, [P, P, P, D, D, I64, POLL_FN, P, P]),
            "cfd_projection_rhs": (I, [P, P, D, P, P]),
            "cfd_precondition": (I, [P, P, P]),
            "cfd_apply_device": (I, [P, P, P]),
            "cfd_solve_scaled_device": (I, [P, P, P, D, I64, POLL_FN, P, P, P]),
            "cfd_precondition_device": (I, [P, P, P]),
            "cfd_project_device": (I, [P, P, P, D, D, I64, POLL_FN, P, P]),
            "cfd_device_alloc": (I, [ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]),
            "cfd_device_free": (I, [P]),
            "cfd_copy_to_device": (I, [P, P, ctypes.c_size_t]),
            "cfd_copy_to_host": (I, [P, P, ctypes.c_size_t]),
            "cfd_device_synchronize": (I, []),
            "cfd_last_error": (ctypes.c_char_p, []),
            "cfd_case_topology": (I, [P, P, P, P, P, I64, P]),
            "cfd_case_topology_device": (I, [P, P, P, P, P, I64]),
            "cfd_profile_kernel": (I, [P, I, I, P, P]),
            "cfd_profile_solve": (I, [P, P, P, D, I64, P, I, P, P]),
            "cfd_server_create": (I, [I, ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]),
            "cfd_server_destroy": (None, [P]),
            "cfd_server_set_batching": (I, [P, I]),
            "cfd_server_admit": (I, [P, P, ctypes.POINTER(ctypes.c_void_p)]),
            "cfd_server_admit_async": (I, [P, P, ctypes.POINTER(ctypes.c_uint64)]),
            "cfd_server_admitted": (I, [P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_void_p)]),
            "cfd_server_release": (I, [P, P]),
            "cfd_server_usage": (I, [P, P, P]),
            "cfd_grid_footprint": (ctypes.c_size_t, [P]),
            "cfd_server_submit": (I, [P, P, P, P, D, I64, I, P, ctypes.POINTER(ctypes.c_uint64)]),
            "cfd_server_progress": (I, [P]),
            "cfd_server_wait": (I, [P, ctypes.c_uint64]),
            "cfd_server_finished": (I, [P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int)]),
            "cfd_grid_perturb": (I, [P, P, P, D, ctypes.c_uint64]),
            "cfd_server_drain": (I, [P]),
            "cfd_server_result": (I, [P, ctypes.c_uint64, P, P, P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _check(rc: int):
    if rc:
        raise CfdError(rc, lib().cfd_last_error().decode())


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _host_vector(a, n: int, name: str, copy: bool = False) -> np.ndarray:
    """A contiguous float64 copy or view of `a`, which must hold exactly n values:
    the library reads and writes n doubles through the pointer."""
    v = np.array(a, dtype=np.float64, copy=True) if copy else np.ascontiguousarray(a, dtype=np.float64)
    if v.ndim != 1 or v.size != n:
        raise ValueError(f"{name} must be a vector of {n} values, got shape {np.shape(a)}")
    return v


def _device_vector(d, n: int, name: str, optional: bool = False):
    """The device pointer of a DeviceArray of exactly n float64 values (the
    kernels index n doubles from it)."""
    if d is None:
        if optional:
            return None
        raise ValueError(f"{name} is required")
    if not isinstance(d, DeviceArray) or d.dtype != np.float64 or d.n != n:
        got = f"{getattr(d, 'n', '?')} x {getattr(d, 'dtype', type(d).__name__)}"
        raise ValueError(f"{name} must be a DeviceArray of {n} float64 values, got {got}")
    if d.ptr is None:
        raise ValueError(f"{name} has been freed")
    return d.ptr


def case_topology(case: Case, ic0: bool = False):
    """Host-only Grid::build topology (no device needed): dict with counts, codes,
    component per cell, anchored per component and optionally the IC(0) diagonal."""
    c, keep = case.to_c()
    counts = np.zeros(8, np.int64)
    _check(lib().cfd_case_topology(ctypes.byref(c), _ptr(counts), None, None, None, 0, None))
    n = int(counts[0])
    codes = np.zeros(n, np.uint8)
    comp = np.zeros(n, np.int64)
    anch = np.zeros(max(1, int(counts[3])), np.uint8)
    diag = np.zeros(n) if ic0 else None
    _check(lib().cfd_case_topology(ctypes.byref(c), _ptr(counts), _ptr(codes), _ptr(comp), _ptr(anch), anch.size,
                                   _ptr(diag) if diag is not None else None))
    return dict(counts=counts, codes=codes, component=comp, anchored=anch[: int(counts[3])].astype(bool), ic0_diagonal=diag)


def case_topology_device(case: Case):
    """The topology of case_topology built on the GPU (rasterization, face codes,
    union-find components), as a dict of the same fields."""
    c, keep = case.to_c()
    n = int(np.prod(case.cells))
    counts = np.zeros(8, np.int64)
    codes = np.zeros(n, np.uint8)
    comp = np.zeros(n, np.int64)
    anch = np.zeros(n, np.uint8)
    _check(lib().cfd_case_topology_device(ctypes.byref(c), _ptr(counts), _ptr(codes), _ptr(comp), _ptr(anch), anch.size))
    return dict(counts=counts, codes=codes, component=comp, anchored=anch[: int(counts[3])].astype(bool))


def case_multigrid_stats(case: Case, cycle: int = 0):
    """Host-only level sizes of the multigrid hierarchy a cycle builds (no device)."""
    c, keep = case.to_c()
    s = MgStats()
    _check(lib().cfd_case_multigrid_stats(ctypes.byref(c), int(cycle), ctypes.byref(s)))
    d = s.as_dict()
    del d["allocated_bytes"]  # nothing is allocated on a device here
    d["digest"] = int(s.digest)
    return d


class Control:
    """Control::check (solver.rs:122-139) as seen by the pressure solve: a wall-time
    deadline and an optional cancel callback, polled every 8 PCG iterations. With
    ``profile=True`` the pressure calls record phase spans (profiling.rs) that
    ``profile()`` returns as a ``cfd-craft.phase-profile/0.1`` dict."""

    def __init__(self, deadline_s: float = float("inf"), cancel=None, profile: bool = False):
        self.deadline = time.monotonic() + deadline_s
        self.cancel = cancel
        self.polls = 0
        self.collector = profiling.Collector() if profile else None

    def span(self, phase: str) -> profiling.Span:
        return profiling.Span(self.collector, phase)

    def profile(self):
        """The phase profile so far, or None without profiling (solver.rs Control::profile)."""
        return self.collector.snapshot() if self.collector is not None else None

    def check(self) -> int:
        self.polls += 1
        if self.cancel is not None and self.cancel():
            return 1
        if time.monotonic() > self.deadline:
            return 2
        return 0


class DeviceArray:
    """A float64 (or other dtype) buffer in HBM owned by the library's allocator."""

    def __init__(self, n: int, dtype=np.float64):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = max(1, self.n * self.dtype.itemsize)
        ptr = ctypes.c_void_p()
        _check(lib().cfd_device_alloc(self.nbytes, ctypes.byref(ptr)))
        self.ptr = ptr

    @classmethod
    def from_host(cls, a: np.ndarray):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        _check(lib().cfd_copy_to_device(d.ptr, _ptr(a), a.nbytes))
        return d

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        _check(lib().cfd_copy_to_device(self.ptr, _ptr(a), a.nbytes))

    def numpy(self) -> np.ndarray:
        out = np.empty(self.n, self.dtype)
        _check(lib().cfd_copy_to_host(_ptr(out), self.ptr, out.nbytes))
        return out

    def free(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value:
            lib().cfd_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Grid:
    """A device-resident pressure grid: topology, factor or hierarchy, work vectors."""

    def __init__(self, case: Case):
        self.case = case
        c, keep = case.to_c()
        h = ctypes.c_void_p()
        began = time.perf_counter()
        _check(lib().cfd_grid_create(ctypes.byref(c), ctypes.byref(h)))
        self._attach(h, case, began)

    @classmethod
    def from_descriptor(cls, desc: "GridDescriptor") -> "Grid":
        """The grid of a Grid the caller built (cfd_grid_create_desc)."""
        g = cls.__new__(cls)
        d = desc.to_c()
        h = ctypes.c_void_p()
        began = time.perf_counter()
        _check(lib().cfd_grid_create_desc(ctypes.byref(d), ctypes.byref(h)))
        g._attach(h, desc, began)
        return g

    def matches_descriptor(self, desc: "GridDescriptor") -> bool:
        """Signature::matches against a caller's Grid (cfd_grid_matches_desc)."""
        d = desc.to_c()
        same = ctypes.c_int(0)
        _check(lib().cfd_grid_matches_desc(self.h, ctypes.byref(d), ctypes.byref(same)))
        return bool(same.value)

    def _attach(self, h, case: Case, began: float):
        self.case = case
        self.h = h
        self._build_began = began  # when the hierarchy build reported next began
        self._cycles_seen = {int(os.environ.get("CFD_MG_CYCLE", "0"))}
        counts = np.zeros(8, np.int64)
        _check(lib().cfd_grid_counts(self.h, _ptr(counts)))
        self.count, self.face_count, self.fluid_count, self.n_components, self.n_anchored = (int(v) for v in counts[:5])
        self.tolerance = case.tolerance
        self.max_iterations = case.max_iterations

    def close(self):
        if getattr(self, "owner", None) is not None:
            return  # released by its server
        if getattr(self, "h", None) is not None and self.h.value:
            lib().cfd_grid_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- topology
    def codes(self) -> np.ndarray:
        out = np.zeros(self.count, np.uint8)
        _check(lib().cfd_grid_codes(self.h, _ptr(out)))
        return out

    def components(self):
        comp = np.zeros(self.count, np.int64)
        anch = np.zeros(max(1, self.n_components), np.uint8)
        _check(lib().cfd_grid_components(self.h, _ptr(comp), _ptr(anch)))
        return comp, anch[: self.n_components].astype(bool)

    def ic0_diagonal(self) -> np.ndarray:
        out = np.zeros(self.count)
        _check(lib().cfd_ic0_diagonal(self.h, _ptr(out)))
        return out

    def multigrid_stats(self) -> dict:
        s = MgStats()
        _check(lib().cfd_multigrid_stats(self.h, ctypes.byref(s)))
        return s.as_dict()

    def profile_kernel(self, kernel: int, reps: int = 20):
        """(average seconds per launch, algorithmic bytes per launch) of a hot kernel:
        0 operator, 1 fine smoother half-sweep, 2 x/r update, 3 IC(0) apply, 4 V-cycle."""
        sec = ctypes.c_double(0.0)
        nbytes = ctypes.c_double(0.0)
        _check(lib().cfd_profile_kernel(self.h, int(kernel), int(reps), ctypes.byref(sec), ctypes.byref(nbytes)))
        return sec.value, nbytes.value

    def profile_solve(self, d_b: "DeviceArray", d_x: "DeviceArray", tol: float, kernel, max_iterations: int = 3000,
                      d_scale: "DeviceArray" = None):
        """One solve with the launches of one kernel class (a name or number of
        PROBE_KERNELS) timed by HIP event pairs: dict with launches, seconds,
        bytes (algorithmic) and the solve's info."""
        pb = _device_vector(d_b, self.count, "d_b")
        px = _device_vector(d_x, self.count, "d_x")
        ps = _device_vector(d_scale, self.count, "d_scale", optional=True)
        k = PROBE_KERNELS[kernel] if isinstance(kernel, str) else int(kernel)
        out = KernelProbe()
        info = LinearInfo()
        _check(lib().cfd_profile_solve(self.h, pb, ps, float(tol), int(max_iterations), px, k, ctypes.byref(out),
                                       ctypes.byref(info)))
        return dict(launches=out.launches, seconds=out.seconds, bytes=out.bytes, dropped=out.dropped,
                    truncated=bool(out.truncated), info=info.as_dict())

    def set_options(self, precision: int = PRECISION_FP64, warm_start: bool = False, cycle: int = 0):
        """precision: PRECISION_FP64 or PRECISION_MIXED; cycle (multigrid): CYCLE_RB,
        CYCLE_REFERENCE (the reference's hierarchy and lexicographic smoothing, FP64),
        or CYCLE_ZEBRA (the reference's hierarchy, its line and plane
        blocks included, smoothed by colour classes; FP64)."""
        o = Options(int(precision), int(bool(warm_start)), 8, int(cycle))
        began = time.perf_counter()
        _check(lib().cfd_grid_set_options(self.h, ctypes.byref(o)))
        if self.case.preconditioner == MULTIGRID and int(cycle) not in self._cycles_seen:
            self._cycles_seen.add(int(cycle))
            self._build_began = began

    # ---- operator and gauges
    def apply(self, x) -> np.ndarray:
        x = _host_vector(x, self.count, "x")
        y = np.empty(self.count)
        _check(lib().cfd_apply(self.h, _ptr(x), _ptr(y)))
        return y

    def center_nullspaces(self, x) -> np.ndarray:
        x = _host_vector(x, self.count, "x", copy=True)
        _check(lib().cfd_center_nullspaces(self.h, _ptr(x)))
        return x

    def matches(self, case: Case) -> bool:
        """Signature::matches (multigrid.rs:158-222): whether `case` gives this
        grid's operator exactly (counts, spacing and face-weight bits, fluid mask,
        anchoring, interior and outlet faces, periodic axes, preconditioner)."""
        c, keep = case.to_c()
        same = ctypes.c_int(0)
        _check(lib().cfd_grid_matches(self.h, ctypes.byref(c), ctypes.byref(same)))
        return bool(same.value)

    def precondition(self, r) -> np.ndarray:
        r = _host_vector(r, self.count, "r")
        z = np.empty(self.count)
        _check(lib().cfd_precondition(self.h, _ptr(r), _ptr(z)))
        return z

    # ---- solves
    @staticmethod
    def _poll(control):
        if control is None:
            return POLL_FN(0), None
        fn = POLL_FN(lambda _user: control.check())
        return fn, fn

    def _measured(self, control, call):
        """Run ``call(info)`` inside a pressure_solve span of ``control`` and record the
        attempted work, also of a failed solve (pressure.rs:350-410)."""
        info = LinearInfo()
        span = control.span("pressure_solve") if control is not None else profiling.Span(None, "pressure_solve")
        with span:
            rc = call(info)
            d = info.as_dict()
            # the multigrid work exists once the solve passed its compatibility check
            mg = self.case.preconditioner == MULTIGRID and d["hierarchy_levels"] > 0
            profiling.record_solve(span, d, mg, self._build_began)
        _check(rc)
        return d

    def solve_scaled(self, b, scale=None, tol=None, max_iterations=None, control=None):
        tol = self.tolerance if tol is None else tol
        max_iterations = self.max_iterations if max_iterations is None else max_iterations
        b = _host_vector(b, self.count, "b")
        s = _host_vector(scale, self.count, "scale") if scale is not None else None
        x = np.empty(self.count)
        fn, keep = self._poll(control)
        info = self._measured(
            control,
            lambda info: lib().cfd_solve_scaled(
                self.h, _ptr(b), _ptr(s) if s is not None else None, float(tol), int(max_iterations), fn, None, _ptr(x),
                ctypes.byref(info),
            ),
        )
        return x, info

    def solve(self, b, tol=None, max_iterations=None, control=None):
        return self.solve_scaled(b, None, tol, max_iterations, control)

    def solve_device(self, d_b: DeviceArray, d_x: DeviceArray, tol=None, max_iterations=None, d_scale=None, control=None):
        tol = self.tolerance if tol is None else tol
        max_iterations = self.max_iterations if max_iterations is None else max_iterations
        pb = _device_vector(d_b, self.count, "d_b")
        px = _device_vector(d_x, self.count, "d_x")
        ps = _device_vector(d_scale, self.count, "d_scale", optional=True)
        fn, keep = self._poll(control)
        return self._measured(
            control,
            lambda info: lib().cfd_solve_scaled_device(self.h, pb, ps, float(tol), int(max_iterations), fn, None, px,
                                                       ctypes.byref(info)),
        )

    def perturb(self, d_src: DeviceArray, d_dst: DeviceArray, rel: float, seed: int):
        """d_dst = d_src + rel rms(d_src) xi on the fluid cells, xi a seeded standard
        normal field (cfd_grid_perturb): a chained campaign right-hand side in HBM."""
        ps = _device_vector(d_src, self.count, "d_src")
        pd = _device_vector(d_dst, self.count, "d_dst")
        _check(lib().cfd_grid_perturb(self.h, ps, pd, float(rel), ctypes.c_uint64(int(seed) & ((1 << 64) - 1))))

    def projection_rhs(self, u, dt):
        u = _host_vector(u, self.face_count, "u")
        b = np.empty(self.count)
        t = np.empty(self.count)
        _check(lib().cfd_projection_rhs(self.h, _ptr(u), float(dt), _ptr(b), _ptr(t)))
        return b, t

    def project(self, u, p, dt, tol=None, max_iterations=None, control=None, static=False):
        """Returns (u, p, info); the inputs are not modified (the reference mutates in place
        and only on success, pressure.rs:468). static=True calls cfd_project_static:
        no device allocation (u and p staged in the library's static device array)."""
        tol = self.tolerance if tol is None else tol
        max_iterations = self.max_iterations if max_iterations is None else max_iterations
        u = _host_vector(u, self.face_count, "u", copy=True)
        p = _host_vector(p, self.count, "p", copy=True)
        fn, keep = self._poll(control)
        # pressure_projection encloses the solve (pressure.rs:478); one library call
        # forms the right-hand side, solves and corrects, all inside pressure_solve
        outer = control.span("pressure_projection") if control is not None else profiling.Span(None, "")
        with outer:
            info = self._measured(
                control,
                lambda info: (lib().cfd_project_static if static else lib().cfd_project)(
                    self.h, _ptr(u), _ptr(p), float(dt), float(tol), int(max_iterations), fn, None, ctypes.byref(info)),
            )
        return u, p, info


class GridCache:
    """The reference's per-grid preconditioner cache (pressure.rs:55-97): the
    device context is reused while the case gives the same operator
    (Grid.matches, the Signature check) and rebuilt when it changes; the first
    solve after a rebuild reports hierarchy_builds = 1."""

    def __init__(self):
        self.grid = None

    def get(self, case: Case) -> Grid:
        if self.grid is None or not self.grid.matches(case):
            if self.grid is not None:
                self.grid.close()
            self.grid = None
            self.grid = Grid(case)
        return self.grid


class Server:
    """Continuous batching of pressure solves on resident grids (cfd_server_*).

    lanes: requests on red-black multigrid grids without closed components run
    in that many concurrent lanes, each a batch of up to 8 solves per precision
    (one launch per iteration stage for the batch); 0: each request on its
    grid's own stream."""

    def __init__(self, max_active: int = 8, hbm_budget_bytes: int = 0, lanes: int = 3):
        h = ctypes.c_void_p()
        _check(lib().cfd_server_create(int(max_active), int(hbm_budget_bytes), ctypes.byref(h)))
        self.h = h
        self.grids = []
        self.pending = {}
        _check(lib().cfd_server_set_batching(self.h, int(lanes)))

    def admit(self, case: Case) -> "Grid":
        """Grid::build under the server's HBM budget (CfdError ENOMEM when it does not fit)."""
        c, keep = case.to_c()
        gh = ctypes.c_void_p()
        began = time.perf_counter()
        _check(lib().cfd_server_admit(self.h, ctypes.byref(c), ctypes.byref(gh)))
        g = Grid.__new__(Grid)
        g._attach(gh, case, began)
        g.owner = self
        self.grids.append(g)
        return g

    def admit_async(self, case: Case) -> int:
        """Queued admission: the grid is built on the server's builder thread once
        the HBM budget has room for it (admissions in order); a ticket for admitted()."""
        c, keep = case.to_c()
        t = ctypes.c_uint64(0)
        _check(lib().cfd_server_admit_async(self.h, ctypes.byref(c), ctypes.byref(t)))
        self.pending[t.value] = (case, time.perf_counter())
        return t.value

    def admitted(self, ticket: int) -> "Grid":
        """The grid of a queued admission once built, else None; raises CfdError
        if its build failed (ENOMEM: larger than the whole budget)."""
        done = ctypes.c_int(0)
        gh = ctypes.c_void_p()
        rc = lib().cfd_server_admitted(self.h, ctypes.c_uint64(ticket), ctypes.byref(done), ctypes.byref(gh))
        if done.value:
            case, began = self.pending.pop(ticket)
        _check(rc)
        if not done.value:
            return None
        g = Grid.__new__(Grid)
        g._attach(gh, case, began)
        g.owner = self
        self.grids.append(g)
        return g

    def release(self, g: "Grid"):
        _check(lib().cfd_server_release(self.h, g.h))
        g.h = None
        self.grids.remove(g)

    def usage(self):
        used = ctypes.c_size_t(0)
        budget = ctypes.c_size_t(0)
        _check(lib().cfd_server_usage(self.h, ctypes.byref(used), ctypes.byref(budget)))
        return used.value, budget.value

    def submit(self, g: "Grid", d_b: "DeviceArray", d_x: "DeviceArray", tol: float, max_iterations: int = 2000,
               precision: int = PRECISION_MIXED, d_scale: "DeviceArray" = None) -> int:
        pb = _device_vector(d_b, g.count, "d_b")
        px = _device_vector(d_x, g.count, "d_x")
        ps = _device_vector(d_scale, g.count, "d_scale", optional=True)
        t = ctypes.c_uint64(0)
        _check(lib().cfd_server_submit(self.h, g.h, pb, ps, float(tol), int(max_iterations), int(precision), px,
                                       ctypes.byref(t)))
        return t.value

    def progress(self) -> int:
        n = lib().cfd_server_progress(self.h)
        if n < 0:
            _check(-n)
        return n

    def wait(self, ticket: int):
        _check(lib().cfd_server_wait(self.h, ctypes.c_uint64(ticket)))

    def finished(self, ticket: int) -> bool:
        """True once the request's result can be taken (result() does not wait)."""
        done = ctypes.c_int(0)
        _check(lib().cfd_server_finished(self.h, ctypes.c_uint64(ticket), ctypes.byref(done)))
        return bool(done.value)

    def drain(self):
        _check(lib().cfd_server_drain(self.h))

    def result(self, ticket: int):
        """(info dict, latency s, queueing s); raises CfdError if the solve failed."""
        info = LinearInfo()
        lat = ctypes.c_double(0.0)
        q = ctypes.c_double(0.0)
        _check(lib().cfd_server_result(self.h, ctypes.c_uint64(ticket), ctypes.byref(info), ctypes.byref(lat),
                                       ctypes.byref(q)))
        return info.as_dict(), lat.value, q.value

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().cfd_server_drain(self.h)
            for g in list(self.grids):
                lib().cfd_server_release(self.h, g.h)
                g.h = None
            self.grids = []
            lib().cfd_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def synchronize():
    _check(lib().cfd_device_synchronize())
