"""ctypes wrapper of the CPU oracle (liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline. The
product (cfd-craft_amd/) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PKG = os.path.join(os.path.dirname(_HERE), "cfd-craft_amd")
if _PKG not in sys.path:
    sys.path.insert(0, _PKG)

from cfdgpu.case import Case  # noqa: E402  (input format only)

_LIB = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liborc.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        D = ctypes.POINTER(ctypes.c_double)
        L.orc_grid_build.restype = P
        L.orc_grid_build.argtypes = [P]
        L.orc_grid_free.argtypes = [P]
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_grid_counts.argtypes = [P, P]
        L.orc_grid_faces.argtypes = [P, P, P, P, P, P]
        L.orc_grid_fluid.argtypes = [P, P]
        L.orc_grid_components.argtypes = [P, P, P]
        L.orc_apply.argtypes = [P, P, P]
        L.orc_operator_apply.argtypes = [P, P, P]
        L.orc_center_nullspaces.argtypes = [P, P]
        L.orc_solve_scaled.argtypes = [P, P, P, ctypes.c_double, ctypes.c_int64, P, P]
        L.orc_solve_scaled.restype = ctypes.c_int
        L.orc_projection_rhs.argtypes = [P, P, ctypes.c_double, P, P]
        L.orc_project.argtypes = [P, P, P, ctypes.c_double, ctypes.c_double, ctypes.c_int64, P]
        L.orc_ic0_apply.argtypes = [P, P, P]
        L.orc_ic0_diagonal.argtypes = [P, P]
        L.orc_mg_apply.argtypes = [P, P, P]
        L.orc_mg_stats.argtypes = [P, P, P, ctypes.c_int, P]
        L.orc_mg_shape.argtypes = [P, P, P, ctypes.c_int, P]
        L.orc_divergence.argtypes = [P, P, P]
        L.orc_gradient.argtypes = [P, P, P]
        L.orc_enforce_velocity.argtypes = [P, P]
        L.orc_flow_vector.argtypes = [P, ctypes.c_double, D]
        L.orc_poly_contains.argtypes = [P, ctypes.c_int32, ctypes.c_double, ctypes.c_double]
        L.orc_poly_contains.restype = ctypes.c_int
        L.orc_masked_level_sizes.argtypes = [P, ctypes.c_int, P]
        L.orc_masked_level_sizes.restype = ctypes.c_int
        L.orc_masked_level.argtypes = [P, ctypes.c_int] + [P] * 8
        L.orc_masked_level.restype = ctypes.c_int
        _LIB = L
    return _LIB


class Info(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int64),
        ("relative_residual", ctypes.c_double),
        ("operator_applications", ctypes.c_int64),
        ("preconditioner_applications", ctypes.c_int64),
        ("true_residual_checks", ctypes.c_int64),
        ("cycles", ctypes.c_int64),
        ("smoothing_cell_updates", ctypes.c_int64),
        ("residual_cell_updates", ctypes.c_int64),
        ("coarse_solves", ctypes.c_int64),
        ("hierarchy_builds", ctypes.c_int64),
        ("restriction_cell_updates", ctypes.c_int64),
        ("prolongation_cell_updates", ctypes.c_int64),
        ("block_smoothing_cell_updates", ctypes.c_int64),
        ("block_solves", ctypes.c_int64),
        ("block_level_cell_visits", ctypes.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class OracleError(RuntimeError):
    def __init__(self, code, text):
        super().__init__(text)
        self.code = code


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def poly_contains(polygon, p) -> bool:
    """geometry.rs:13-26 contains(polygon, p) for one point."""
    xy = np.ascontiguousarray(np.asarray(polygon, dtype=np.float64).reshape(-1))
    return bool(lib().orc_poly_contains(_ptr(xy), len(xy) // 2, float(p[0]), float(p[1])))


class Grid:
    """The oracle's Grid (grid.rs:114-162) with its cached factor / hierarchy."""

    def __init__(self, case: Case):
        self.case = case
        c, keep = case.to_c()
        h = lib().orc_grid_build(ctypes.byref(c))
        if not h:
            raise OracleError(0, lib().orc_last_error().decode())
        self.h = ctypes.c_void_p(h)
        counts = np.zeros(8, dtype=np.int64)
        lib().orc_grid_counts(self.h, _ptr(counts))
        self.count, self.face_count, self.fluid_count, self.n_components, self.n_anchored = (int(v) for v in counts[:5])

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_grid_free(self.h)
            self.h = None

    def faces(self):
        n = self.face_count
        axis = np.zeros(n, np.int32)
        lo = np.zeros(n, np.int64)
        hi = np.zeros(n, np.int64)
        kind = np.zeros(n, np.int32)
        w = np.zeros(n, np.float64)
        lib().orc_grid_faces(self.h, _ptr(axis), _ptr(lo), _ptr(hi), _ptr(kind), _ptr(w))
        return dict(axis=axis, lo=lo, hi=hi, kind=kind, weight=w)

    def fluid(self):
        f = np.zeros(self.count, np.uint8)
        lib().orc_grid_fluid(self.h, _ptr(f))
        return f.astype(bool)

    def components(self):
        comp = np.zeros(self.count, np.int64)
        anch = np.zeros(max(1, self.n_components), np.uint8)
        lib().orc_grid_components(self.h, _ptr(comp), _ptr(anch))
        return comp, anch[: self.n_components].astype(bool)

    def apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.count)
        lib().orc_apply(self.h, _ptr(x), _ptr(y))
        return y

    def operator_apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.count)
        lib().orc_operator_apply(self.h, _ptr(x), _ptr(y))
        return y

    def center_nullspaces(self, x):
        x = np.array(x, dtype=np.float64)
        lib().orc_center_nullspaces(self.h, _ptr(x))
        return x

    def solve_scaled(self, b, scale=None, tol=1e-10, max_iterations=1200):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.count)
        info = Info()
        s = np.ascontiguousarray(scale, dtype=np.float64) if scale is not None else None
        code = lib().orc_solve_scaled(
            self.h, _ptr(b), _ptr(s) if s is not None else None, float(tol), int(max_iterations), _ptr(x), ctypes.byref(info)
        )
        if code:
            raise OracleError(code, lib().orc_last_error().decode())
        return x, info.as_dict()

    def solve(self, b, tol=1e-10, max_iterations=1200):
        return self.solve_scaled(b, None, tol, max_iterations)

    def projection_rhs(self, u, dt):
        u = np.ascontiguousarray(u, dtype=np.float64)
        b = np.empty(self.count)
        t = np.empty(self.count)
        lib().orc_projection_rhs(self.h, _ptr(u), float(dt), _ptr(b), _ptr(t))
        return b, t

    def project(self, u, p, dt, tol=1e-10, max_iterations=1200):
        u = np.array(u, dtype=np.float64)
        p = np.array(p, dtype=np.float64)
        info = Info()
        code = lib().orc_project(self.h, _ptr(u), _ptr(p), float(dt), float(tol), int(max_iterations), ctypes.byref(info))
        if code:
            raise OracleError(code, lib().orc_last_error().decode())
        return u, p, info.as_dict()

    def ic0_apply(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.empty(self.count)
        lib().orc_ic0_apply(self.h, _ptr(r), _ptr(z))
        return z

    def ic0_diagonal(self):
        d = np.empty(self.count)
        lib().orc_ic0_diagonal(self.h, _ptr(d))
        return d

    def mg_apply(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.empty(self.count)
        code = lib().orc_mg_apply(self.h, _ptr(r), _ptr(z))
        if code:
            raise OracleError(code, lib().orc_last_error().decode())
        return z

    def mg_stats(self):
        cells = np.zeros(64, np.int64)
        nnz = np.zeros(64, np.int64)
        masked = ctypes.c_int(0)
        n = lib().orc_mg_stats(self.h, _ptr(cells), _ptr(nnz), 64, ctypes.byref(masked))
        if n < 0:
            raise OracleError(-n, lib().orc_last_error().decode())
        return dict(level_cells=cells[:n].tolist(), level_nonzeros=nnz[:n].tolist(), masked=bool(masked.value))

    def mg_shape(self):
        """Level dimensions (unmasked hierarchy) and block smoothers
        (level, axes, block_unknowns, block_count, half_bandwidth)."""
        dims = np.zeros(64 * 3, np.int64)
        blocks = np.zeros(64 * 6, np.int64)
        nb = ctypes.c_int(0)
        n = lib().orc_mg_shape(self.h, _ptr(dims), _ptr(blocks), 64, ctypes.byref(nb))
        if n < 0:
            raise OracleError(-n, lib().orc_last_error().decode())
        out = []
        for k in range(nb.value):
            lvl, mask, unknowns, count, bw, _ = (int(v) for v in blocks[6 * k: 6 * k + 6])
            out.append(dict(level=lvl, axes=[a for a in range(3) if mask >> a & 1], block_unknowns=unknowns,
                            block_count=count, half_bandwidth=bw))
        return dict(level_dimensions=dims[: 3 * n].reshape(n, 3).tolist(), block_smoothers=out)

    def masked_levels(self):
        """The masked aggregation hierarchy (masked.rs:181-291) level by level: dicts
        with n, parent (None at the coarsest level), CSR offsets / columns / values,
        diagonal, closed (list of member arrays in the level's order), and, on
        level 0, fluid (the cell of each unknown)."""
        sizes = np.zeros(6, np.int64)
        nlev = lib().orc_masked_level_sizes(self.h, -1, _ptr(sizes))
        if nlev < 0:
            raise OracleError(-nlev, lib().orc_last_error().decode())
        out = []
        for l in range(nlev):
            lib().orc_masked_level_sizes(self.h, l, _ptr(sizes))
            n, nnz, ncl, nmem, has_parent, nblocks = (int(v) for v in sizes)
            parent = np.zeros(max(1, n), np.int64)
            offsets = np.zeros(n + 1, np.int64)
            columns = np.zeros(max(1, nnz), np.int64)
            values = np.zeros(max(1, nnz))
            diag = np.zeros(max(1, n))
            cptr = np.zeros(ncl + 1, np.int64)
            cmem = np.zeros(max(1, nmem), np.int64)
            fluid = np.zeros(max(1, self.fluid_count), np.int64) if l == 0 else None
            code = lib().orc_masked_level(self.h, l, _ptr(parent), _ptr(offsets), _ptr(columns), _ptr(values), _ptr(diag),
                                          _ptr(cptr), _ptr(cmem), _ptr(fluid) if fluid is not None else None)
            if code:
                raise OracleError(-code, lib().orc_last_error().decode())
            out.append(dict(n=n, parent=parent[:n] if has_parent else None, offsets=offsets, columns=columns[:nnz],
                            values=values[:nnz], diagonal=diag[:n],
                            closed=[cmem[cptr[c]:cptr[c + 1]] for c in range(ncl)], blocks=nblocks,
                            fluid=fluid[: self.fluid_count] if fluid is not None else None))
        return out

    def divergence(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        d = np.empty(self.count)
        lib().orc_divergence(self.h, _ptr(u), _ptr(d))
        return d

    def gradient(self, p):
        p = np.ascontiguousarray(p, dtype=np.float64)
        g = np.empty(self.face_count)
        lib().orc_gradient(self.h, _ptr(p), _ptr(g))
        return g

    def enforce_velocity(self, u):
        u = np.array(u, dtype=np.float64)
        lib().orc_enforce_velocity(self.h, _ptr(u))
        return u
