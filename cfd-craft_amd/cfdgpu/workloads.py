"""Seeded synthetic pressure-solve workloads and the reference's own fixtures.

Fixtures of the reference (inputs of its published iteration counts):
  * ``cavity(n)`` / ``cube(n)``: cases/verification/lid-cavity.json with walls on x
    and z and a periodic y axis (docs/evidence/pressure-multigrid.md:25-38);
  * ``manufactured(case, fluid)``: the performance probe's field
    sin(2 pi x/Lx) + 0.3 cos(4 pi z/Lz) on fluid cells (performance_probe.rs:82-95).

Benchmark configurations (pressure solves of staggered Cartesian grids):
  * ``lid_cavity_48`` : 48^3 closed cavity, IC(0), RHS = -div(u) of U(-1,1) faces;
  * ``urban_block``   : 128x128x64, h = 2 m, 24 ground boxes, open x with a log-law
                        inflow, RHS from the perturbed log-law velocity;
  * ``urban_prisms``  : 256x256x128, 200 convex polygon prisms;
  * ``served_mix``    : requests over {64^3, 96x96x48, 128x128x64, 192x192x96}.
Velocity fields follow Grid::enforce_velocity (grid.rs:683-691) before the
right-hand side -div(u)/dt of projection_rhs (pressure.rs:440-459) is formed.
"""
from __future__ import annotations

import math
from typing import List, Tuple

import numpy as np

from .case import (
    BLOCKED,
    FLOW_NEUTRAL_RANS,
    GROUND_OPEN_TOP,
    IC0,
    INLET,
    INTERIOR,
    MULTIGRID,
    OPEN,
    OUTLET,
    PERIODIC,
    SLIP,
    SLIP_FACE,
    WALL,
    WALL_FACE,
    Block,
    Case,
)


# ---------------------------------------------------------------- geometry in numpy
def _centers(n, dx):
    return (np.arange(n) + 0.5) * dx


def _in_polygon(poly, x, y):
    """Even-odd crossings (geometry.rs:13-26), vectorised over points."""
    inside = np.zeros(np.broadcast(x, y).shape, bool)
    k = len(poly)
    for i in range(k):
        a = poly[i]
        b = poly[(i + 1) % k]
        crosses = (a[1] > y) != (b[1] > y)
        lo, hi = (a, b) if a[1] < b[1] else (b, a)
        with np.errstate(divide="ignore", invalid="ignore"):
            xi = lo[0] + (y - lo[1]) * (hi[0] - lo[0]) / (hi[1] - lo[1])
        inside ^= crosses & (x < xi)
    return inside


def fluid_mask(case: Case) -> np.ndarray:
    """Cell-center mask of the case's boxes and prisms (geometry.rs:113-193), x fastest."""
    n = [int(v) for v in case.cells]
    dx = case.dx
    fluid = np.ones((n[2], n[1], n[0]), bool)
    for b in case.blocks:
        c = [_centers(n[a], dx[a]) for a in range(3)]
        sel = [np.nonzero((c[a] >= b.min_m[a]) & (c[a] < b.max_m[a]))[0] for a in range(3)]
        if any(s.size == 0 for s in sel):
            continue
        X, Y = np.meshgrid(c[0][sel[0]], c[1][sel[1]])
        cols = np.ones(X.shape, bool)
        if b.footprint:
            cols = _in_polygon(b.footprint, X, Y)
        sub = fluid[sel[2][0] : sel[2][-1] + 1, sel[1][0] : sel[1][-1] + 1, sel[0][0] : sel[0][-1] + 1]
        sub[:, cols] = False
    return fluid.reshape(-1)


def face_kinds(case: Case, fluid: np.ndarray):
    """Face kinds and face-center heights in the reference's face order (grid.rs:341-429)."""
    n = [int(v) for v in case.cells]
    dx = case.dx
    F = fluid.reshape(n[2], n[1], n[0])
    kinds, zs = [], []
    for a in range(3):
        periodic = case.boundaries[a] == PERIODIC
        dims = list(n)
        if not periodic:
            dims[a] += 1
        ax = 2 - a  # numpy axis of grid axis a
        shape = (dims[2], dims[1], dims[0])
        # lo cell fluid / exists, hi cell fluid / exists
        lo_f = np.zeros(shape, bool)
        hi_f = np.zeros(shape, bool)
        lo_e = np.zeros(shape, bool)
        hi_e = np.zeros(shape, bool)
        t = np.arange(dims[a])
        idx = [slice(None)] * 3
        if periodic:
            lo_f[...] = np.roll(F, 1, axis=ax)
            hi_f[...] = F
            lo_e[...] = True
            hi_e[...] = True
        else:
            idx_lo = [slice(None)] * 3
            idx_lo[ax] = slice(1, None)
            lo_f[tuple(idx_lo)] = F
            lo_e[tuple(idx_lo)] = True
            idx_hi = [slice(None)] * 3
            idx_hi[ax] = slice(0, n[a])
            hi_f[tuple(idx_hi)] = F
            hi_e[tuple(idx_hi)] = True
        kind = np.full(shape, BLOCKED, np.int32)
        kind[lo_f & hi_f] = INTERIOR
        kind[(lo_f ^ hi_f) & lo_e & hi_e] = WALL_FACE
        boundary = (lo_f ^ hi_f) & ~(lo_e & hi_e)
        zc = np.broadcast_to(((np.arange(dims[2]) + (0.0 if a == 2 else 0.5)) * dx[2])[:, None, None], shape)
        bnd = case.boundaries[a]
        if bnd == WALL:
            kind[boundary] = WALL_FACE
        elif bnd == SLIP:
            kind[boundary] = SLIP_FACE
        elif bnd == GROUND_OPEN_TOP:
            tt = np.broadcast_to(t.reshape([-1 if d == ax else 1 for d in range(3)]), shape)
            kind[boundary & (tt == 0)] = WALL_FACE
            kind[boundary & (tt != 0)] = SLIP_FACE
        elif bnd == OPEN:
            tt = np.broadcast_to(t.reshape([-1 if d == ax else 1 for d in range(3)]), shape)
            normal = np.where(tt == 0, -1.0, 1.0)[boundary]
            zb = np.ascontiguousarray(zc)[boundary]
            uz, inv = np.unique(zb, return_inverse=True)
            va = np.array([case.flow_vector(z)[a] for z in uz])[inv]
            v = va * normal
            kind[boundary] = np.where(v < -1e-12, INLET, np.where(v > 1e-12, OUTLET, SLIP_FACE))
        kinds.append(kind.reshape(-1))
        zs.append(np.ascontiguousarray(zc).reshape(-1))
        del idx
    return np.concatenate(kinds), np.concatenate(zs), [sum(k.size for k in kinds[:a]) for a in range(3)]


def enforce_velocity(case: Case, kinds, zc, offsets, u):
    """Grid::enforce_velocity (grid.rs:683-691): zero on wall, slip and blocked
    faces, the case inflow on inlet faces."""
    u = np.array(u, dtype=np.float64)
    u[(kinds == WALL_FACE) | (kinds == SLIP_FACE) | (kinds == BLOCKED)] = 0.0
    for a in range(3):
        end = offsets[a + 1] if a < 2 else u.size
        sel = np.nonzero(kinds[offsets[a] : end] == INLET)[0] + offsets[a]
        if sel.size:
            uz, inv = np.unique(zc[sel], return_inverse=True)
            u[sel] = np.array([case.flow_vector(z)[a] for z in uz])[inv]
    return u


def projection_rhs(case: Case, fluid: np.ndarray, u: np.ndarray, dt: float = 1.0):
    """-div(u)/dt and its round-off scale (pressure.rs:440-459), in numpy."""
    n = [int(v) for v in case.cells]
    dx = case.dx
    off = 0
    div = np.zeros((n[2], n[1], n[0]))
    scale = np.zeros((n[2], n[1], n[0]))
    for a in range(3):
        periodic = case.boundaries[a] == PERIODIC
        dims = list(n)
        if not periodic:
            dims[a] += 1
        m = math.prod(dims)
        U = u[off : off + m].reshape(dims[2], dims[1], dims[0])
        off += m
        ax = 2 - a
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax] = slice(0, n[a])
        if periodic:
            Uh = np.roll(U, -1, axis=ax)
            ul, uh = U, Uh
        else:
            hi[ax] = slice(1, n[a] + 1)
            ul, uh = U[tuple(lo)], U[tuple(hi)]
        div = div + (uh - ul) / dx[a]
        scale = scale + (np.abs(ul) + np.abs(uh)) / dx[a]
    b = (-div / dt).reshape(-1)
    t = (scale / dt).reshape(-1)
    b[~fluid] = -0.0
    t[~fluid] = 0.0
    return b, t


# ---------------------------------------------------------------- reference fixtures
def cavity(cells, preconditioner=IC0) -> Case:
    """lid-cavity.json with its grid changed: walls on x and z, periodic y, unit domain."""
    return Case(cells=list(cells), size_m=[1.0, 1.0, 1.0], boundaries=[WALL, PERIODIC, WALL], preconditioner=preconditioner,
                tolerance=1e-11, max_iterations=10000)


def manufactured(case: Case, fluid: np.ndarray) -> np.ndarray:
    """sin(2 pi x/Lx) + 0.3 cos(4 pi z/Lz) on fluid cells (performance_probe.rs:82-95),
    before its gauge is removed."""
    n = [int(v) for v in case.cells]
    dx = case.dx
    i = np.arange(case.count)
    x = (i % n[0] + 0.5) * dx[0]
    z = (i // (n[0] * n[1]) + 0.5) * dx[2]
    ex = np.sin(2.0 * math.pi * x / case.size_m[0]) + 0.3 * np.cos(4.0 * math.pi * z / case.size_m[2])
    ex[~fluid] = 0.0
    return ex


# ---------------------------------------------------------------- benchmark configurations
def lid_cavity_48(seed: int = 1, n: int = 48) -> Tuple[Case, np.ndarray, np.ndarray]:
    """Closed 48^3 cavity, IC(0), tol 1e-10: RHS = -div(u) of interior faces U(-1,1).
    The unit lid is tangential and reaches only momentum, not this RHS."""
    case = Case(cells=[n, n, n], size_m=[1.0, 1.0, 1.0], boundaries=[WALL, WALL, WALL], preconditioner=IC0,
                tolerance=1e-10, max_iterations=1200)
    rng = np.random.default_rng(seed)
    fluid = np.ones(case.count, bool)
    kinds, zc, offs = face_kinds(case, fluid)
    u = rng.uniform(-1.0, 1.0, kinds.size)
    u = enforce_velocity(case, kinds, zc, offs, u)
    b, t = projection_rhs(case, fluid, u)
    return case, b, t


def _neutral_rans(case: Case, speed=5.0, zref=10.0, z0=0.1, direction=270.0):
    case.flow_kind = FLOW_NEUTRAL_RANS
    case.speed_m_s = speed
    case.reference_height_m = zref
    case.roughness_m = z0
    case.direction_deg_from = direction


def _log_law_velocity(case: Case, kinds, zc, offs, rng, noise=0.05):
    u = np.empty(kinds.size)
    for a in range(3):
        end = offs[a + 1] if a < 2 else kinds.size
        z = zc[offs[a] : end]
        uzs, inv = np.unique(z, return_inverse=True)
        vec = np.array([case.flow_vector(v) for v in uzs])
        base = vec[inv, a]
        mag = np.hypot(vec[:, 0], vec[:, 1])[inv]
        u[offs[a] : end] = base + noise * mag * rng.standard_normal(z.size)
    return enforce_velocity(case, kinds, zc, offs, u)


def urban_block(seed: int = 2, n=(128, 128, 64), h: float = 2.0, boxes: int = 24, preconditioner=MULTIGRID):
    """Urban block: ground boxes with footprints 8-24 cells and heights 10-40 cells,
    at least 4 cells from the x boundaries; open x (log-law inflow from the west),
    slip y, ground_open_top z."""
    n = [int(v) for v in n]
    case = Case(cells=n, size_m=[n[0] * h, n[1] * h, n[2] * h], boundaries=[OPEN, SLIP, GROUND_OPEN_TOP],
                preconditioner=preconditioner, tolerance=1e-10, max_iterations=1200)
    _neutral_rans(case)
    rng = np.random.default_rng(seed)
    for k in range(boxes):
        wx, wy = (int(v) for v in rng.integers(8, 25, 2))
        hz = int(rng.integers(10, 41))
        x0 = int(rng.integers(4, n[0] - 4 - wx + 1))
        y0 = int(rng.integers(0, n[1] - wy + 1))
        hz = min(hz, n[2] - 1)
        case.blocks.append(Block(min_m=[x0 * h, y0 * h, 0.0], max_m=[(x0 + wx) * h, (y0 + wy) * h, hz * h], id=f"b{k}"))
    return _finish(case, rng)


def wind_campaign(direction_deg: float, seed: int = 5, n=(128, 128, 64), h: float = 2.0, boxes: int = 24,
                  preconditioner=MULTIGRID):
    """One direction of a wind-direction campaign: the same seeded boxes (at least
    4 cells from every side boundary), x and y open with the log-law wind blowing
    from direction_deg, ground_open_top z. The direction decides which open faces
    are outlets, so each direction is its own operator."""
    n = [int(v) for v in n]
    case = Case(cells=n, size_m=[n[0] * h, n[1] * h, n[2] * h], boundaries=[OPEN, OPEN, GROUND_OPEN_TOP],
                preconditioner=preconditioner, tolerance=1e-8, max_iterations=2000)
    _neutral_rans(case, direction=direction_deg)
    rng = np.random.default_rng(seed)
    for k in range(boxes):
        wx, wy = (int(v) for v in rng.integers(8, 25, 2))
        hz = min(int(rng.integers(10, 41)), n[2] - 1)
        x0 = int(rng.integers(4, n[0] - 4 - wx + 1))
        y0 = int(rng.integers(4, n[1] - 4 - wy + 1))
        case.blocks.append(Block(min_m=[x0 * h, y0 * h, 0.0], max_m=[(x0 + wx) * h, (y0 + wy) * h, hz * h], id=f"b{k}"))
    return _finish(case, rng)


def _convex_polygon(rng, cx, cy, r, k):
    ang = np.sort(rng.uniform(0, 2 * math.pi, k))
    rad = r * rng.uniform(0.7, 1.0, k)
    return [[cx + rad[i] * math.cos(ang[i]), cy + rad[i] * math.sin(ang[i])] for i in range(k)]


def _valid_convex(p):
    k = len(p)
    sgn = 0
    for i in range(k):
        a, b, c = p[i], p[(i + 1) % k], p[(i + 2) % k]
        cr = (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])
        if cr == 0:
            return False
        s = 1 if cr > 0 else -1
        if sgn == 0:
            sgn = s
        elif s != sgn:
            return False
    return True


def urban_prisms(seed: int = 3, n=(256, 256, 128), h: float = 1.0, prisms: int = 200, preconditioner=MULTIGRID,
                 radius=(13.0, 27.0), heights=(5, 60), margin=4):
    """Prism buildings: random convex 4-8-vertex footprints, heights 5-60 cells;
    the default radii give about 20 % solid cells at 256x256x128."""
    n = [int(v) for v in n]
    case = Case(cells=n, size_m=[n[0] * h, n[1] * h, n[2] * h], boundaries=[OPEN, SLIP, GROUND_OPEN_TOP],
                preconditioner=preconditioner, tolerance=1e-10, max_iterations=1200)
    _neutral_rans(case)
    rng = np.random.default_rng(seed)
    made = 0
    while made < prisms:
        r = rng.uniform(*radius) * h
        k = int(rng.integers(4, 9))
        cx = rng.uniform(margin * h + r, n[0] * h - margin * h - r)
        cy = rng.uniform(r, n[1] * h - r)
        poly = _convex_polygon(rng, cx, cy, r, k)
        if not _valid_convex(poly):
            continue
        xs = [p[0] for p in poly]
        ys = [p[1] for p in poly]
        hz = min(int(rng.integers(heights[0], heights[1] + 1)), n[2] - 1)
        b = Block(min_m=[min(xs), min(ys), 0.0], max_m=[max(xs), max(ys), hz * h], footprint=poly, id=f"p{made}")
        # the grid must represent it (CFD-E028): at least one contained cell center
        cxs = _centers(n[0], h)
        cys = _centers(n[1], h)
        X, Y = np.meshgrid(cxs[(cxs >= b.min_m[0]) & (cxs < b.max_m[0])], cys[(cys >= b.min_m[1]) & (cys < b.max_m[1])])
        if X.size == 0 or not _in_polygon(poly, X, Y).any():
            continue
        case.blocks.append(b)
        made += 1
    return _finish(case, rng)


def street_canyon(direction_deg: float, seed: int = 5, n=(384, 384, 160), h: float = 1.0, lots: int = 20,
                  street: float = 3.0, heights=(30, 155), margin=4, preconditioner=MULTIGRID):
    """One direction of the street-canyon wind campaign: lots x lots building lots
    (400 by default) between streets of `street` metres, each holding one prism
    with a random convex 4-8-vertex footprint that nearly fills the lot and a
    height of heights cells (about 25 % solid cells at 384x384x160); x and y open
    with the log-law wind blowing from direction_deg, ground_open_top z. The same
    seed gives the same city for every direction; the direction decides which
    open faces are outlets, so each direction is its own operator. Returns the
    case, the right-hand side -div(u) of the seeded log-law field and its scale."""
    return _finish(*_street_canyon_city(direction_deg, seed, n, h, lots, street, heights, margin, preconditioner))


def street_canyon_case(direction_deg: float, seed: int = 5, n=(384, 384, 160), h: float = 1.0, lots: int = 20,
                       street: float = 3.0, heights=(30, 155), margin=4, preconditioner=MULTIGRID) -> Case:
    """The case of street_canyon alone (the same city and wind), without forming
    its right-hand side."""
    return _street_canyon_city(direction_deg, seed, n, h, lots, street, heights, margin, preconditioner)[0]


def _street_canyon_city(direction_deg, seed, n, h, lots, street, heights, margin, preconditioner):
    n = [int(v) for v in n]
    case = Case(cells=n, size_m=[n[0] * h, n[1] * h, n[2] * h], boundaries=[OPEN, OPEN, GROUND_OPEN_TOP],
                preconditioner=preconditioner, tolerance=1e-8, max_iterations=2000)
    _neutral_rans(case, direction=direction_deg)
    rng = np.random.default_rng(seed)
    lx = (n[0] - 2 * margin) * h / lots
    ly = (n[1] - 2 * margin) * h / lots
    cxs = _centers(n[0], h)
    cys = _centers(n[1], h)
    for a in range(lots):
        for b in range(lots):
            cx = margin * h + (a + 0.5) * lx
            cy = margin * h + (b + 0.5) * ly
            hx, hy = 0.5 * (lx - street), 0.5 * (ly - street)
            while True:
                k = int(rng.integers(4, 9))
                # one vertex per sector of 2 pi / k, on the lot's rectangle (less the
                # street), pulled in by up to 15 %
                ang = (rng.uniform(0, 2 * math.pi / k) + 2 * math.pi * (np.arange(k) + rng.uniform(0.15, 0.85, k)) / k)
                rad = 1.0 / np.maximum(np.abs(np.cos(ang)) / hx, np.abs(np.sin(ang)) / hy)
                rad = rad * rng.uniform(0.85, 1.0, k)
                poly = [[cx + rad[i] * math.cos(ang[i]), cy + rad[i] * math.sin(ang[i])] for i in range(k)]
                if not _valid_convex(poly):
                    continue
                xs = [p[0] for p in poly]
                ys = [p[1] for p in poly]
                X, Y = np.meshgrid(cxs[(cxs >= min(xs)) & (cxs < max(xs))], cys[(cys >= min(ys)) & (cys < max(ys))])
                if X.size and _in_polygon(poly, X, Y).any():
                    break
            hz = min(int(rng.integers(heights[0], heights[1] + 1)), n[2] - 1)
            case.blocks.append(Block(min_m=[min(xs), min(ys), 0.0], max_m=[max(xs), max(ys), hz * h], footprint=poly,
                                     id=f"c{a}_{b}"))
    return case, rng


def _finish(case: Case, rng):
    fluid = fluid_mask(case)
    kinds, zc, offs = face_kinds(case, fluid)
    u = _log_law_velocity(case, kinds, zc, offs, rng)
    b, t = projection_rhs(case, fluid, u)
    return case, b, t


SERVED_SHAPES = [(64, 64, 64), (96, 96, 48), (128, 128, 64), (192, 192, 96)]
SERVED_TOLERANCES = [1e-6, 1e-8, 1e-10]


def served_request(seed: int, request: int) -> Tuple[Case, np.ndarray, np.ndarray, float]:
    """One request of the served mix: a shape from SERVED_SHAPES, seeded boxes giving
    5-25 % solid cells, and a tolerance from SERVED_TOLERANCES."""
    rng = np.random.default_rng([seed, request])
    shape = SERVED_SHAPES[int(rng.integers(0, len(SERVED_SHAPES)))]
    tol = SERVED_TOLERANCES[int(rng.integers(0, len(SERVED_TOLERANCES)))]
    target = rng.uniform(0.05, 0.25)
    n = list(shape)
    h = 2.0
    case = Case(cells=n, size_m=[n[0] * h, n[1] * h, n[2] * h], boundaries=[OPEN, SLIP, GROUND_OPEN_TOP],
                preconditioner=MULTIGRID, tolerance=tol, max_iterations=2000)
    _neutral_rans(case)
    solid = 0.0
    k = 0
    while solid < target and k < 400:
        wx, wy = (int(v) for v in rng.integers(max(3, n[0] // 16), max(4, n[0] // 5), 2))
        hz = int(rng.integers(max(2, n[2] // 8), max(3, (2 * n[2]) // 3)))
        x0 = int(rng.integers(4, n[0] - 4 - wx + 1))
        y0 = int(rng.integers(0, n[1] - wy + 1))
        case.blocks.append(Block(min_m=[x0 * h, y0 * h, 0.0], max_m=[(x0 + wx) * h, (y0 + wy) * h, hz * h], id=f"s{k}"))
        k += 1
        solid = 1.0 - fluid_mask(case).mean() if k % 4 == 0 else solid + wx * wy * hz / math.prod(n)
    case2, b, t = _finish(case, rng)
    return case2, b, t, tol
