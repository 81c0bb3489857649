"""Pins the oracle's prism containment (oracle.cpp poly_contains / Raster,
restating crates/cfd-craft/src/geometry.rs:13-26,113-193 and case.rs:436-443)
and checks the library's host rasterizer against the same rules.

  * geometry.rs:278-327: a cell center on the shared slanted edge of two
    abutting footprints lies in exactly one of them, whatever the winding and
    first vertex (the on-edge rule: the footprint whose interior lies on the
    +x side of the edge, or the +y side of a horizontal edge);
  * geometry.rs:328-335: a concave footprint;
  * every convex footprint of the prism-city and street-canyon workloads at
    full size, against an independent half-plane test of each cell center.
"""
import numpy as np
import pytest

import cfdgpu
import oracle
from cfdgpu import workloads as W
from cfdgpu.case import MULTIGRID, OPEN, WALL, Block, Case

SLANTS = [(28, 30, 12, (2, -1)), (27, 0, 10, (2, 1)), (3, 5, 9, (1, 3)), (11, 40, 7, (3, -1)), (0, 0, 12, (1, 1))]


def abutting(dx, i0, j0, steps, slope):
    at = lambda m: [(i0 + 0.5 + m * slope[0]) * dx, (j0 + 0.5 + m * slope[1]) * dx]  # noqa: E731
    a, b = at(0), at(steps)
    left = [a, b, [b[0] - 50.0, b[1]], [a[0] - 50.0, a[1]]]
    right = [b, a, [a[0] + 50.0, a[1]], [b[0] + 50.0, b[1]]]
    return at, left, right


def test_shared_slanted_edge_is_in_exactly_one_footprint():
    checked = 0
    for dx in (0.1, 0.3, 0.7):
        for i0, j0, steps, slope in SLANTS:
            at, left, right = abutting(dx, i0, j0, steps, slope)
            for m in range(1, steps):
                p = at(m)
                assert oracle.poly_contains(left, p) != oracle.poly_contains(right, p), (dx, p)
                assert oracle.poly_contains(left, p) == oracle.poly_contains(left[::-1], p)
                assert oracle.poly_contains(left, p) == oracle.poly_contains(left[1:] + left[:1], p)
                checked += 1
    assert checked > 100
    # the reference's first recorded failure (dx = 0.7)
    a, b = [19.95, 21.349999999999998], [36.75, 12.95]
    p = [21.349999999999998, 20.65]
    left = [a, b, [b[0] - 50.0, b[1]], [a[0] - 50.0, a[1]]]
    right = [b, a, [a[0] + 50.0, a[1]], [b[0] + 50.0, b[1]]]
    assert oracle.poly_contains(left, p) != oracle.poly_contains(right, p)


def test_on_edge_points_go_to_the_plus_x_side():
    """The edge rule itself: a point on a slanted edge belongs to the footprint
    whose interior lies toward +x, on a horizontal edge toward +y."""
    at, left, right = abutting(0.5, 3, 4, 8, (1, 2))
    p = at(3)  # interior of `right` lies toward +x of the edge
    assert oracle.poly_contains(right, p) and not oracle.poly_contains(left, p)
    below = [[0.0, 0.0], [4.0, 0.0], [4.0, 2.0], [0.0, 2.0]]
    above = [[0.0, 2.0], [4.0, 2.0], [4.0, 4.0], [0.0, 4.0]]
    assert oracle.poly_contains(above, [1.5, 2.0]) and not oracle.poly_contains(below, [1.5, 2.0])


def test_concave_footprint():
    p = [[0.0, 0.0], [2.0, 0.0], [2.0, 1.0], [1.0, 1.0], [1.0, 2.0], [0.0, 2.0]]
    assert oracle.poly_contains(p, [0.5, 1.5])
    assert not oracle.poly_contains(p, [1.5, 1.5])
    assert oracle.poly_contains(p, [1.5, 0.5])


def _prism_case(cells, dx, footprints):
    blocks = []
    for fp in footprints:
        xs = [v[0] for v in fp]
        ys = [v[1] for v in fp]
        blocks.append(Block([min(xs), min(ys), 0.0], [max(xs), max(ys), dx], footprint=fp))
    return Case(cells=[cells[0], cells[1], 2], size_m=[cells[0] * dx, cells[1] * dx, 2 * dx], boundaries=[WALL, WALL, WALL],
                blocks=blocks, preconditioner=MULTIGRID)


@pytest.mark.parametrize("dx", [0.3, 0.7])
def test_abutting_prisms_leave_no_hole_on_the_grid(dx):
    """Two prisms sharing a slanted edge through cell centers: the oracle and
    the library's host build mark every cell center on the edge solid in
    exactly one of the two single-prism grids, and in the two-prism grid."""
    for i0, j0, steps, slope in SLANTS:
        at, left, right = abutting(dx, i0, j0, steps, slope)
        n = [int(np.ceil(max(abs(v[0]) for v in left + right) / dx)) + 2, 60]
        n[0] = max(n[0], 60)
        pts = [at(m) for m in range(1, steps)]
        cells = [int(p[0] / dx) + n[0] * int(p[1] / dx) for p in pts if 0 <= p[1] < n[1] * dx and 0 <= p[0] < n[0] * dx]
        masks = []
        for fps in ([left], [right], [left, right]):
            case = _prism_case(n, dx, fps)
            fo = oracle.Grid(case).fluid()
            fh = (cfdgpu.case_topology(case)["codes"] & 0x40) != 0
            assert np.array_equal(fo, fh)
            masks.append(fo[: n[0] * n[1]])
        for c in cells:
            assert masks[0][c] != masks[1][c]
            assert not masks[2][c]


def _half_plane_raster(footprints, n, dx):
    """Solid cell centers by an independent test: a center strictly inside a
    convex footprint is on the inner side of every edge; centers within 1e-9 m of
    an edge line are reported separately (the on-edge rule decides those)."""
    xc = (np.arange(n[0]) + 0.5) * dx
    yc = (np.arange(n[1]) + 0.5) * dx
    X, Y = np.meshgrid(xc, yc)
    solid = np.zeros(X.shape, bool)
    near = np.zeros(X.shape, bool)
    for fp in footprints:
        P = np.asarray(fp, float)
        k = len(P)
        area2 = sum(P[i, 0] * P[(i + 1) % k, 1] - P[(i + 1) % k, 0] * P[i, 1] for i in range(k))
        orient = 1.0 if area2 > 0 else -1.0
        inside = np.ones(X.shape, bool)
        close = np.zeros(X.shape, bool)
        for i in range(k):
            a, b = P[i], P[(i + 1) % k]
            cr = orient * ((b[0] - a[0]) * (Y - a[1]) - (b[1] - a[1]) * (X - a[0]))
            length = np.hypot(b[0] - a[0], b[1] - a[1])
            inside &= cr > 0
            close |= np.abs(cr) / length < 1e-9
        lo = P.min(axis=0)
        hi = P.max(axis=0)
        box = (X >= lo[0]) & (X < hi[0]) & (Y >= lo[1]) & (Y < hi[1])
        solid |= inside & box
        near |= close & box
    return solid, near


@pytest.mark.parametrize("workload", ["prisms", "canyon"])
def test_convex_footprints_match_half_planes(workload):
    """Every footprint of the full-size workloads (200 prisms on 256x256, 400 on
    384x384), rasterized by the oracle and by the library's host build, against
    the half-plane test of each cell center (one cell layer: the footprints are
    two-dimensional)."""
    if workload == "prisms":
        full = W.urban_prisms()[0]
    else:
        full = W.street_canyon(0.0)[0]
    n = [int(v) for v in full.cells[:2]]
    dx = full.size_m[0] / full.cells[0]
    fps = [b.footprint for b in full.blocks]
    case = _prism_case(n, dx, fps)
    fo = oracle.Grid(case).fluid()[: n[0] * n[1]].reshape(n[1], n[0])
    fh = ((cfdgpu.case_topology(case)["codes"] & 0x40) != 0)[: n[0] * n[1]].reshape(n[1], n[0])
    assert np.array_equal(fo, fh)
    solid, near = _half_plane_raster(fps, n, dx)
    decided = ~near
    assert decided.sum() > 0.99 * decided.size
    assert np.array_equal(~fo[decided], solid[decided])
    assert solid.mean() > 0.15
