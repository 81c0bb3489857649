"""Pins the oracle's masked aggregation hierarchy (oracle.cpp masked_build,
restating crates/cfd-craft/src/multigrid/masked.rs:90-291) with checks that do
not share its code: a brute-force reading of the aggregation rule, dense
Galerkin products, and the reference's own unit tests restated on the same
fixtures.

The rule being pinned (masked.rs:1-6,127-156,181-291): level 0 is the fluid
cells in index order; every level groups its unknowns into the connected
pieces of the geometric bins of twice its coordinate spacing (growing the bins
while that reduces nothing), numbers the pieces by their lowest member, sums
the couplings between distinct pieces and the outlet weights of every piece
(Galerkin with a piecewise-constant prolongation), and stops at the first
level whose fluid components all have at most 32 unknowns, which is solved
by one dense factor per component.

A mistake of the kind the reference's comments warn about fails here:
aggregating by bin without splitting it into connected pieces (a non-connected
aggregate), a different bin growth (the minimal-shift check), a different
stop rule (the 32-unknown check on the terminal and non-terminal levels), a
coupling dropped or double counted (the dense Galerkin product), or a sealed
component merged with another (the closed-component bookkeeping).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from cfdgpu import workloads as W
from cfdgpu.case import MULTIGRID, OPEN, PERIODIC, WALL, Block, Case

COARSE_CELLS = 32


def lcg_mask(count, seed=12345, percent=45):
    """The reference's porous mask generator (masked.rs:553-560): a 64-bit LCG,
    solid where (state >> 33) % 100 < percent."""
    state = seed
    out = np.zeros(count, bool)
    for k in range(count):
        state = (state * 6364136223846793005 + 1442695040888963407) % (1 << 64)
        out[k] = (state >> 33) % 100 < percent
    return out


def topology(cells, boundaries, solid):
    """masked.rs:475-507: unit cells of lid-cavity.json with the given boundaries,
    flow [1, 0, 0] (an open x axis has its inlet at x-lo and its outlet at x-hi)
    and one unit box per solid cell."""
    blocks = []
    for z in range(cells[2]):
        for y in range(cells[1]):
            for x in range(cells[0]):
                if solid((x, y, z)):
                    blocks.append(Block([float(x), float(y), float(z)], [x + 1.0, y + 1.0, z + 1.0]))
    return Case(cells=list(cells), size_m=[float(v) for v in cells], boundaries=list(boundaries),
                velocity_m_s=[1.0, 0.0, 0.0], blocks=blocks, preconditioner=MULTIGRID)


def barrier():
    """masked.rs:348-363: the 8x2x6 cavity (periodic y) cut by a full barrier."""
    return Case(cells=[8, 2, 6], size_m=[8.0, 2.0, 6.0], boundaries=[WALL, PERIODIC, WALL],
                blocks=[Block([3.0, 0.0, 0.0], [4.0, 2.0, 6.0])], preconditioner=MULTIGRID)


def thin_passage():
    return topology([10, 2, 7], [OPEN, PERIODIC, WALL], lambda q: q[0] == 4 and q[2] != 3)


POCKET = (6, 1, 3)


def sealed_side():
    def solid(q):
        around = sum(abs(q[a] - POCKET[a]) for a in range(3)) == 1
        return q[0] == 3 or around or tuple(q) in ((7, 0, 0), (8, 1, 0), (8, 0, 1))

    return topology([9, 3, 6], [OPEN, WALL, WALL], solid)


def periodic_wrap():
    return topology([9, 1, 8], [PERIODIC, PERIODIC, WALL], lambda q: q[0] in (2, 6))


def porous_periodic():
    mask = lcg_mask(8 * 6 * 6)
    return topology([8, 6, 6], [PERIODIC, PERIODIC, PERIODIC], lambda q: mask[q[0] + 8 * (q[1] + 6 * q[2])])


def courtyards():
    """Three sealed courtyards (full-height rings of wall) beside an
    outlet-anchored street in one grid."""
    blocks = []
    for x0, y0, w in ((3, 3, 8), (14, 4, 7), (4, 12, 7)):
        blocks += [Block([x0, y0, 0.0], [x0 + w, y0 + 1.0, 9.0]), Block([x0, y0 + w - 1.0, 0.0], [x0 + w, y0 + w, 9.0]),
                   Block([x0, y0, 0.0], [x0 + 1.0, y0 + w, 9.0]), Block([x0 + w - 1.0, y0, 0.0], [x0 + w, y0 + w, 9.0])]
    blocks = [Block([float(v) for v in b.min_m], [float(v) for v in b.max_m]) for b in blocks]
    return Case(cells=[24, 20, 9], size_m=[24.0, 20.0, 9.0], boundaries=[OPEN, WALL, WALL], velocity_m_s=[1.0, 0, 0],
                blocks=blocks, preconditioner=MULTIGRID)


def random_masks():
    out = []
    for seed, pct, bnd in ((7, 30, [WALL, WALL, WALL]), (11, 38, [OPEN, PERIODIC, WALL]),
                           (13, 20, [PERIODIC, PERIODIC, PERIODIC]), (17, 42, [OPEN, WALL, PERIODIC])):
        n = [14, 12, 10]
        mask = lcg_mask(int(np.prod(n)), seed=seed, percent=pct)
        out.append((f"lcg{seed}-{pct}", topology(n, bnd, lambda q, m=mask, n=n: m[q[0] + n[0] * (q[1] + n[1] * q[2])])))
    return out


def fixtures():
    urban = W.urban_block(n=(32, 32, 16), boxes=6)[0]
    prisms = W.urban_prisms(n=(40, 36, 20), prisms=10, radius=(3.0, 6.0), heights=(3, 12))[0]
    canyon = W.street_canyon(200.0, n=(48, 48, 20), lots=4, street=3.0, heights=(5, 15))[0]
    return [("barrier", barrier()), ("thin-passage", thin_passage()), ("sealed-side", sealed_side()),
            ("periodic-wrap", periodic_wrap()), ("porous-periodic", porous_periodic()), ("courtyards", courtyards()),
            ("urban32", urban), ("prisms40", prisms), ("canyon48", canyon)] + random_masks()


FIXTURES = fixtures()
IDS = [n for n, _ in FIXTURES]


# ---------------------------------------------------------------- brute-force helpers
def csr_matrix(level):
    """The level's operator as a scipy matrix (stored off-diagonals plus the diagonal)."""
    n = level["n"]
    rows = np.repeat(np.arange(n), np.diff(level["offsets"]))
    A = sp.csr_matrix((level["values"], (rows, level["columns"])), shape=(n, n))
    return (A + sp.diags(level["diagonal"])).tocsr()


def dense_fine_operator(g):
    """The fluid part of A from the face list alone (grid.rs faces and weights)."""
    f = g.faces()
    fluid_cells = np.nonzero(g.fluid())[0]
    where = -np.ones(g.count, np.int64)
    where[fluid_cells] = np.arange(fluid_cells.size)
    rows, cols, vals = [], [], []
    for lo, hi, kind, w in zip(f["lo"], f["hi"], f["kind"], f["weight"]):
        if kind == 0 and lo != hi:  # interior
            a, b = where[lo], where[hi]
            rows += [a, b, a, b]
            cols += [a, b, b, a]
            vals += [w, w, -w, -w]
        elif kind == 4:  # outlet
            a = where[lo if lo >= 0 else hi]
            rows.append(a)
            cols.append(a)
            vals.append(w)
    n = fluid_cells.size
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n)).tocsr(), fluid_cells


def graph_components(n, pairs):
    """Connected components of an undirected graph given as (i, j) pairs: labels
    numbered by lowest member (union-find, independent of any BFS order)."""
    parent = np.arange(n)

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    for i, j in pairs:
        ri, rj = find(i), find(j)
        if ri != rj:
            parent[max(ri, rj)] = min(ri, rj)
    roots = np.array([find(i) for i in range(n)])
    order = {}
    labels = np.empty(n, np.int64)
    for i in range(n):
        labels[i] = order.setdefault(roots[i], len(order))
    return labels


def edges_of(level):
    n = level["n"]
    rows = np.repeat(np.arange(n), np.diff(level["offsets"]))
    cols = level["columns"]
    keep = rows < cols
    return list(zip(rows[keep].tolist(), cols[keep].tolist()))


def anchors_of(level):
    """Outlet weight of each unknown: the diagonal less its couplings (values are -w)."""
    n = level["n"]
    rows = np.repeat(np.arange(n), np.diff(level["offsets"]))
    off = np.zeros(n)
    np.add.at(off, rows, -level["values"])
    return level["diagonal"] - off


def walk(case):
    """(oracle grid, levels, per-level coordinates of every unknown in fine cells,
    per-level set of fine cells) for a case."""
    g = oracle.Grid(case)
    levels = g.masked_levels()
    n = case.cells
    cells = levels[0]["fluid"]
    coords = np.stack([cells % n[0], cells // n[0] % n[1], cells // (n[0] * n[1])], axis=1)
    # members[l][k]: fine (level-0) unknowns of node k of level l
    owner = np.arange(levels[0]["n"])
    owners = [owner]
    for lv in levels[:-1]:
        owner = lv["parent"][owner]
        owners.append(owner)
    return g, levels, coords, owners


# ---------------------------------------------------------------- tests
@pytest.mark.parametrize("name,case", FIXTURES, ids=IDS)
def test_level0_is_the_fluid_operator(name, case):
    g, levels, _, _ = walk(case)
    A, fluid_cells = dense_fine_operator(g)
    assert np.array_equal(levels[0]["fluid"], fluid_cells)
    D = csr_matrix(levels[0])
    assert abs(A - D).max() <= 1e-12 * abs(A).max()


@pytest.mark.parametrize("name,case", FIXTURES, ids=IDS)
def test_aggregates_are_the_connected_pieces_of_minimal_bins(name, case):
    """Every aggregate is connected (masked.rs:389-404); the partition is exactly
    the connected pieces of the level graph inside bins of the smallest size that
    reduces the level (the bins double per level and grow while nothing merges,
    masked.rs:246-257); aggregates are numbered by their lowest member."""
    g, levels, coords, owners = walk(case)
    shift = 0  # coordinates of level l are fine coordinates >> shift
    for l, lv in enumerate(levels[:-1]):
        n, parent = lv["n"], lv["parent"]
        m = levels[l + 1]["n"]
        assert parent.min() == 0 and parent.max() == m - 1 and m < n
        # every fine cell of a node shares the node's coordinates at this level's spacing
        node_coord = np.zeros((n, 3), np.int64)
        node_coord[owners[l]] = coords >> shift
        assert np.array_equal(node_coord[owners[l]], coords >> shift)
        pairs = edges_of(lv)
        pa, pb = (np.array([p[i] for p in pairs], np.int64) for i in (0, 1))
        found = None
        for grow in range(1, 40):
            u = shift + grow
            same = np.all((node_coord[pa] >> grow) == (node_coord[pb] >> grow), axis=1) if pairs else np.zeros(0, bool)
            labels = graph_components(n, [(int(a), int(b)) for a, b, s in zip(pa, pb, same) if s])
            if labels.max() + 1 < n:
                found = (u, labels)
                break
        assert found is not None, name
        u, labels = found
        assert np.array_equal(parent, labels), (name, l)
        # numbered by lowest member: first occurrences in index order are 0, 1, 2, ...
        _, first = np.unique(parent, return_index=True)
        assert np.all(np.diff(first) > 0)
        shift = u


@pytest.mark.parametrize("name,case", FIXTURES, ids=IDS)
def test_coarse_operator_is_the_galerkin_product(name, case):
    """A_c = P^T A P with the piecewise-constant P, densely (masked.rs:367-387):
    couplings between distinct aggregates summed, those inside one aggregate
    dropped, outlet weights summed into the aggregate's diagonal."""
    g, levels, _, _ = walk(case)
    for l, lv in enumerate(levels[:-1]):
        n, m = lv["n"], levels[l + 1]["n"]
        P = sp.csr_matrix((np.ones(n), (np.arange(n), lv["parent"])), shape=(n, m))
        want = (P.T @ csr_matrix(lv) @ P).tocsr()
        got = csr_matrix(levels[l + 1])
        scale = abs(want).max()
        assert abs(want - got).max() <= 1e-12 * scale, (name, l)
        # the stored entries are exactly the distinct coupled pairs, none on the diagonal
        rows = np.repeat(np.arange(m), np.diff(levels[l + 1]["offsets"]))
        assert np.all(rows != levels[l + 1]["columns"])
        np.testing.assert_allclose(anchors_of(levels[l + 1]), P.T @ anchors_of(lv), rtol=1e-12, atol=1e-12 * scale)


@pytest.mark.parametrize("name,case", FIXTURES, ids=IDS)
def test_components_closed_lists_and_stop_rule(name, case):
    """Per level: the closed lists are the components without an outlet weight,
    each starting at its lowest unknown (breadth-first from there, masked.rs:90-126),
    their number is preserved from level to level and each coarse closed
    component is the image of a fine one (masked.rs:388); the hierarchy stops at
    the first level whose components all have at most 32 unknowns and solves it
    with one dense factor per component (masked.rs:233-239)."""
    g, levels, _, owners = walk(case)
    comp, anch = g.components()
    fine_closed = sorted(tuple(np.nonzero(comp == c)[0]) for c in range(len(anch)) if not anch[c])
    for l, lv in enumerate(levels):
        labels = graph_components(lv["n"], edges_of(lv))
        anchored = np.zeros(labels.max() + 1, bool)
        np.logical_or.at(anchored, labels, anchors_of(lv) > 0.0)
        want = [np.nonzero(labels == c)[0] for c in range(len(anchored)) if not anchored[c]]
        got = lv["closed"]
        assert len(got) == len(want) == len(fine_closed), (name, l)
        for a, b in zip(got, want):
            assert a[0] == b[0] and np.array_equal(np.sort(a), b)
        # image of the grid's closed components in fine cells
        cells_of_node = {}
        for f, node in enumerate(owners[l]):
            cells_of_node.setdefault(int(node), []).append(int(levels[0]["fluid"][f]))
        images = sorted(tuple(sorted(c for k in members for c in cells_of_node[int(k)])) for members in got)
        assert images == fine_closed
        sizes = np.bincount(labels)
        terminal = l == len(levels) - 1
        assert (sizes.max() <= COARSE_CELLS) == terminal, (name, l, sizes.max())
        if terminal:
            assert lv["parent"] is None and lv["blocks"] == len(sizes)


# ---------------------------------------------------------------- masked.rs:346-573 restated
def test_reference_barrier_grid_facts():
    g, levels, _, _ = walk(barrier())
    assert len(levels) > 1
    comp, anch = g.components()
    assert list(anch) == [False, False]


def test_reference_topology_facts():
    """The component topologies of masked.rs:509-573."""
    g = oracle.Grid(thin_passage())
    assert list(g.components()[1]) == [True]
    g = oracle.Grid(sealed_side())
    comp, anch = g.components()
    n = [9, 3, 6]
    pocket = POCKET[0] + n[0] * (POCKET[1] + n[1] * POCKET[2])
    regions = [np.nonzero(comp == c)[0] for c in range(len(anch))]
    assert any(len(r) == 1 and r[0] == pocket for r in regions)
    assert any(len(r) == 1 and a for r, a in zip(regions, anch))
    assert not anch.all()
    g = oracle.Grid(periodic_wrap())
    comp, anch = g.components()
    assert list(anch) == [False, False]
    assert any(comp[0] == c and comp[8] == c for c in range(2))
    g = oracle.Grid(porous_periodic())
    comp, anch = g.components()
    sizes = np.bincount(comp[comp >= 0])
    assert (sizes == 1).any() and (sizes > COARSE_CELLS).any() and not anch.any()


def gauged_cycle(g):
    """masked.rs:405-436: the cycle on the gauged fluid space (every unknown of an
    anchored component, differences to the first cell of a sealed one)."""
    comp, anch = g.components()
    fluid = g.fluid()
    basis = []
    for c in range(len(anch)):
        region = np.nonzero(comp == c)[0]
        for i in region[(0 if anch[c] else 1):]:
            v = np.zeros(g.count)
            v[i] = 1.0
            if not anch[c]:
                v[region[0]] = -1.0
            basis.append(v)
    B = np.array(basis)
    M = np.zeros((len(basis), len(basis)))
    for j, b in enumerate(basis):
        y = g.mg_apply(b)
        assert np.all(np.isfinite(y))
        assert np.all(y[~fluid] == 0.0)
        for c in range(len(anch)):
            if not anch[c]:
                assert abs(y[comp == c].mean()) < 1e-12
        M[:, j] = B @ y
    return M


@pytest.mark.parametrize("make", [thin_passage, sealed_side, periodic_wrap, porous_periodic, barrier, courtyards],
                         ids=["thin-passage", "sealed-side", "periodic-wrap", "porous-periodic", "barrier", "courtyards"])
def test_masked_cycle_is_spd_for_every_component_topology(make):
    g = oracle.Grid(make())
    M = gauged_cycle(g)
    scale = np.abs(M).max()
    assert np.max(np.abs(M - M.T)) < 1e-11 * scale
    np.linalg.cholesky(0.5 * (M + M.T))  # raises unless positive definite
    assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() > 1e-12 * scale
