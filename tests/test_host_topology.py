"""The product's host set-up (cfd_case_topology, no device) against the oracle's
Grid::build (grid.rs:246-464): per-cell coupling codes, components, anchoring and
the IC(0) factor diagonal (pressure.rs:191-214) must be identical, bit for bit.
Also checks that libcfdgpu.so exports every entry point include/cfdgpu.h declares,
and that the numpy workload helpers classify faces exactly as the reference does.
"""
import ctypes
import json
import os
import re

import numpy as np
import pytest

import oracle
import cfdgpu
from cfdgpu import workloads as W
from cfdgpu.case import INTERIOR, OPEN, OUTLET, PERIODIC, WALL, Block, Case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def codes_from_oracle(g):
    f = g.faces()
    codes = np.zeros(g.count, np.uint8)
    codes[g.fluid()] |= 64
    for ax, lo, hi, kind in zip(f["axis"], f["lo"], f["hi"], f["kind"]):
        if kind == INTERIOR and lo != hi:
            codes[lo] |= 1 << (2 * ax + 1)
            codes[hi] |= 1 << (2 * ax)
        elif kind == OUTLET:
            if lo >= 0:
                codes[lo] |= 1 << (2 * ax + 1)
            else:
                codes[hi] |= 1 << (2 * ax)
    comp, anch = g.components()
    for c in range(len(anch)):
        if not anch[c]:
            codes[comp == c] |= 128
    return codes


def cases():
    out = [("cavity24", W.cavity([24, 2, 24])), ("cube12", W.cavity([12, 12, 12]))]
    for n in ([9, 2, 7], [8, 5, 6], [2, 2, 2], [9, 1, 7], [6, 4, 5]):
        for bnd in (WALL, PERIODIC, OPEN):
            out.append((f"lid{n}-{bnd}", Case(cells=n, size_m=[2.7, 1.9, 3.1], boundaries=[bnd, PERIODIC, WALL],
                                              velocity_m_s=[1.0, 0.0, 0.0])))
    with open(os.path.join(ROOT, "tests", "golden", "canyon-93k.json")) as f:
        out.append(("canyon93k", Case.from_json(json.load(f))))
    out.append(("urban32", W.urban_block(n=(32, 32, 16), boxes=6)[0]))
    out.append(("prisms48", W.urban_prisms(n=(48, 48, 24), prisms=12, radius=(3.0, 6.0), heights=(3, 12))[0]))
    out.append(("canyon-streets", W.street_canyon(200.0, n=(72, 72, 30), lots=6, street=2.0, heights=(5, 28))[0]))
    out.append(("barrier", Case(cells=[9, 2, 5], size_m=[9, 2, 5], boundaries=[OPEN, WALL, WALL], velocity_m_s=[1.0, 0, 0],
                                blocks=[Block([3.0, 0, 0], [4.0, 2, 5])])))
    out.append(("rotated-wind", Case(cells=[20, 16, 8], size_m=[20, 16, 8], boundaries=[OPEN, OPEN, 4], flow_kind=1,
                                     speed_m_s=3.0, reference_height_m=5.0, roughness_m=0.05, direction_deg_from=225.0,
                                     blocks=[Block([5.0, 4, 0], [9.0, 9, 5])])))
    return out


@pytest.mark.parametrize("name,case", cases(), ids=[n for n, _ in cases()])
def test_topology_matches_oracle(name, case):
    o = oracle.Grid(case)
    t = cfdgpu.case_topology(case, ic0=True)
    assert np.array_equal(t["codes"], codes_from_oracle(o))
    comp, anch = o.components()
    assert np.array_equal(t["component"], comp)
    assert np.array_equal(t["anchored"], anch)
    assert list(t["counts"][:5]) == [o.count, o.face_count, o.fluid_count, o.n_components, o.n_anchored]
    assert np.array_equal(t["ic0_diagonal"].view(np.uint64), o.ic0_diagonal().view(np.uint64))


def fill_cases():
    """Grids with a periodic axis of exactly 3 cells, the topology where two lower
    neighbours of a cell are joined themselves and IC(0) entries take the dot
    term of Factor::build (pressure.rs:197-209)."""
    return [
        ("fill-x", Case(cells=[3, 8, 6], size_m=[0.3, 0.8, 0.6], boundaries=[PERIODIC, WALL, WALL])),
        ("fill-y", Case(cells=[7, 3, 5], size_m=[0.7, 0.3, 0.5], boundaries=[OPEN, PERIODIC, WALL],
                        velocity_m_s=[1.0, 0.0, 0.0])),
        ("fill-z", Case(cells=[6, 5, 3], size_m=[0.6, 0.5, 0.3], boundaries=[WALL, WALL, PERIODIC],
                        blocks=[Block([0.2, 0.2, 0.0], [0.4, 0.3, 0.1])])),
        ("fill-xyz", Case(cells=[3, 3, 3], size_m=[0.3, 0.3, 0.3], boundaries=[PERIODIC, PERIODIC, PERIODIC])),
        ("fill-3-2", Case(cells=[3, 2, 9], size_m=[0.3, 0.2, 0.9], boundaries=[PERIODIC, PERIODIC, OPEN],
                          velocity_m_s=[0.0, 0.0, 1.0])),
        ("fill-block", Case(cells=[9, 3, 8], size_m=[0.9, 0.3, 0.8], boundaries=[OPEN, PERIODIC, WALL],
                            velocity_m_s=[1.0, 0.0, 0.0], blocks=[Block([0.3, 0.0, 0.2], [0.5, 0.1, 0.5])])),
    ]


@pytest.mark.parametrize("name,case", fill_cases(), ids=[n for n, _ in fill_cases()])
def test_ic0_fill_diagonal_matches_oracle(name, case):
    """The host factor with the dot terms gives the oracle's IC(0) diagonal bit for bit."""
    o = oracle.Grid(case)
    t = cfdgpu.case_topology(case, ic0=True)
    assert np.array_equal(t["ic0_diagonal"].view(np.uint64), o.ic0_diagonal().view(np.uint64))


@pytest.mark.parametrize("name,case", cases(), ids=[n for n, _ in cases()])
def test_numpy_workload_helpers_match_oracle(name, case):
    o = oracle.Grid(case)
    fluid = W.fluid_mask(case)
    assert np.array_equal(fluid, o.fluid())
    kinds, zc, offs = W.face_kinds(case, fluid)
    assert np.array_equal(kinds, o.faces()["kind"])
    u = np.random.default_rng(0).standard_normal(o.face_count)
    assert np.array_equal(W.enforce_velocity(case, kinds, zc, offs, u), o.enforce_velocity(u))
    b1, t1 = o.projection_rhs(u, 0.5)
    b2, t2 = W.projection_rhs(case, fluid, u, 0.5)
    assert np.array_equal(b1, b2) and np.array_equal(t1, t2)


@pytest.mark.parametrize("name,case", cases(), ids=[n for n, _ in cases()])
def test_reference_hierarchy_matches_oracle(name, case):
    """The hierarchy of the reference cycle, built by the library on the host
    (masked aggregation, masked.rs:181-291, or the unmasked trilinear Galerkin
    hierarchy, multigrid.rs:530-645): unknowns and stored entries per level equal
    the oracle's, which the published hierarchy sizes pin."""
    case.preconditioner = 1
    o = oracle.Grid(case)
    try:
        ref = o.mg_stats()
    except oracle.OracleError as e:
        pytest.skip(f"oracle: {e}")
    got = cfdgpu.case_multigrid_stats(case, cfdgpu.CYCLE_REFERENCE)
    assert got["level_cells"] == ref["level_cells"]
    assert got["level_nonzeros"] == ref["level_nonzeros"]


def _aniso_cases():
    from test_oracle_goldens import _aniso, aniso_case

    return [(k, aniso_case(v, "mg")) for k, v in _aniso()["shapes"].items()]


@pytest.mark.parametrize("name,case", _aniso_cases(), ids=[n for n, _ in _aniso_cases()])
def test_anisotropic_hierarchy_matches_oracle(name, case):
    """The 18 anisotropy shapes: semicoarsened levels (weak axes wait until the
    stiff ones collapse, multigrid.rs:582-611) with the oracle's sizes and entries."""
    o = oracle.Grid(case)
    ref = o.mg_stats()
    got = cfdgpu.case_multigrid_stats(case, cfdgpu.CYCLE_REFERENCE)
    assert got["level_cells"] == ref["level_cells"]
    assert got["level_nonzeros"] == ref["level_nonzeros"]
    # MultigridStats shape: level dimensions and block smoothers (line.rs:200-209)
    shape = o.mg_shape()
    assert got.get("level_dimensions", []) == shape["level_dimensions"]
    blocks = [{k: v for k, v in b.items() if k != "allocated_bytes"} for b in got.get("block_smoothers", [])]
    assert blocks == shape["block_smoothers"]


def _lid(n, size):
    return Case(cells=list(n), size_m=list(size), boundaries=[WALL, PERIODIC, WALL], velocity_m_s=[1.0, 0.0, 0.0],
                preconditioner=1)


def test_block_smoother_stats_of_the_reference_tests():
    """tests/multigrid.rs:297-377: the 12x4x12 cavity of size [1, 1, 0.1] takes z
    lines on every smoothed level; a stiff two-cell y axis (16x2x16, y 0.001 m)
    takes two-unknown y blocks; 8x3x8 takes blocks only once dz shrinks 100-fold.
    multigrid.rs:940-960: while a stiff axis still coarsens, the weak axes keep
    their dimensions."""
    s = cfdgpu.case_multigrid_stats(_lid([12, 4, 12], [1.0, 1.0, 0.1]), cfdgpu.CYCLE_REFERENCE)
    assert s["block_smoothers"] and all(b["axes"] == [2] for b in s["block_smoothers"])
    dims = s["level_dimensions"]
    for n, nxt in zip(dims, dims[1:]):
        if n[2] > 2:
            assert (n[0], n[1]) == (nxt[0], nxt[1])
    s = cfdgpu.case_multigrid_stats(_lid([16, 2, 16], [1.0, 0.001, 1.0]), cfdgpu.CYCLE_REFERENCE)
    assert s["block_smoothers"]
    assert all(b["axes"] == [1] and b["block_unknowns"] == 2 for b in s["block_smoothers"])
    assert "block_smoothers" not in cfdgpu.case_multigrid_stats(_lid([8, 3, 8], [1.0, 1.0, 1.0]), cfdgpu.CYCLE_REFERENCE)
    assert cfdgpu.case_multigrid_stats(_lid([8, 3, 8], [1.0, 1.0, 0.01]), cfdgpu.CYCLE_REFERENCE)["block_smoothers"]
    # the aggregation hierarchy has irregular levels: no dimensions (masked.rs:262)
    s = cfdgpu.case_multigrid_stats(_lid([12, 4, 12], [1.0, 1.0, 0.1]), cfdgpu.CYCLE_RB)
    assert "level_dimensions" not in s and "block_smoothers" not in s


def test_reference_hierarchy_published_sizes():
    """48x2x48 cavity: 5 levels, 6138 cells and 78084 stored entries in all
    (pressure-multigrid.md:141-142); operator complexity of cube48 2.601."""
    s = cfdgpu.case_multigrid_stats(W.cavity([48, 2, 48]), cfdgpu.CYCLE_REFERENCE)
    assert s["level_cells"] == [4608, 1152, 288, 72, 18]
    assert sum(s["level_cells"]) == 6138 and sum(s["level_nonzeros"]) == 78084
    s = cfdgpu.case_multigrid_stats(W.cavity([48, 48, 48]), cfdgpu.CYCLE_REFERENCE)
    assert round(sum(s["level_nonzeros"]) / s["level_nonzeros"][0], 3) == 2.601
    assert round(sum(s["level_cells"]) / s["level_cells"][0], 3) == 1.143


def test_grid_build_errors_match_reference_codes():
    with pytest.raises(cfdgpu.CfdError, match="CFD-E028"):
        cfdgpu.case_topology(Case(cells=[8, 8, 8], size_m=[8, 8, 8], boundaries=[WALL] * 3,
                                  blocks=[Block([1.1, 1.1, 1.1], [1.2, 1.2, 1.2])]))
    with pytest.raises(cfdgpu.CfdError, match="CFD-E030"):
        cfdgpu.case_topology(Case(cells=[4, 4, 4], size_m=[4, 4, 4], boundaries=[WALL] * 3,
                                  blocks=[Block([0.0, 0, 0], [4.0, 4, 4])]))
    with pytest.raises(cfdgpu.CfdError, match="CFD-E004"):
        cfdgpu.case_topology(Case(cells=[1, 4, 4], size_m=[4, 4, 4], boundaries=[WALL] * 3))


def test_library_exports_every_declared_entry_point():
    with open(os.path.join(ROOT, "include", "cfdgpu.h")) as f:
        header = f.read()
    names = set(re.findall(r"^\s*(?:int|void|const char\*)\s+(cfd_\w+)\s*\(", header, re.M))
    assert len(names) >= 25
    lib = ctypes.CDLL(cfdgpu.LIBRARY)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_case_json_keys():
    with open(os.path.join(ROOT, "tests", "golden", "lid-cavity.json")) as f:
        c = json.load(f)
    case = Case.from_json(c)
    assert case.preconditioner == cfdgpu.IC0 and case.tolerance == 1e-10 and case.max_iterations == 1200
    c["numerics"]["pressure_preconditioner"] = "multigrid"
    assert Case.from_json(c).preconditioner == cfdgpu.MULTIGRID


def test_street_canyon_city_is_the_same_for_every_direction():
    """The campaign's city (lots x lots convex prisms between streets) depends on
    the seed only; the wind direction changes the open faces' kinds and the
    right-hand side."""
    a, ba, _ = W.street_canyon(0.0, n=(96, 96, 40), lots=5, heights=(8, 36))
    b, bb, _ = W.street_canyon(90.0, n=(96, 96, 40), lots=5, heights=(8, 36))
    assert len(a.blocks) == len(b.blocks) == 25
    for p, q in zip(a.blocks, b.blocks):
        assert p.footprint == q.footprint and p.max_m == q.max_m
        xs = [v[0] for v in p.footprint]
        ys = [v[1] for v in p.footprint]
        assert 4.0 <= min(xs) and max(xs) <= 92.0 and 4.0 <= min(ys) and max(ys) <= 92.0
        assert 4 <= len(p.footprint) <= 8
    fluid = W.fluid_mask(a)
    assert np.array_equal(fluid, W.fluid_mask(b))
    assert 0.1 < 1.0 - fluid.mean() < 0.4
    assert not np.array_equal(ba, bb)



def test_red_black_hierarchy_semicoarsens_weak_axes():
    """The red-black cycle's aggregation bins an axis only while its coupling is
    within 25x of the strongest axis that still spans several nodes: a lid cavity
    of 2 or 4 cells over the unit y extent keeps its y layers apart (its y
    couplings are 2300x weaker), a 100:1 stiff z axis is binned alone first;
    cubic cells bin 2x2x2 as the reference's masked aggregation does."""
    from cfdgpu.case import MULTIGRID, PERIODIC

    rb = lambda case: cfdgpu.case_multigrid_stats(case, cfdgpu.CYCLE_RB)["level_cells"]
    assert rb(W.cavity([48, 2, 48], MULTIGRID)) == [4608, 1152]
    assert rb(W.cavity([96, 4, 96], MULTIGRID)) == [36864, 9216, 2304, 576]
    assert rb(W.cavity([48, 48, 48], MULTIGRID)) == [110592, 13824, 1728]
    aniso = Case(cells=[96, 16, 96], size_m=[1, 1, 0.1], boundaries=[WALL, PERIODIC, WALL], preconditioner=MULTIGRID)
    assert rb(aniso)[:3] == [147456, 73728, 36864]
    # the reference cycle keeps the reference's hierarchy
    assert cfdgpu.case_multigrid_stats(W.cavity([48, 2, 48], MULTIGRID), cfdgpu.CYCLE_REFERENCE)["level_cells"] == \
        [4608, 1152, 288, 72, 18]
