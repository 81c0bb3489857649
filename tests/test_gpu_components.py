"""GPU parity on grids with several fluid components: sealed rooms and
courtyards (closed components, one gauge each), anchored singletons, sealed
singletons without an unknown, and components that close through a periodic
wrap (crates/cfd-craft/src/multigrid/masked.rs:509-573; pressure.rs:316-325
for the per-component gauges, 351-362 for CFD-E031 per component).

Each grid is solved by IC(0), the reference's multigrid cycle and the
red-black cycle in FP64 and mixed precision (with the dense coarsest level and
with coarse levels forced), and IC(0) in FP32 inside the FP64 PCG, against the oracle at tol 1e-13: fields to 1e-10
relative, IC(0) and reference-cycle iteration counts within +-1.
"""
import os
from contextlib import contextmanager

import numpy as np
import pytest

import cfdgpu
import oracle
from cfdgpu.case import IC0, MULTIGRID, OPEN, WALL, Block, Case
from test_oracle_masked import barrier, courtyards, periodic_wrap, porous_periodic, sealed_side, thin_passage

pytestmark = pytest.mark.gpu


@contextmanager
def env(**kv):
    old = {k: os.environ.get(k) for k in kv}
    os.environ.update({k: str(v) for k, v in kv.items()})
    try:
        yield
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def large_courtyards():
    """Open x street with three sealed courtyards of 2.9k to 8.6k cells (more
    than one 2048-member chunk each): full-height rings of wall."""
    blocks = []
    for x0, y0, w in ((6, 6, 22), (34, 8, 18), (8, 34, 16)):
        blocks += [Block([x0, y0, 0.0], [x0 + w, y0 + 1.0, 24.0]), Block([x0, y0 + w - 1.0, 0.0], [x0 + w, y0 + w, 24.0]),
                   Block([x0, y0, 0.0], [x0 + 1.0, y0 + w, 24.0]), Block([x0 + w - 1.0, y0, 0.0], [x0 + w, y0 + w, 24.0])]
    blocks = [Block([float(v) for v in b.min_m], [float(v) for v in b.max_m]) for b in blocks]
    return Case(cells=[64, 56, 24], size_m=[64.0, 56.0, 24.0], boundaries=[OPEN, WALL, WALL], velocity_m_s=[1.0, 0, 0],
                blocks=blocks, preconditioner=MULTIGRID)


GRIDS = [("barrier", barrier), ("thin-passage", thin_passage), ("sealed-side", sealed_side),
         ("periodic-wrap", periodic_wrap), ("porous-periodic", porous_periodic), ("courtyards", courtyards),
         ("large-courtyards", large_courtyards)]
IDS = [n for n, _ in GRIDS]


def test_fixtures_have_several_closed_components():
    closed = {}
    for name, make in GRIDS:
        comp, anch = oracle.Grid(make()).components()
        closed[name] = int((~anch).sum())
    assert closed["barrier"] == 2 and closed["periodic-wrap"] == 2
    assert closed["courtyards"] >= 3 and closed["large-courtyards"] >= 3 and closed["porous-periodic"] >= 3


def _exact(o, seed):
    rng = np.random.default_rng(seed)
    return o.center_nullspaces(np.where(o.fluid(), rng.standard_normal(o.count), 0.0))


@pytest.mark.parametrize("name,make", GRIDS, ids=IDS)
def test_center_nullspaces_per_component(gpu_available, name, make):
    case = make()
    o = oracle.Grid(case)
    g = cfdgpu.Grid(case)
    x = np.random.default_rng(3).standard_normal(o.count) + 5.0
    xo = o.center_nullspaces(x)
    xg = g.center_nullspaces(x)
    assert np.max(np.abs(xg - xo)) <= 1e-13 * np.max(np.abs(x)), name
    comp, anch = o.components()
    for c in range(len(anch)):
        if not anch[c]:
            assert abs(xg[comp == c].mean()) <= 1e-13 * np.max(np.abs(x))


SOLVERS = [("ic0", IC0, None, None, None), ("ic0-mixed", IC0, None, cfdgpu.PRECISION_MIXED, None), ("reference", MULTIGRID, cfdgpu.CYCLE_REFERENCE, cfdgpu.PRECISION_FP64, None),
           ("rb-fp64", MULTIGRID, cfdgpu.CYCLE_RB, cfdgpu.PRECISION_FP64, None),
           ("rb-mixed", MULTIGRID, cfdgpu.CYCLE_RB, cfdgpu.PRECISION_MIXED, None),
           ("rb-fp64-levels", MULTIGRID, cfdgpu.CYCLE_RB, cfdgpu.PRECISION_FP64, 8),
           ("rb-mixed-levels", MULTIGRID, cfdgpu.CYCLE_RB, cfdgpu.PRECISION_MIXED, 8)]


@pytest.mark.parametrize("solver", SOLVERS, ids=[s[0] for s in SOLVERS])
@pytest.mark.parametrize("name,make", GRIDS, ids=IDS)
def test_solve_matches_oracle(gpu_available, name, make, solver):
    label, pc, cycle, precision, dense = solver
    case = make()
    case.preconditioner = pc
    o = oracle.Grid(case)
    ex = _exact(o, 17)
    b = o.apply(ex)
    xo, io = o.solve(b, 1e-13, 5000)
    with env(CFD_MG_DENSE=dense or 2048):
        g = cfdgpu.Grid(case)
    if precision is not None:
        g.set_options(precision, cycle=cycle or 0)
    print(f"[components] {name} {label}", flush=True)
    x, info = g.solve(b, 1e-13, 5000)
    assert info["relative_residual"] <= 1.01e-12, (name, label, info)
    assert rel(x, xo) <= 1e-10, (name, label, rel(x, xo), info["iterations"], io["iterations"])
    if label in ("ic0", "reference"):
        assert abs(info["iterations"] - io["iterations"]) <= 1, (info["iterations"], io["iterations"])
    comp, anch = o.components()
    for c in range(len(anch)):
        if not anch[c]:
            assert abs(x[comp == c].sum()) <= 1e-12 * np.abs(x).sum()
    assert np.all(x[~o.fluid()] == 0.0)


@pytest.mark.parametrize("pc", [IC0, MULTIGRID], ids=["ic0", "multigrid"])
@pytest.mark.parametrize("name,make", [g for g in GRIDS if g[0] in ("barrier", "courtyards", "large-courtyards")],
                         ids=["barrier", "courtyards", "large-courtyards"])
def test_incompatible_flux_in_one_component_is_e031(gpu_available, name, make, pc):
    """The compatibility check runs per closed component (pressure.rs:351-362):
    flux added to one sealed component only is CFD-E031 on the device as in the
    oracle; a compatible right-hand side with round-off sized offsets is accepted."""
    case = make()
    case.preconditioner = pc
    o = oracle.Grid(case)
    g = cfdgpu.Grid(case)
    b = o.apply(_exact(o, 5))
    comp, anch = o.components()
    closed = [c for c in range(len(anch)) if not anch[c]]
    b_ok = b + np.where(comp == closed[0], 1e-13, 0.0)
    o.solve(b_ok, 1e-10, 5000)
    g.solve(b_ok, 1e-10, 5000)
    for c in closed:
        bad = b + np.where(comp == c, 0.1, 0.0)
        with pytest.raises(oracle.OracleError) as eo:
            o.solve(bad, 1e-10, 5000)
        with pytest.raises(cfdgpu.CfdError) as eg:
            g.solve(bad, 1e-10, 5000)
        assert eo.value.code == 31 and eg.value.code == 31 and "CFD-E031" in str(eg.value)


@pytest.mark.parametrize("name,make", GRIDS, ids=IDS)
def test_reference_vcycle_matches_oracle(gpu_available, name, make):
    """One V-cycle of the reference's masked cycle (per-component centering on
    every level, sealed singletons left at zero) equals the oracle's serial one
    up to the order of the centering sums."""
    case = make()
    o = oracle.Grid(case)
    r = o.center_nullspaces(np.where(o.fluid(), np.random.default_rng(23).standard_normal(o.count), 0.0))
    zo = o.mg_apply(r)
    g = cfdgpu.Grid(case)
    g.set_options(cfdgpu.PRECISION_FP64, cycle=cfdgpu.CYCLE_REFERENCE)
    z = g.precondition(r)
    assert rel(z, zo) <= 1e-12, (name, rel(z, zo))
