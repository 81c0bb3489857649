"""The grid context from a caller's built Grid on the device (cfd_grid_create_desc,
cfd_grid_matches_desc): a descriptor grid is the case's grid, and the
reference's weight-edit test is restated through the C-ABI."""
import numpy as np
import pytest

import cfdgpu
import oracle
from cfdgpu import workloads as W
from cfdgpu.case import IC0, MULTIGRID

from test_grid_descriptor import EUNSUPPORTED, descriptor, first, outlet_case

pytestmark = pytest.mark.gpu

CASES = [
    ("outlet-cavity", lambda: outlet_case((9, 2, 7))),
    ("urban", lambda: W.urban_block(n=(48, 40, 24), boxes=6)[0]),
    ("cavity-ic0", lambda: W.cavity([24, 2, 24], IC0)),
    ("courtyards", None),
]


@pytest.mark.parametrize("name,make", CASES, ids=[n for n, _ in CASES])
def test_descriptor_grid_is_the_case_grid(gpu_available, name, make):
    """Built from the Grid's own faces, the context has the case grid's codes and
    components, applies the operator bit for bit as the oracle does, and solves
    with the case grid's iterations and bits; it matches its own descriptor."""
    if make is None:
        from test_gpu_components import large_courtyards

        case = large_courtyards()
    else:
        case = make()
    d = descriptor(case)
    d.preconditioner = case.preconditioner
    print(f"[descriptor] {name}", flush=True)
    gd = cfdgpu.Grid.from_descriptor(d)
    gc = cfdgpu.Grid(case)
    o = oracle.Grid(case)
    assert np.array_equal(gd.codes(), gc.codes())
    for a, b in zip(gd.components(), gc.components()):
        assert np.array_equal(a, b)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(case.count)
    assert np.array_equal(gd.apply(x), o.apply(x))
    ex = o.center_nullspaces(np.where(o.fluid(), rng.standard_normal(o.count), 0.0))
    b = o.apply(ex)
    xd, infod = gd.solve(b, 1e-11, 5000)
    xc, infoc = gc.solve(b, 1e-11, 5000)
    assert infod["iterations"] == infoc["iterations"], (infod, infoc)
    assert np.array_equal(xd, xc)
    assert gd.matches_descriptor(d)
    assert gd.matches(case)


def test_reused_hierarchies_follow_weight_edits(gpu_available):
    """tests/multigrid.rs:235-264 through the C-ABI: on the 9x2x7 outlet cavity
    the signature is kept by a repeated solve and by an edited wall face (walls
    carry no pressure weight), and lost by a few-ulp interior distance edit and
    by a doubled outlet area; the reference rebuilds its hierarchy there (3
    builds in all). Here the same decisions are made by cfd_grid_matches_desc;
    a rebuild from an edited interior or outlet face is refused with
    CFD_EUNSUPPORTED (one weight per axis and face kind on the device)."""
    d = descriptor(outlet_case())
    g = cfdgpu.Grid.from_descriptor(d)
    exact = np.sin(0.37 * np.arange(g.count))
    control = cfdgpu.Control(profile=True)

    def solve(grid):
        b = grid.apply(exact)
        _, info = grid.solve(b, 1e-12, 2000, control=control)
        assert info["relative_residual"] <= 1.01e-11
        return info

    builds = 0
    builds += solve(g)["hierarchy_builds"]  # build
    assert g.matches_descriptor(d)
    builds += solve(g)["hierarchy_builds"]  # reuse
    d.faces["area"][first(d, cfdgpu.FACE_WALL)] *= 2.0
    assert g.matches_descriptor(d)  # walls carry no pressure weight
    builds += solve(g)["hierarchy_builds"]  # reuse
    d.faces["distance"][first(d, cfdgpu.FACE_INTERIOR)] *= 1.0 + 1e-15
    decisions = [g.matches_descriptor(d)]  # a few-ulp weight change: rebuild
    with pytest.raises(cfdgpu.CfdError) as e:
        cfdgpu.Grid.from_descriptor(d)
    assert e.value.code == EUNSUPPORTED
    d.faces["area"][first(d, cfdgpu.FACE_OUTLET)] *= 2.0
    decisions.append(g.matches_descriptor(d))  # rebuild
    assert decisions == [False, False]
    assert builds == 1
    assert builds + decisions.count(False) == 3  # the reference's count
