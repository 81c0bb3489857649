"""cfd_project_static, the projection that allocates nothing on the device,
against cfd_project through the C-ABI: the same bits for u, p and the
iteration count, on two calls in a row (the second on a smaller grid, which
reuses the front of the static staging array), and a clear refusal for a grid
beyond the array's size."""
import numpy as np
import pytest

import cfdgpu
from cfdgpu import workloads as W
from cfdgpu.case import IC0

pytestmark = pytest.mark.gpu


def _inputs(case, seed):
    fluid = W.fluid_mask(case)
    kinds, zc, offs = W.face_kinds(case, fluid)
    rng = np.random.default_rng(seed)
    u = W.enforce_velocity(case, kinds, zc, offs, rng.standard_normal(kinds.size))
    p = rng.standard_normal(case.count)
    return u, p


@pytest.mark.parametrize("pc", ["multigrid", "ic0"])
def test_static_projection_equals_project(gpu_available, pc):
    for step, (n, boxes) in enumerate((((64, 64, 32), 10), ((32, 32, 16), 6))):
        case = W.urban_block(n=n, boxes=boxes)[0]
        if pc == "ic0":
            case.preconditioner = IC0
        g = cfdgpu.Grid(case)
        u, p = _inputs(case, 40 + step)
        print(f"[static] {pc} step {step} cells={n}", flush=True)
        ua, pa, ia = g.project(u, p, 0.5, 1e-10, 3000)
        ub, pb, ib = g.project(u, p, 0.5, 1e-10, 3000, static=True)
        assert ia["iterations"] == ib["iterations"] > 0
        assert np.array_equal(ua, ub) and np.array_equal(pa, pb)
        assert not np.array_equal(ua, u)


def test_static_projection_refuses_a_grid_beyond_its_array(gpu_available):
    case = W.urban_block(n=(192, 192, 96), boxes=4, preconditioner=IC0)[0]
    g = cfdgpu.Grid(case)
    u = np.zeros(g.face_count)
    p = np.ones(g.count)
    with pytest.raises(cfdgpu.CfdError) as e:
        g.project(u, p, 0.5, 1e-10, 100, static=True)
    assert e.value.code == 200 and "static scratch" in str(e.value)
