"""Grid::build on the device (grid.rs:246-464): mask rasterization, face codes and
union-find components in HIP kernels must give the host build's topology bit for
bit (which the CPU tests hold to the oracle), the same errors, and grids that
solve identically."""
import numpy as np
import pytest

import cfdgpu
from cfdgpu import workloads as W
from cfdgpu.case import MULTIGRID, OPEN, PERIODIC, WALL, Block, Case

from test_host_topology import cases

pytestmark = pytest.mark.gpu


def more_cases():
    out = list(cases())
    out.append(("urban128", W.urban_block()[0]))
    out.append(("prisms96", W.urban_prisms(n=(96, 96, 48), prisms=60, radius=(5.0, 10.0), heights=(3, 22))[0]))
    for d in (0.0, 90.0, 200.0, 290.0):
        out.append((f"wind{int(d)}", W.wind_campaign(d, n=(64, 64, 32), boxes=10)[0]))
    # pockets: closed components inside a wall box, among anchored fluid
    out.append(("pockets", Case(cells=[12, 10, 8], size_m=[12, 10, 8], boundaries=[OPEN, WALL, WALL],
                                velocity_m_s=[1.0, 0, 0],
                                blocks=[Block([2.0, 2, 0], [5.0, 5, 8]), Block([3.0, 0, 0], [4.0, 10, 8]),
                                        Block([7.0, 1, 1], [10.0, 9, 7])])))
    out.append(("periodic-odd", Case(cells=[7, 5, 3], size_m=[7, 5, 3], boundaries=[PERIODIC, PERIODIC, PERIODIC])))
    return out


@pytest.mark.parametrize("name,case", more_cases(), ids=[n for n, _ in more_cases()])
def test_device_topology_equals_host(gpu_available, name, case):
    print(f"[setup] {name} cells={case.cells}", flush=True)
    host = cfdgpu.case_topology(case)
    dev = cfdgpu.case_topology_device(case)
    assert list(dev["counts"]) == list(host["counts"])
    assert np.array_equal(dev["codes"], host["codes"])
    assert np.array_equal(dev["component"], host["component"])
    assert np.array_equal(dev["anchored"], host["anchored"])


def test_device_build_errors_match_host(gpu_available):
    for case, code in (
        (Case(cells=[8, 8, 8], size_m=[8, 8, 8], boundaries=[WALL] * 3, blocks=[Block([1.1, 1.1, 1.1], [1.2, 1.2, 1.2])]),
         "CFD-E028"),
        (Case(cells=[4, 4, 4], size_m=[4, 4, 4], boundaries=[WALL] * 3, blocks=[Block([0.0, 0, 0], [4.0, 4, 4])]),
         "CFD-E030"),
        (Case(cells=[1, 4, 4], size_m=[4, 4, 4], boundaries=[WALL] * 3), "CFD-E004"),
    ):
        with pytest.raises(cfdgpu.CfdError, match=code) as host:
            cfdgpu.case_topology(case)
        with pytest.raises(cfdgpu.CfdError, match=code) as dev:
            cfdgpu.case_topology_device(case)
        assert str(host.value) == str(dev.value)


def test_grids_from_device_setup_solve_identically(gpu_available):
    import os

    case, b, t = W.urban_block(n=(64, 64, 32), boxes=10)
    out = []
    for flag in ("1", "0"):
        os.environ["CFD_GPU_SETUP"] = flag
        try:
            g = cfdgpu.Grid(case)
        finally:
            os.environ.pop("CFD_GPU_SETUP", None)
        x, info = g.solve_scaled(b, t, 1e-10, 2000)
        out.append((x, info["iterations"]))
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][0], out[1][0])
