"""The first aggregation of the masked hierarchy taken from the codes bin by bin
(2x2x2 bins: a union within each bin, each coarse edge summed over the faces of
the one bin interface it lies on) builds the hierarchy the general path builds
from level 0's face-ordered edge list (union-find over every edge, coarse sums
in edge order): the same digest of every index and value of every level, on
grids with outlets, closed components, prisms, odd and periodic axes (periodic
axes of 1, 2 and 3+ bins; two bins take the general path), for the red-black
and the reference hierarchies. Host only."""
import os

import pytest

import cfdgpu
from cfdgpu import workloads as W
from cfdgpu.case import GROUND_OPEN_TOP, MULTIGRID, OPEN, PERIODIC, SLIP, WALL, Block, Case

CASES = {
    "urban": lambda: W.urban_block(n=(64, 48, 40), boxes=8)[0],
    "prisms": lambda: W.urban_prisms(n=(64, 64, 32))[0],
    "odd-periodic": lambda: Case(cells=[27, 20, 18], size_m=[2.7, 2.0, 1.8], boundaries=[PERIODIC, PERIODIC, WALL],
                                 blocks=[Block([0.5, 0.5, 0.0], [1.1, 1.2, 0.9])], preconditioner=MULTIGRID),
    "periodic5": lambda: Case(cells=[10, 9, 7], size_m=[1.0, 0.9, 0.7], boundaries=[PERIODIC, PERIODIC, OPEN],
                              velocity_m_s=(0.0, 0.0, 1.0), blocks=[Block([0.2, 0.2, 0.0], [0.5, 0.4, 0.3])],
                              preconditioner=MULTIGRID),
    "periodic-odd-z": lambda: Case(cells=[12, 8, 9], size_m=[1.2, 0.8, 0.9], boundaries=[OPEN, WALL, PERIODIC],
                                   velocity_m_s=(1.0, 0.0, 0.0), blocks=[Block([0.3, 0.2, 0.3], [0.7, 0.5, 0.6])],
                                   preconditioner=MULTIGRID),
    "periodic2": lambda: Case(cells=[16, 2, 12], size_m=[1.6, 0.2, 1.2], boundaries=[OPEN, PERIODIC, WALL],
                              velocity_m_s=(1.0, 0.0, 0.0), preconditioner=MULTIGRID),
    "periodic4": lambda: Case(cells=[16, 4, 12], size_m=[1.6, 0.4, 1.2], boundaries=[OPEN, PERIODIC, WALL],
                              velocity_m_s=(1.0, 0.0, 0.0), preconditioner=MULTIGRID),
    "odd-open": lambda: Case(cells=[33, 21, 15], size_m=[3.3, 2.1, 1.5], boundaries=[OPEN, SLIP, GROUND_OPEN_TOP],
                             velocity_m_s=(1.0, 0.0, 0.0), blocks=[Block([0.3, 0.2, 0.0], [1.7, 1.5, 0.9])],
                             preconditioner=MULTIGRID),
    "split": lambda: Case(cells=[24, 12, 12], size_m=[2.4, 1.2, 1.2], boundaries=[WALL, WALL, WALL],
                          blocks=[Block([1.1, 0.0, 0.0], [1.3, 1.2, 1.2]), Block([0.3, 0.3, 0.3], [0.5, 0.5, 0.5])],
                          preconditioner=MULTIGRID),
}


@pytest.mark.parametrize("cycle", [cfdgpu.CYCLE_RB, cfdgpu.CYCLE_REFERENCE])
@pytest.mark.parametrize("name", list(CASES))
def test_binned_first_aggregation_is_the_general_one(name, cycle):
    case = CASES[name]()
    fast = cfdgpu.case_multigrid_stats(case, cycle)
    os.environ["CFD_AGG_GENERAL"] = "1"
    try:
        general = cfdgpu.case_multigrid_stats(case, cycle)
    finally:
        del os.environ["CFD_AGG_GENERAL"]
    assert fast["level_cells"] == general["level_cells"]
    assert fast["level_nonzeros"] == general["level_nonzeros"]
    assert fast["digest"] == general["digest"]
