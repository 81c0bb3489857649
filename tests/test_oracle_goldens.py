"""Pins the CPU oracle against the reference's published numbers for this path.

Goldens (no import of the reference needed):
  * IC(0) / multigrid PCG iterations of the manufactured kernels at tol 1e-11
    (docs/evidence/pressure-multigrid.md:34-38) and their recovery bounds
    (validation/benchmark_pressure.py:102-104: true residual <= 1.01e-10, RMS <= 1e-8);
  * multigrid grid / operator complexity and the 48x2x48 hierarchy sizes
    (pressure-multigrid.md:129-142);
  * canyon-93k: 166 IC(0) iterations, true residual 6.9540e-12, RMS 3.1094e-12
    (docs/evidence/native-cpu-performance.md:210-211).
"""
import json
import math
import os

import numpy as np
import pytest

import oracle
from cfdgpu import workloads as W
from cfdgpu.case import IC0, MULTIGRID, Case

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

KERNELS = [
    ("cavity24", [24, 2, 24], 53, 6),
    ("cavity48", [48, 2, 48], 102, 6),
    ("cavity96", [96, 2, 96], 193, 6),
    ("cube24", [24, 24, 24], 61, 7),
    ("cube48", [48, 48, 48], 113, 7),
]


def _solve(case):
    g = oracle.Grid(case)
    ex = g.center_nullspaces(W.manufactured(case, g.fluid()))
    b = g.apply(ex)
    x, info = g.solve(b, 1e-11, 10000)
    return g, x, info, math.sqrt(float(np.mean((x - ex) ** 2)))


@pytest.mark.parametrize("name,cells,ic0,mg", KERNELS, ids=[k[0] for k in KERNELS])
def test_published_iteration_counts(name, cells, ic0, mg):
    for pc, expected in ((IC0, ic0), (MULTIGRID, mg)):
        g, x, info, rms = _solve(W.cavity(cells, pc))
        assert info["iterations"] == expected, (name, pc, info)
        assert info["relative_residual"] <= 1.01e-10
        assert rms <= 1e-8


COMPLEXITY = {
    "cavity24": (1.328, 2.267),
    "cavity48": (1.332, 2.450),
    "cavity96": (1.333, 2.553),
    "cube24": (1.143, 2.443),
    "cube48": (1.143, 2.601),
}


@pytest.mark.parametrize("name,cells,ic0,mg", KERNELS, ids=[k[0] for k in KERNELS])
def test_published_hierarchy_complexity(name, cells, ic0, mg):
    s = oracle.Grid(W.cavity(cells, MULTIGRID)).mg_stats()
    gc = sum(s["level_cells"]) / s["level_cells"][0]
    oc = sum(s["level_nonzeros"]) / s["level_nonzeros"][0]
    assert (round(gc, 3), round(oc, 3)) == COMPLEXITY[name]


def test_published_48x2x48_hierarchy_sizes():
    s = oracle.Grid(W.cavity([48, 2, 48], MULTIGRID)).mg_stats()
    assert s["level_cells"] == [4608, 1152, 288, 72, 18]
    assert sum(s["level_cells"]) == 6138
    assert sum(s["level_nonzeros"]) == 78084


def test_published_canyon_93k():
    with open(os.path.join(GOLDEN, "canyon-93k.json")) as f:
        case = Case.from_json(json.load(f))
    g, x, info, rms = _solve(case)
    assert info["iterations"] == 166
    assert f"{info['relative_residual']:.4e}" == "6.9540e-12"
    assert f"{rms:.4e}" == "3.1094e-12"
