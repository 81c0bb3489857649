"""Pins the CPU oracle against the reference's published numbers for this path.

Goldens (no import of the reference needed):
  * IC(0) / multigrid PCG iterations of the manufactured kernels at tol 1e-11
    (docs/evidence/pressure-multigrid.md:34-38) and their recovery bounds
    (validation/benchmark_pressure.py:102-104: true residual <= 1.01e-10, RMS <= 1e-8);
  * multigrid grid / operator complexity and the 48x2x48 hierarchy sizes
    (pressure-multigrid.md:129-142);
  * canyon-93k: 166 IC(0) iterations, true residual 6.9540e-12, RMS 3.1094e-12
    (docs/evidence/native-cpu-performance.md:210-211);
  * the 18-shape anisotropy suite: IC(0) and multigrid (line/plane block smoothers)
    iterations at tol 1e-14 and the suite's maximum recovery error and true residual
    (docs/evidence/anisotropic-pressure.md).
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


# ---------------------------------------------------------------- anisotropy suite
def _aniso():
    with open(os.path.join(os.path.dirname(__file__), "golden", "anisotropy_shapes.json")) as f:
        return json.load(f)


def aniso_case(spec, preconditioner):
    from cfdgpu.case import IC0, MULTIGRID, OPEN, PERIODIC, WALL, Case

    bx = {"wall": WALL, "periodic": PERIODIC, "open": OPEN}[spec["x"]]
    return Case(cells=spec["cells"], size_m=spec["size_m"], boundaries=[bx, PERIODIC, WALL],
                velocity_m_s=[1.0, 0.0, 0.0] if bx == OPEN else [0.0, 0.0, 0.0],
                preconditioner=IC0 if preconditioner == "ic0" else MULTIGRID)


def aniso_rhs(spec, count, apply):
    i = np.arange(count, dtype=float)
    ex = np.sin(0.73 * i) + 0.3 * np.cos(0.17 * i)
    if spec["x"] != "open":
        ex -= ex.mean()
    return ex, apply(ex)


def test_published_anisotropy_ic0():
    """The 18 shapes of the anisotropy suite: IC(0)-PCG iterations at tol 1e-14 equal
    the published ones; the largest recovery error and true residual over the
    suite equal the published maxima (anisotropic-pressure.md:27-30,43-60)."""
    data = _aniso()
    rms_max = res_max = 0.0
    for name, spec in data["shapes"].items():
        o = oracle.Grid(aniso_case(spec, "ic0"))
        ex, b = aniso_rhs(spec, o.count, o.apply)
        x, info = o.solve(b, 1e-14, 1200)
        assert info["iterations"] == spec["ic0_iterations"], name
        rms = float(np.sqrt(np.mean((x - ex) ** 2)))
        assert rms <= data["gates"]["rms_max"] and info["relative_residual"] <= data["gates"]["true_residual_max"]
        rms_max = max(rms_max, rms)
        res_max = max(res_max, info["relative_residual"])
    assert float(f"{rms_max:.3e}") in (1.844e-10, 1.843e-10, 1.845e-10), rms_max
    assert abs(res_max - data["gates"]["published_max_true_residual"]) <= 0.005e-14, res_max


def test_published_anisotropy_multigrid():
    """Multigrid PCG iterations of the 18 shapes at tol 1e-14: the published
    candidate column (anisotropic shapes take line or small-plane block
    smoothers with semicoarsening, multigrid/line.rs and multigrid.rs:573-611;
    the three shapes below the threshold keep point Gauss-Seidel)."""
    data = _aniso()
    for name, spec in data["shapes"].items():
        o = oracle.Grid(aniso_case(spec, "mg"))
        ex, b = aniso_rhs(spec, o.count, o.apply)
        x, info = o.solve(b, 1e-14, 1200)
        assert info["iterations"] == spec["mg_iterations"], name
        assert float(np.sqrt(np.mean((x - ex) ** 2))) <= data["gates"]["rms_max"]
        assert info["relative_residual"] <= data["gates"]["true_residual_max"]
        if spec["point_mg_iterations"] is not None:
            assert spec["point_mg_iterations"] == spec["mg_iterations"]
