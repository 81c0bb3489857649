"""The benchmark configurations on the device against the oracle:

  * the urban block at its full size (128x128x64, 24 boxes, open x outlet):
    red-black multigrid PCG in FP64 and mixed precision and the reference's own
    multigrid cycle, each against the oracle's solve of the same system
    (pressure.rs:342-436); IC(0)-PCG on a 96x96x48 urban block;
  * the prism city (convex 4-8-vertex prisms) at 96x96x48 in mixed precision:
    the FP64 path's iteration count and the oracle's field;
  * the street-canyon wind campaign at 96x96x40: two directions (two
    operators), a cold FP64 solve, then a chain of right-hand sides
    b_{k+1} = b_k + 0.02 rms(b_k) xi solved warm-started in mixed precision,
    the last step against the oracle.
"""
import math

import numpy as np
import pytest

import cfdgpu
import oracle
from cfdgpu import workloads as W
from cfdgpu.case import IC0

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def urban():
    case, b, t = W.urban_block()
    o = oracle.Grid(case)
    print("[configs] urban block 128x128x64: oracle multigrid solve at 1e-13", flush=True)
    xo, io = o.solve_scaled(b, t, 1e-13, 3000)
    return case, b, t, o, xo, io


@pytest.mark.parametrize("precision", [cfdgpu.PRECISION_FP64, cfdgpu.PRECISION_MIXED], ids=["fp64", "mixed"])
def test_urban_block_full_size_matches_oracle(gpu_available, urban, precision):
    case, b, t, o, xo, io = urban
    g = cfdgpu.Grid(case)
    g.set_options(precision)
    x, info = g.solve_scaled(b, t, 1e-13, 3000)
    assert info["relative_residual"] <= 1.01e-12
    assert rel(x, xo) <= 1e-10, (rel(x, xo), info["iterations"], io["iterations"])
    # the configuration's own tolerance: both gates of the reference
    x10, i10 = g.solve_scaled(b, t, 1e-10, 3000)
    assert i10["recurrent_residual"] <= 1e-10 and i10["relative_residual"] <= 1e-9 + 1e-12
    assert rel(x10, xo) <= 1e-7


def test_urban_block_full_size_reference_cycle_iterations(gpu_available, urban):
    """The reference's own multigrid cycle takes the oracle's iteration count on
    the full urban block (rel-res 1e-8)."""
    case, b, t, o, xo, io = urban
    xo8, io8 = o.solve_scaled(b, t, 1e-8, 3000)
    g = cfdgpu.Grid(case)
    g.set_options(cfdgpu.PRECISION_FP64, cycle=cfdgpu.CYCLE_REFERENCE)
    x, info = g.solve_scaled(b, t, 1e-8, 3000)
    assert abs(info["iterations"] - io8["iterations"]) <= 1, (info["iterations"], io8["iterations"])
    assert rel(x, xo8) <= 1e-9


def test_urban_block_ic0_matches_oracle(gpu_available):
    """The urban block with the IC(0) preconditioner at the configuration's 1e-10
    (at 96x96x48 with 14 boxes: the oracle's serial IC(0) takes about 80 s at the
    full size): the oracle's iteration count within one and its field."""
    case, b, t = W.urban_block(n=(96, 96, 48), boxes=14, preconditioner=IC0)
    print("[configs] urban block 96x96x48: oracle IC(0) solve at 1e-10", flush=True)
    xo, io = oracle.Grid(case).solve_scaled(b, t, 1e-10, 3000)
    x, info = cfdgpu.Grid(case).solve_scaled(b, t, 1e-10, 3000)
    assert abs(info["iterations"] - io["iterations"]) <= 1, (info["iterations"], io["iterations"])
    assert rel(x, xo) <= 1e-10


def test_lid_cavity_48_mixed_ic0(gpu_available):
    """The closed 48^3 cavity with IC(0) applied in FP32 inside the FP64 PCG
    (pressure.rs:223-241 in single precision over FP64 pivots): the FP64 path's
    iteration count within a few at 1e-10, and the oracle's field at 1e-13."""
    case, b, t = W.lid_cavity_48()
    o = oracle.Grid(case)
    xo, io = o.solve_scaled(b, t, 1e-13, 3000)
    g = cfdgpu.Grid(case)
    its = {}
    for precision in (cfdgpu.PRECISION_FP64, cfdgpu.PRECISION_MIXED):
        g.set_options(precision)
        x, info = g.solve_scaled(b, t, 1e-10, 3000)
        assert info["relative_residual"] <= 1.01e-9
        its[precision] = info["iterations"]
    assert abs(its[cfdgpu.PRECISION_MIXED] - its[cfdgpu.PRECISION_FP64]) <= 3, its
    x, info = g.solve_scaled(b, t, 1e-13, 3000)
    assert info["relative_residual"] <= 1.01e-12
    assert rel(x, xo) <= 1e-10, rel(x, xo)


def test_prism_city_mixed_precision(gpu_available):
    """Mixed precision (FP32 V-cycle inside FP64 PCG) on convex prisms: the FP64
    path's iteration count at 1e-8 and 1e-10, and the oracle's field at 1e-13."""
    case, b, t = W.urban_prisms(n=(96, 96, 48), prisms=60, radius=(5.0, 10.0), heights=(3, 22))
    o = oracle.Grid(case)
    print("[configs] prisms 96x96x48: oracle solve at 1e-13", flush=True)
    xo, io = o.solve_scaled(b, t, 1e-13, 3000)
    g = cfdgpu.Grid(case)
    for tol in (1e-8, 1e-10):
        its = {}
        for precision in (cfdgpu.PRECISION_FP64, cfdgpu.PRECISION_MIXED):
            g.set_options(precision)
            x, info = g.solve_scaled(b, t, tol, 3000)
            assert info["relative_residual"] <= 10 * tol + 1e-12
            its[precision] = info["iterations"]
        assert abs(its[cfdgpu.PRECISION_MIXED] - its[cfdgpu.PRECISION_FP64]) <= 1, (tol, its)
    g.set_options(cfdgpu.PRECISION_MIXED)
    x, info = g.solve_scaled(b, t, 1e-13, 3000)
    assert info["relative_residual"] <= 1.01e-12
    assert rel(x, xo) <= 1e-10, rel(x, xo)


@pytest.mark.parametrize("direction", [0.0, 200.0])
def test_street_canyon_chain(gpu_available, direction):
    case, b, t = W.street_canyon(direction, n=(96, 96, 40), lots=5, street=3.0, heights=(8, 38))
    fluid = W.fluid_mask(case)
    assert 0.15 < 1.0 - fluid.mean() < 0.35
    o = oracle.Grid(case)
    g = cfdgpu.Grid(case)
    print(f"[configs] street canyon 96x96x40 direction {direction}", flush=True)
    x0, i0 = g.solve_scaled(b, t, 1e-8, 3000)  # cold, FP64
    assert i0["recurrent_residual"] <= 1e-8 and i0["relative_residual"] <= 1e-7 + 1e-12
    rng = np.random.default_rng(5)
    g.set_options(cfdgpu.PRECISION_MIXED, warm_start=True)
    g.solve_scaled(b, t, 1e-8, 3000)  # the resident solution the chain starts from
    chain = b
    its = []
    for _ in range(5):
        rms = math.sqrt(float(np.mean(chain[fluid] ** 2)))
        chain = chain + np.where(fluid, 0.02 * rms * rng.standard_normal(chain.size), 0.0)
        x, info = g.solve_scaled(chain, None, 1e-8, 3000)
        assert info["recurrent_residual"] <= 1e-8 and info["relative_residual"] <= 1e-7 + 1e-12
        its.append(info["iterations"])
    assert max(its) <= i0["iterations"] + 1, (its, i0["iterations"])
    xo, io = o.solve(chain, 1e-13, 3000)
    assert rel(x, xo) <= 1e-6, rel(x, xo)
    xt, it = g.solve_scaled(chain, None, 1e-13, 3000)
    assert it["relative_residual"] <= 1.01e-12
    assert rel(xt, xo) <= 1e-10, rel(xt, xo)
