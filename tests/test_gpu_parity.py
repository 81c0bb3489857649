"""GPU parity of the pressure path against the CPU oracle (the reference's serial
FP64 arithmetic restated in oracle/), through the C-ABI of libcfdgpu.so.

Bars: the operator and one IC(0) application are bitwise equal to the oracle
(same per-cell arithmetic, FP contraction off); PCG-IC(0) solves take the same
iteration count within +-1 and match the oracle's field to 1e-10 relative; the
red-black multigrid cycles differ from the reference's lexicographic one, so
their solves are held to the same field bound and to the acceptance gates.
"""
import math
import os
from contextlib import contextmanager

import numpy as np
import pytest

import oracle
import cfdgpu
from cfdgpu import workloads as W
from cfdgpu.case import GROUND_OPEN_TOP, IC0, MULTIGRID, OPEN, PERIODIC, SLIP, WALL, Block, Case

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@contextmanager
def env(**kv):
    """Set-up knobs read when a grid is created (CFD_MG_DENSE, CFD_MG_RB_PAIRS, ...)."""
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


def small_cases():
    cases = [("cavity24", W.cavity([24, 2, 24])), ("cube12", W.cavity([12, 12, 12]))]
    for n in ([9, 2, 7], [8, 5, 6], [2, 2, 2], [9, 1, 7], [6, 4, 5]):
        for bnd in (WALL, PERIODIC, OPEN):
            cases.append((f"lid{n}-{bnd}", Case(cells=n, size_m=[2.7, 1.9, 3.1], boundaries=[bnd, PERIODIC, WALL],
                                                velocity_m_s=[1.0, 0.0, 0.0])))
    cases.append(("barrier", Case(cells=[9, 2, 5], size_m=[9, 2, 5], boundaries=[OPEN, WALL, WALL],
                                  velocity_m_s=[1.0, 0, 0], blocks=[Block([3.0, 0, 0], [4.0, 2, 5])])))
    c, _, _ = W.urban_block(n=(32, 32, 16), boxes=6)
    cases.append(("urban32", c))
    return cases


@pytest.mark.parametrize("name,case", small_cases(), ids=[n for n, _ in small_cases()])
def test_operator_bitwise(gpu_available, name, case):
    g = cfdgpu.Grid(case)
    o = oracle.Grid(case)
    rng = np.random.default_rng(7)
    x = rng.standard_normal(g.count)
    y = g.apply(x)
    assert np.array_equal(y, o.apply(x)), name
    assert np.array_equal(y, o.operator_apply(x)), name


@pytest.mark.parametrize("name,case", small_cases(), ids=[n for n, _ in small_cases()])
def test_ic0_apply_bitwise(gpu_available, name, case):
    case.preconditioner = IC0
    g = cfdgpu.Grid(case)
    o = oracle.Grid(case)
    rng = np.random.default_rng(11)
    r = rng.standard_normal(g.count)
    z = g.precondition(r)
    assert np.array_equal(z, o.ic0_apply(r)), name


@pytest.mark.parametrize("name,case", small_cases() + [("lid20-odd", Case(cells=[20, 13, 9], size_m=[2.0, 1.3, 0.9],
                                                                       boundaries=[PERIODIC, WALL, PERIODIC]))],
                         ids=[n for n, _ in small_cases()] + ["lid20-odd"])
def test_ic0_tiles_equal_hyperplanes(gpu_available, name, case):
    """The tiled wavefront of the IC(0) sweeps gives the bits of one launch per hyperplane."""
    case.preconditioner = IC0
    r = np.random.default_rng(12).standard_normal(case.count)
    out = []
    for tiles in (1, 0):
        with env(CFD_IC0_TILES=tiles):
            g = cfdgpu.Grid(case)
        out.append(g.precondition(r))
    assert np.array_equal(out[0], out[1])


def _fill_cases():
    from test_host_topology import fill_cases

    return fill_cases()


@pytest.mark.parametrize("name,case", _fill_cases(), ids=[n for n, _ in _fill_cases()])
def test_ic0_with_fill_coupling(gpu_available, name, case):
    """A periodic axis of exactly 3 cells: the IC(0) entries with their dot terms
    (pressure.rs:197-209) come from the host; the application is bitwise the
    oracle's (tiled and per hyperplane), FP64 PCG takes the oracle's iterations
    and field, and the FP32 application (mixed) reaches the same field."""
    case.preconditioner = IC0
    o = oracle.Grid(case)
    rng = np.random.default_rng(17)
    r = rng.standard_normal(o.count)
    for tiles in (1, 0):
        with env(CFD_IC0_TILES=tiles):
            g = cfdgpu.Grid(case)
        print(f"[ic0 fill] {name} tiles={tiles}", flush=True)
        assert np.array_equal(g.precondition(r), o.ic0_apply(r)), (name, tiles)
    ex = o.center_nullspaces(np.where(o.fluid(), rng.standard_normal(o.count), 0.0))
    b = o.apply(ex)
    xo, io = o.solve(b, 1e-13, 5000)
    x, info = g.solve(b, 1e-13, 5000)
    assert abs(info["iterations"] - io["iterations"]) <= 1, (info, io)
    assert rel(x, xo) <= 1e-10
    g.set_options(cfdgpu.PRECISION_MIXED)
    xm, im = g.solve(b, 1e-13, 5000)
    assert rel(xm, xo) <= 1e-10, (im, io)


def _manufactured(case):
    o = oracle.Grid(case)
    ex = W.manufactured(case, o.fluid())
    ex = o.center_nullspaces(ex)
    return o, ex, o.apply(ex)


@pytest.mark.parametrize("cells,its", [([24, 2, 24], 53), ([48, 2, 48], 102), ([24, 24, 24], 61)])
def test_ic0_pcg_matches_published_and_oracle(gpu_available, cells, its):
    case = W.cavity(cells, IC0)
    o, ex, b = _manufactured(case)
    xo, io = o.solve(b, 1e-11, 10000)
    g = cfdgpu.Grid(case)
    x, info = g.solve(b, 1e-11, 10000)
    assert abs(info["iterations"] - its) <= 1
    assert abs(info["iterations"] - io["iterations"]) <= 1
    assert info["relative_residual"] <= 1.01e-10
    assert rel(x, xo) <= 1e-10
    assert math.sqrt(np.mean((x - ex) ** 2)) <= 1e-8


@pytest.mark.parametrize("name,case", small_cases(), ids=[n for n, _ in small_cases()])
@pytest.mark.parametrize("pc", [IC0, MULTIGRID, "multigrid-levels"])
def test_solve_field_parity_small(gpu_available, name, case, pc):
    """multigrid: small grids fall under the dense coarsest size and are solved
    directly; multigrid-levels forces coarse levels down to 8 nodes per component."""
    dense = 8 if pc == "multigrid-levels" else 2048
    pc = MULTIGRID if pc == "multigrid-levels" else pc
    case.preconditioner = pc
    o = oracle.Grid(case)
    rng = np.random.default_rng(5)
    ex = o.center_nullspaces(np.where(o.fluid(), rng.standard_normal(o.count), 0.0))
    b = o.apply(ex)
    xo, io = o.solve(b, 1e-13, 5000)
    with env(CFD_MG_DENSE=dense):
        g = cfdgpu.Grid(case)
    x, info = g.solve(b, 1e-13, 5000)
    assert info["relative_residual"] <= 1e-12 + 1e-12
    assert rel(x, xo) <= 1e-10, (name, info, io)
    if pc == IC0:
        assert abs(info["iterations"] - io["iterations"]) <= 1


def test_multigrid_urban_block_field(gpu_available):
    case, b, t = W.urban_block(n=(64, 64, 32), boxes=10)
    o = oracle.Grid(case)
    xo, io = o.solve_scaled(b, t, 1e-10, 2000)
    for precision in (cfdgpu.PRECISION_FP64, cfdgpu.PRECISION_MIXED):
        g = cfdgpu.Grid(case)
        g.set_options(precision)
        x, info = g.solve_scaled(b, t, 1e-10, 2000)
        assert info["relative_residual"] <= 1.01e-9
        # two solves at residual 1e-10 differ by at most ~cond * 1e-10; the field bound is looser
        assert rel(x, xo) <= 1e-7, (precision, info, io)


def test_zero_rhs_returns_zero_iterations(gpu_available):
    g = cfdgpu.Grid(W.cavity([8, 2, 8]))
    x, info = g.solve(np.zeros(g.count), 1e-10, 100)
    assert info["iterations"] == 0 and np.all(x == 0.0)


def test_incompatible_flux_is_e031(gpu_available):
    case = Case(cells=[9, 2, 7], size_m=[2.7, 1.9, 3.1], boundaries=[WALL, PERIODIC, WALL])
    g = cfdgpu.Grid(case)
    b = np.cos(0.3 * np.arange(g.count))
    b = g.center_nullspaces(b) + 1e-11
    g.solve(b, 1e-12, 2000)  # round-off incompatibility is accepted
    with pytest.raises(cfdgpu.CfdError) as e:
        g.solve(b + 0.1, 1e-12, 2000)
    assert e.value.code == 31 and "CFD-E031" in str(e.value)


def test_iteration_limit_is_e033_and_cancel(gpu_available):
    case = W.cavity([24, 2, 24])
    o, ex, b = _manufactured(case)
    g = cfdgpu.Grid(case)
    with pytest.raises(cfdgpu.CfdError) as e:
        g.solve(b, 1e-11, 1)
    assert e.value.code == 33 and "CFD-E033" in str(e.value)
    calls = []
    ctl = cfdgpu.Control(cancel=lambda: calls.append(1) or len(calls) > 2)
    with pytest.raises(cfdgpu.CfdError) as e:
        g.solve(b, 1e-11, 10000, control=ctl)
    assert e.value.code == cfdgpu.CANCELLED and str(e.value) == "cancelled"


def test_projection_matches_oracle(gpu_available):
    case, b, t = W.urban_block(n=(32, 32, 16), boxes=6)
    case.preconditioner = IC0
    o = oracle.Grid(case)
    rng = np.random.default_rng(3)
    fluid = o.fluid()
    kinds, zc, offs = W.face_kinds(case, fluid)
    u = W.enforce_velocity(case, kinds, zc, offs, rng.standard_normal(o.face_count))
    p = o.center_nullspaces(rng.standard_normal(o.count))
    g = cfdgpu.Grid(case)
    bg, tg = g.projection_rhs(u, 0.25)
    bo, to = o.projection_rhs(u, 0.25)
    assert np.array_equal(bg, bo) and np.array_equal(tg, to)
    uo, po, io = o.project(u, p, 0.25, 1e-12, 3000)
    ug, pg, ig = g.project(u, p, 0.25, 1e-12, 3000)
    assert abs(ig["iterations"] - io["iterations"]) <= 1
    assert rel(ug, uo) <= 1e-10 and rel(pg, po) <= 1e-10
    div = o.divergence(ug)
    assert np.max(np.abs(div[fluid])) < 1e-8 * max(1.0, np.max(np.abs(bo)))


def test_deterministic_repeat(gpu_available):
    case, b, t = W.urban_block(n=(32, 32, 16), boxes=6)
    g = cfdgpu.Grid(case)
    x1, i1 = g.solve_scaled(b, t, 1e-10, 2000)
    x2, i2 = g.solve_scaled(b, t, 1e-10, 2000)
    assert np.array_equal(x1, x2) and i1["iterations"] == i2["iterations"]


@pytest.mark.parametrize("pc", [MULTIGRID, IC0], ids=["multigrid", "ic0"])
def test_queued_iterations_do_not_change_results(gpu_available, pc):
    """The solver queues as many iterations as the last such solve needed and
    extrapolates the rest; iterations queued past convergence do nothing. A grid
    that learned a long solve (tight tolerance) and then solves an easier system,
    or a harder one after an easy one, gives the bits and counts of a fresh grid,
    for every tolerance in turn and with an iteration limit inside a queued
    segment (CFD-E033 at the same count)."""
    case, b, t = W.urban_block(n=(32, 32, 16), boxes=6)
    case.preconditioner = pc
    seq = [(b, 1e-12), (0.5 * b, 1e-6), (b, 1e-10), (2.0 * b, 1e-12), (b, 1e-8)]
    g = cfdgpu.Grid(case)
    for k, (rhs, tol) in enumerate(seq):
        print(f"[queue] {pc} step {k} tol={tol}", flush=True)
        x, info = g.solve_scaled(rhs, t, tol, 3000)
        fresh = cfdgpu.Grid(case)
        xf, inf = fresh.solve_scaled(rhs, t, tol, 3000)
        assert info["iterations"] == inf["iterations"], (k, info, inf)
        assert np.array_equal(x, xf), k
    # a limit that falls inside the learned segments
    n = info["iterations"]
    for limit in (n - 1, 3):
        with pytest.raises(cfdgpu.CfdError) as e:
            g.solve_scaled(b, t, 1e-8, limit)
        assert e.value.code == 33


@pytest.mark.parametrize("cycle", [cfdgpu.CYCLE_RB])
def test_odd_periodic_multigrid_is_deterministic(gpu_available, cycle):
    """Odd periodic coarse levels (36 -> 18 -> 9): repeated cycles are bitwise equal."""
    case = W.cavity([40, 36, 28], MULTIGRID)
    r = np.random.default_rng(4).standard_normal(case.count)
    with env(CFD_MG_DENSE=8):
        g = cfdgpu.Grid(case)
    g.set_options(cfdgpu.PRECISION_FP64, cycle=cycle)
    a = g.precondition(r)
    b = g.precondition(r)
    assert np.array_equal(a, b)
    x, info = g.solve(g.center_nullspaces(r), 1e-10, 500)
    assert info["relative_residual"] <= 1.01e-9


MARCH_CASES = [
    ("urban", lambda: W.urban_block(n=(70, 50, 40), boxes=8)[0]),
    ("periodic-odd", lambda: Case(cells=[67, 33, 21], size_m=[6.7, 3.3, 2.1], boundaries=[PERIODIC, PERIODIC, PERIODIC],
                                  blocks=[Block([1.0, 1.0, 0.0], [2.0, 2.0, 2.1])], preconditioner=MULTIGRID)),
    ("closed", lambda: W.cavity([40, 18, 36], MULTIGRID)),
    ("ic0-closed", lambda: W.cavity([36, 20, 18], IC0)),
    ("courtyards", None),
]


@pytest.mark.parametrize("name,make", MARCH_CASES, ids=[n for n, _ in MARCH_CASES])
def test_marching_operator_solves_like_cell_by_cell(gpu_available, name, make):
    """The operator kernel that marches through z (tiles of 64 x 16 columns over
    16 layers, p of the layers above and below in registers, the layer in LDS;
    used on large grids, forced here) forms every p and (A p)_c with the
    cell-by-cell kernel's arithmetic: the same iteration counts and the same
    fields up to the order of the p . Ap sums, on partial tiles, periodic axes
    and closed components."""
    if make is None:
        from test_gpu_components import large_courtyards

        case = large_courtyards()
    else:
        case = make()
    o = oracle.Grid(case)
    ex = o.center_nullspaces(np.where(o.fluid(), np.random.default_rng(4).standard_normal(o.count), 0.0))
    b = o.apply(ex)
    out = []
    for march in (2, 0):
        with env(CFD_MARCH=march):
            g = cfdgpu.Grid(case)
        print(f"[march] {name} march={march}", flush=True)
        out.append(g.solve(b, 1e-11, 3000))
    (x2, i2), (x0, i0) = out
    assert i2["iterations"] == i0["iterations"], (i2["iterations"], i0["iterations"])
    assert rel(x2, x0) <= 1e-11, rel(x2, x0)
    assert rel(x2, ex) <= 1e-8


@pytest.mark.parametrize("name", ["odd-periodic-coarse", "odd-periodic-fine", "urban"])
def test_coarse_restriction_zero_residual_of_the_last_class(gpu_available, name):
    """The coarse restriction takes the residual of the colour class relaxed last
    as zero (after its relaxation it is zero to rounding). On levels coloured
    greedily because of odd periodic seams (36 -> 18 -> 9 cells) as on the
    two-coloured ones, one FP64 V-cycle with that shortcut equals the V-cycle
    that forms every residual to rounding; taking any other class as zero would
    change it at order one."""
    if name == "odd-periodic-coarse":
        case = W.cavity([40, 36, 28], MULTIGRID)
    elif name == "odd-periodic-fine":
        case = Case(cells=[27, 20, 18], size_m=[2.7, 2.0, 1.8], boundaries=[PERIODIC, PERIODIC, WALL],
                    blocks=[Block([0.5, 0.5, 0.0], [1.1, 1.2, 0.9])], preconditioner=MULTIGRID)
    else:
        case = W.urban_block(n=(64, 48, 40), boxes=8)[0]
    rng = np.random.default_rng(13)
    r = rng.standard_normal(case.count)
    out = []
    for zero_last in (1, 0):
        with env(CFD_MG_DENSE=8, CFD_MG_ZERO_LAST=zero_last):
            g = cfdgpu.Grid(case)
        if name != "urban":
            r = g.center_nullspaces(r)
        out.append(g.precondition(r))
    assert rel(out[0], out[1]) <= 1e-12, rel(out[0], out[1])


@pytest.mark.parametrize("cells", [[27, 20, 18], [33, 21, 16]])
def test_odd_periodic_fine_level_colours(gpu_available, cells):
    """Periodic axes of odd length on the fine level (seam colour classes): the
    cycle is deterministic, symmetric in effect (PCG converges) and the field
    matches the oracle."""
    case = Case(cells=cells, size_m=[2.7, 2.0, 1.8], boundaries=[PERIODIC, PERIODIC, WALL],
                blocks=[Block([0.5, 0.5, 0.0], [1.1, 1.2, 0.9])])
    case.preconditioner = MULTIGRID
    o = oracle.Grid(case)
    rng = np.random.default_rng(9)
    ex = o.center_nullspaces(np.where(o.fluid(), rng.standard_normal(o.count), 0.0))
    b = o.apply(ex)
    xo, io = o.solve(b, 1e-13, 5000)
    with env(CFD_MG_DENSE=8):
        g = cfdgpu.Grid(case)
    for precision in (cfdgpu.PRECISION_FP64, cfdgpu.PRECISION_MIXED):
        g.set_options(precision)
        assert np.array_equal(g.precondition(b), g.precondition(b))
        x, info = g.solve(b, 1e-13, 5000)
        assert info["relative_residual"] <= 2e-12
        assert rel(x, xo) <= 1e-10, (precision, info, io)


LAYOUT_CASES = [("urban", [64, 48, 40]), ("urban", [33, 27, 21]), ("urban", [130, 20, 9]), ("urban", [260, 22, 7]),
                ("urban", [63, 20, 9]),
                ("cavity", [70, 18, 22]), ("cavity", [40, 2, 30]), ("cavity", [30, 1, 20]), ("cavity", [36, 4, 12])]


@pytest.mark.parametrize("kind,cells", LAYOUT_CASES, ids=[f"{k}{'x'.join(map(str, c))}" for k, c in LAYOUT_CASES])
@pytest.mark.parametrize("precision", [cfdgpu.PRECISION_FP64, cfdgpu.PRECISION_MIXED])
def test_colour_separated_fine_level_is_bitwise(gpu_available, kind, cells, precision):
    """The red-black cycle with the fine level stored colour by colour, with one
    or (nx a multiple of 4) two cells per thread (in tiles, or consecutive
    threads on consecutive pairs), gives the same bits as with the natural
    layout (same per-cell arithmetic), odd sizes and periodic axes of 1 and 2
    cells included."""
    if kind == "urban":
        case = W.urban_block(n=tuple(cells), boxes=6)[0]
    else:
        case = W.cavity(cells, MULTIGRID)
    r = np.random.default_rng(8).standard_normal(case.count)
    if kind == "cavity":
        r -= r.mean()  # one closed component
    out = []
    for layout, pairs, lin in ((1, 0, 1), (1, 1, 1), (0, 0, 1), (1, 1, 0)):
        with env(CFD_MG_RB_LAYOUT=layout, CFD_MG_RB_PAIRS=pairs, CFD_MG_RB_LIN=lin, CFD_MG_DENSE=64):
            g = cfdgpu.Grid(case)
        g.set_options(precision)
        print(f"[layout] {kind} {cells} precision={precision} layout={layout} pairs={pairs} linear={lin}", flush=True)
        out.append(g.precondition(r))
    for k in (1, 2, 3):
        assert np.array_equal(out[0], out[k]), (k, float(np.max(np.abs(out[0] - out[k]))))


def test_split_aggregation_converges_on_prisms(gpu_available):
    """Polygon footprints leave thin solid corners: the aggregation splits each
    2x2x2 bin into connected groups (masked.rs:127-156), which keeps the
    red-black cycle fast where plain bins merge fluid across walls (plain
    2x2x2 bins took 64 PCG iterations on the 256x256x128 prism city, the split
    hierarchy 13-16)."""
    case, b, t = W.urban_prisms(n=(96, 96, 48), prisms=60, radius=(5.0, 10.0), heights=(3, 22))
    o = oracle.Grid(case)
    xo, io = o.solve_scaled(b, t, 1e-11, 3000)
    its = {}
    for cycle in (cfdgpu.CYCLE_RB,):
        g = cfdgpu.Grid(case)
        g.set_options(cfdgpu.PRECISION_FP64, cycle=cycle)
        x, info = g.solve_scaled(b, t, 1e-11, 3000)
        assert info["relative_residual"] <= 1.01e-10
        assert rel(x, xo) <= 1e-8, (cycle, info, io)
        its[cycle] = info["iterations"]
    assert its[cfdgpu.CYCLE_RB] <= 25, its


def test_rb_cycle_deterministic_and_mixed_close(gpu_available):
    case, b, t = W.urban_block(n=(64, 48, 40), boxes=8)
    with env(CFD_MG_DENSE=64):
        g = cfdgpu.Grid(case)
    r = np.random.default_rng(3).standard_normal(case.count)
    z1 = g.precondition(r)
    assert np.array_equal(z1, g.precondition(r))
    g.set_options(cfdgpu.PRECISION_MIXED)
    z2 = g.precondition(r)
    assert np.array_equal(z2, g.precondition(r))
    assert rel(z2, z1) <= 1e-5


# ---------------------------------------------------------------- the reference's cycle
def _ref_grid(case):
    g = cfdgpu.Grid(case)
    g.set_options(cfdgpu.PRECISION_FP64, cycle=cfdgpu.CYCLE_REFERENCE)
    return g


def _ref_cases():
    out = list(small_cases())
    out.append(("prisms40", W.urban_prisms(n=(40, 36, 20), prisms=10, radius=(3.0, 6.0), heights=(3, 12))[0]))
    out.append(("cube20", W.cavity([20, 20, 20])))
    return out


@pytest.mark.parametrize("name,case", _ref_cases(), ids=[n for n, _ in _ref_cases()])
def test_reference_cycle_matches_oracle_vcycle(gpu_available, name, case):
    """One application of the reference's V-cycle (lexicographic Gauss-Seidel by
    dependency level, the reference's hierarchy and centering points) equals the
    oracle's serial cycle up to the order of the centering sums."""
    case.preconditioner = MULTIGRID
    o = oracle.Grid(case)
    rng = np.random.default_rng(21)
    r = o.center_nullspaces(np.where(o.fluid(), rng.standard_normal(o.count), 0.0))
    zo = o.mg_apply(r)
    g = _ref_grid(case)
    z = g.precondition(r)
    assert rel(z, zo) <= 1e-12, (name, rel(z, zo))


def _tile_cases():
    out = [(n, c) for n, c in _ref_cases()]
    out.append(("periodic3", Case(cells=[12, 3, 10], size_m=[1.2, 0.3, 1.0], boundaries=[WALL, PERIODIC, WALL],
                                  blocks=[Block([0.3, 0.0, 0.2], [0.6, 0.3, 0.5])])))
    out.append(("odd-periodic", Case(cells=[27, 20, 18], size_m=[2.7, 2.0, 1.8], boundaries=[PERIODIC, PERIODIC, WALL],
                                     blocks=[Block([0.5, 0.5, 0.0], [1.1, 1.2, 0.9])])))
    out.append(("urban", W.urban_block(n=(48, 40, 24), boxes=6)[0]))
    # trilinear levels of 4096 rows (iterate in LDS) and 32768 rows (in global memory)
    out.append(("cube32", W.cavity([32, 32, 32], MULTIGRID)))
    out.append(("cube64", W.cavity([64, 64, 64], MULTIGRID)))
    return out


@pytest.mark.parametrize("cycle", [cfdgpu.CYCLE_REFERENCE, cfdgpu.CYCLE_ZEBRA])
@pytest.mark.parametrize("name,case", _tile_cases(), ids=[n for n, _ in _tile_cases()])
def test_reference_cycle_tiles_are_bitwise(gpu_available, name, case, cycle):
    """The reference (and zebra) cycle's fine sweeps over 8x8x8 tiles (flat 16x2x16
    ones across an axis of at most 2 cells), one launch per tile diagonal, and its
    coarse levels of at most 65536 rows swept in one workgroup (from sweep-ordered
    copies of their rows, or from the CSR) give the bits of one launch per
    hyperplane and per dependency level: one V-cycle and a whole solve, on
    periodic axes of 1, 2, 3 and odd length, closed components and the
    trilinear hierarchy."""
    case.preconditioner = MULTIGRID
    o = oracle.Grid(case)
    rng = np.random.default_rng(31)
    r = o.center_nullspaces(np.where(o.fluid(), rng.standard_normal(o.count), 0.0))
    b = o.apply(o.center_nullspaces(np.where(o.fluid(), rng.standard_normal(o.count), 0.0)))
    out = []
    for tiles, packed in ((1, 1), (1, 0), (0, 0)):
        with env(CFD_REF_TILES=tiles, CFD_REF_PACKED=packed, CFD_MG_CYCLE=cycle):
            g = cfdgpu.Grid(case)
        g.set_options(cfdgpu.PRECISION_FP64, cycle=cycle)
        print(f"[ref tiles] {name} cycle={cycle} tiles={tiles} packed={packed}", flush=True)
        z = g.precondition(r)
        x, info = g.solve(b, 1e-11, 3000)
        out.append((z, x, info["iterations"]))
    for k in (0, 1):
        assert np.array_equal(out[k][0], out[2][0]), (k, float(np.max(np.abs(out[k][0] - out[2][0]))))
        assert out[k][2] == out[2][2]
        assert np.array_equal(out[k][1], out[2][1])


@pytest.mark.parametrize("name,case", _ref_cases(), ids=[n for n, _ in _ref_cases()])
def test_reference_cycle_iterations_match_oracle(gpu_available, name, case):
    case.preconditioner = MULTIGRID
    o = oracle.Grid(case)
    rng = np.random.default_rng(5)
    ex = o.center_nullspaces(np.where(o.fluid(), rng.standard_normal(o.count), 0.0))
    b = o.apply(ex)
    xo, io = o.solve(b, 1e-12, 3000)
    g = _ref_grid(case)
    x, info = g.solve(b, 1e-12, 3000)
    assert abs(info["iterations"] - io["iterations"]) <= 1, (info, io)
    assert rel(x, xo) <= 1e-10


@pytest.mark.parametrize("cells,its", [([24, 2, 24], 6), ([48, 2, 48], 6), ([24, 24, 24], 7)])
def test_reference_cycle_published_iterations(gpu_available, cells, its):
    """Published multigrid PCG iteration counts of the perf-probe cavities
    (pressure-multigrid.md:34-38) at tol 1e-11."""
    case = W.cavity(cells, MULTIGRID)
    o, ex, b = _manufactured(case)
    g = _ref_grid(case)
    x, info = g.solve(b, 1e-11, 10000)
    assert abs(info["iterations"] - its) <= 1, info
    assert info["relative_residual"] <= 1.01e-10
    assert math.sqrt(np.mean((x - ex) ** 2)) <= 1e-8


def test_lid_cavity_48_ic0_matches_oracle(gpu_available):
    """The reference's CPU-runnable case: closed 48^3 cavity, IC(0)-PCG at 1e-10,
    right-hand side -div(u) of seeded face velocities."""
    case, b, t = W.lid_cavity_48()
    xo, io = oracle.Grid(case).solve_scaled(b, t, 1e-10, 1200)
    x, info = cfdgpu.Grid(case).solve_scaled(b, t, 1e-10, 1200)
    assert abs(info["iterations"] - io["iterations"]) <= 1, (info, io)
    assert rel(x, xo) <= 1e-10


# ---------------------------------------------------------------- anisotropy suite
def _aniso_shapes():
    import json
    import os

    with open(os.path.join(os.path.dirname(__file__), "golden", "anisotropy_shapes.json")) as f:
        return json.load(f)["shapes"]


@pytest.mark.parametrize("name", list(_aniso_shapes()))
def test_published_anisotropy_on_device(gpu_available, name):
    """The 18 published shapes on the device: IC(0)-PCG iterations at tol 1e-14
    (and, where the reference keeps point smoothing, the reference multigrid
    cycle's), recovery RMS <= 1e-8 and true residual <= 1.1e-12."""
    from test_oracle_goldens import aniso_case, aniso_rhs

    spec = _aniso_shapes()[name]
    for pc in ("ic0", "mg"):
        want = spec["ic0_iterations"] if pc == "ic0" else spec["mg_iterations"]
        case = aniso_case(spec, pc)
        g = cfdgpu.Grid(case)
        if pc == "mg":
            g.set_options(cfdgpu.PRECISION_FP64, cycle=cfdgpu.CYCLE_REFERENCE)
        ex, b = aniso_rhs(spec, g.count, g.apply)
        x, info = g.solve(b, 1e-14, 1200)
        assert abs(info["iterations"] - want) <= 1, (pc, info["iterations"], want)
        assert float(np.sqrt(np.mean((x - ex) ** 2))) <= 1e-8
        assert info["relative_residual"] <= 1.1e-12


@pytest.mark.parametrize("name", ["thin100", "cube100", "odd", "periodic", "outlet", "stiff-two-cell", "plate256",
                                  "short-stiff-axis"])
def test_reference_cycle_with_block_smoothers_matches_oracle(gpu_available, name):
    """Anisotropic shapes: the reference cycle with line / small-plane block
    smoothers and semicoarsening equals the oracle's cycle to 1e-11 (the
    centering means are tree sums on the device and sequential in the oracle;
    spacing ratios up to 1000:1 amplify that last-bit difference)."""
    from test_oracle_goldens import aniso_case

    spec = _aniso_shapes()[name]
    case = aniso_case(spec, "mg")
    o = oracle.Grid(case)
    rng = np.random.default_rng(31)
    r = o.center_nullspaces(rng.standard_normal(o.count))
    zo = o.mg_apply(r)
    z = _ref_grid(case).precondition(r)
    assert rel(z, zo) <= 1e-11, rel(z, zo)


ZEBRA_SHAPES = ["thin100", "cube100", "odd", "periodic", "outlet", "stiff-two-cell", "plate256", "short-stiff-axis",
                "isotropic"]


@pytest.mark.parametrize("name", ZEBRA_SHAPES)
def test_zebra_cycle_solves_the_anisotropy_suite(gpu_available, name):
    """The zebra cycle (the reference's hierarchy and block smoothers, swept by
    colour classes) is a different but symmetric smoother: its PCG solve reaches
    the published gates (recovery RMS <= 1e-8, true residual <= 1.1e-12 at tol
    1e-14) with an iteration count near the reference cycle's, and its
    V-cycle is symmetric on the compatible subspace (u . M v = v . M u)."""
    from test_oracle_goldens import aniso_case, aniso_rhs

    spec = _aniso_shapes()[name]
    case = aniso_case(spec, "mg")
    g = cfdgpu.Grid(case)
    g.set_options(cfdgpu.PRECISION_FP64, cycle=cfdgpu.CYCLE_ZEBRA)
    print(f"[zebra] {name} cells={spec['cells']}", flush=True)
    ex, b = aniso_rhs(spec, g.count, g.apply)
    x, info = g.solve(b, 1e-14, 1200)
    assert float(np.sqrt(np.mean((x - ex) ** 2))) <= 1e-8
    assert info["relative_residual"] <= 1.1e-12
    assert info["iterations"] <= 2 * spec["mg_iterations"] + 4, (info["iterations"], spec["mg_iterations"])
    rng = np.random.default_rng(5)
    u = g.center_nullspaces(rng.standard_normal(g.count))
    v = g.center_nullspaces(rng.standard_normal(g.count))
    a, c = float(u @ g.precondition(v)), float(v @ g.precondition(u))
    assert abs(a - c) <= 1e-10 * max(abs(a), abs(c)), (a, c)
    assert float(u @ g.precondition(u)) > 0.0


def test_zebra_cycle_on_a_masked_grid_matches_oracle(gpu_available):
    """With solids the zebra cycle sweeps the masked aggregation levels by colour
    classes (the fine level keeps the reference's order); the solution equals the
    oracle's to the solve tolerance."""
    case, b, t = W.urban_block(n=(32, 32, 16), boxes=6)
    o = oracle.Grid(case)
    xo, io = o.solve_scaled(b, t, 1e-12, 3000)
    g = cfdgpu.Grid(case)
    g.set_options(cfdgpu.PRECISION_FP64, cycle=cfdgpu.CYCLE_ZEBRA)
    x, info = g.solve_scaled(b, t, 1e-12, 3000)
    assert rel(x, xo) <= 1e-9, (info, io)
    with pytest.raises(cfdgpu.CfdError):
        g.set_options(cfdgpu.PRECISION_MIXED, cycle=cfdgpu.CYCLE_ZEBRA)


@pytest.mark.parametrize("kind", ["anchored", "closed"])
def test_device_inverted_block_is_the_exact_inverse(gpu_available, kind):
    """With the dense limit above the cell count the fine level is the coarsest
    level and one dense block per component solves it. Blocks above 128 nodes
    are inverted on the device by blocked Gauss-Jordan (pivot blocks of 64,
    here with a partial last block; a closed component's block centered), so
    one preconditioner application is A^-1 b itself: the oracle's manufactured
    field to rounding."""
    if kind == "anchored":
        case = Case(cells=[17, 13, 9], size_m=[17, 13, 9], boundaries=[OPEN, WALL, WALL], velocity_m_s=[1.0, 0, 0],
                    blocks=[Block([5.0, 4, 0], [8.0, 9, 6])], preconditioner=MULTIGRID)
    else:
        case = W.cavity([13, 11, 9], MULTIGRID)
    o = oracle.Grid(case)
    ex = o.center_nullspaces(np.where(o.fluid(), np.random.default_rng(2).standard_normal(o.count), 0.0))
    b = o.apply(ex)
    with env(CFD_MG_DENSE=4096):
        g = cfdgpu.Grid(case)
    print(f"[inverse] {kind} cells={case.cells}", flush=True)
    z = g.precondition(b)
    assert rel(z, ex) <= 1e-10, rel(z, ex)
    x, info = g.solve(b, 1e-12, 50)
    assert info["iterations"] <= 2, info
    assert rel(x, ex) <= 1e-10


def test_grid_signature_matches_like_the_reference(gpu_available):
    """Signature::matches (multigrid.rs:158-222): a context is reused for a case
    with the same operator and rebuilt when faces, weights, the mask or the
    preconditioner change (pressure.rs:69-92); the first solve after a rebuild
    reports hierarchy_builds = 1 and the next one 0."""
    import copy

    case, b, t = W.wind_campaign(200.0, n=(32, 32, 16), boxes=4)
    g = cfdgpu.Grid(case)
    print("[signature] wind 200", flush=True)
    assert g.matches(case)
    assert g.matches(W.wind_campaign(200.0, n=(32, 32, 16), boxes=4)[0])
    assert not g.matches(W.wind_campaign(90.0, n=(32, 32, 16), boxes=4)[0])  # other outlets
    fewer = copy.deepcopy(case)
    fewer.blocks = fewer.blocks[1:]
    assert not g.matches(fewer)  # another fluid mask
    stretched = copy.deepcopy(case)
    stretched.size_m = [v * 1.5 for v in case.size_m]
    assert not g.matches(stretched)  # same codes, other face weights
    ic0 = copy.deepcopy(case)
    ic0.preconditioner = IC0
    assert not g.matches(ic0)
    cache = cfdgpu.GridCache()
    first = cache.get(case)
    _, i1 = first.solve_scaled(b, t, 1e-8, 300)
    assert cache.get(case) is first
    _, i2 = first.solve_scaled(b, t, 1e-8, 300)
    other, bo, to = W.wind_campaign(90.0, n=(32, 32, 16), boxes=4)
    second = cache.get(other)
    assert second is not first
    _, i3 = second.solve_scaled(bo, to, 1e-8, 300)
    assert (i1["hierarchy_builds"], i2["hierarchy_builds"], i3["hierarchy_builds"]) == (1, 0, 1)


@pytest.mark.parametrize("cells,size,limit", [([48, 2, 48], None, 12), ([96, 2, 96], None, 14),
                                              ([96, 4, 96], None, 18), ([96, 16, 96], [1.0, 1.0, 0.1], 60)])
def test_red_black_cycle_on_stretched_cells(gpu_available, cells, size, limit):
    """Thin lid cavities (weak y couplings) and a 100:1 stiff axis: with
    semicoarsening the red-black cycle converges in a bounded number of PCG
    iterations (without it 47, 92, 86 and 291) and recovers the manufactured
    field; the plain 2x2x2 aggregation cannot represent errors that differ
    between weakly coupled layers."""
    case = W.cavity(cells, MULTIGRID)
    if size is not None:
        case.size_m = size
    g = cfdgpu.Grid(case)
    print(f"[stretched] {cells} {size}", flush=True)
    ex = g.center_nullspaces(W.manufactured(case, W.fluid_mask(case)))
    x, info = g.solve(g.apply(ex), 1e-11, 3000)
    assert info["iterations"] <= limit, info
    assert info["relative_residual"] <= 1.01e-10
    assert rel(x, ex) <= 1e-8


RESIDENT_CASES = [
    ("urban64", lambda: W.urban_block(n=(64, 48, 40), boxes=4)[0]),
    ("urban128", lambda: W.urban_block(n=(128, 40, 12), boxes=4)[0]),
    ("periodic-x", lambda: _with_wind(Case(cells=[32, 20, 16], size_m=[3.2, 2.0, 1.6],
                                           boundaries=[PERIODIC, OPEN, GROUND_OPEN_TOP],
                                           blocks=[Block([0.6, 0.5, 0.0], [1.4, 1.3, 0.9])],
                                           preconditioner=MULTIGRID), 180.0)),
    ("periodic-y", lambda: _with_wind(Case(cells=[32, 24, 20], size_m=[3.2, 2.4, 2.0],
                                           boundaries=[OPEN, PERIODIC, GROUND_OPEN_TOP],
                                           blocks=[Block([0.6, 0.5, 0.0], [1.4, 1.3, 0.9])],
                                           preconditioner=MULTIGRID), 270.0)),
    ("periodic-z", lambda: _with_wind(Case(cells=[36, 20, 16], size_m=[3.6, 2.0, 1.6],
                                           boundaries=[OPEN, WALL, PERIODIC],
                                           blocks=[Block([0.8, 0.4, 0.3], [1.6, 1.2, 0.9])],
                                           preconditioner=MULTIGRID), 270.0)),
]


def _with_wind(case, direction):
    """The log-law wind from `direction`: its outlet faces fix the pressure level
    (no closed component)."""
    W._neutral_rans(case, direction=direction)
    return case


@pytest.mark.parametrize("precision", [cfdgpu.PRECISION_FP64, cfdgpu.PRECISION_MIXED])
@pytest.mark.parametrize("name,make", RESIDENT_CASES, ids=[n for n, _ in RESIDENT_CASES])
def test_resident_residual_solves_bitwise(gpu_available, name, make, precision):
    """With two cells per thread and no closed component, the FP64 red-black
    solve keeps the PCG's r and z in the fine level's colour-separated arrays
    (the residual update scatters into them, the operator reads z from them,
    the first half-sweep starts from them and the last one forms r . z there)
    instead of natural-layout vectors; the mixed-precision solve keeps z as the
    FP32 values it is instead of an FP64 copy (its residual stays FP64). Every value and
    every partial sum keeps its arithmetic and order: the field, the iteration
    count and the residuals equal those of the natural-layout solve bit for bit,
    with the marching and the cell-by-cell operator kernels, pairs in tiles or
    linear, periodic axes included, and on a warm-started second solve.

    With linear pairs the mixed-precision solve also keeps its FP64 residual
    colour by colour and stores p in FP32 (each p value rounded once, A p and
    p . Ap formed from the rounded values): not bitwise the natural-layout
    solve, so it is held to the same iteration count within one and the same
    accuracy (both fields within 1e-9 of each other and 1e-8 of the exact
    field at rel-res 1e-11)."""
    case = make()
    o = oracle.Grid(case)
    rng = np.random.default_rng(21)
    ex = o.center_nullspaces(np.where(o.fluid(), rng.standard_normal(o.count), 0.0))
    b = o.apply(ex)
    b2 = o.apply(o.center_nullspaces(np.where(o.fluid(), rng.standard_normal(o.count), 0.0)))
    for march, lin in ((2, 1), (0, 1), (2, 0)):
        out = []
        for resident in (1, 0):
            with env(CFD_MG_RB_PAIRS=1, CFD_MG_RB_LIN=lin, CFD_MARCH=march, CFD_RB_RESIDENT=resident):
                g = cfdgpu.Grid(case)
            g.set_options(precision)
            print(f"[resident] {name} precision={precision} march={march} linear={lin} resident={resident}",
                  flush=True)
            d_b, d_x = cfdgpu.DeviceArray.from_host(b), cfdgpu.DeviceArray(case.count)
            if precision == cfdgpu.PRECISION_FP64:
                # the path taken: the first half-sweep of a cycle reads r in the
                # natural layout (21 bytes a cell) only in the set-up's cycle when
                # r is kept colour by colour (8.5 bytes a cell in the others)
                first = g.profile_solve(d_b, d_x, 1e-11, "fine_first_half_sweep")
                n = first["launches"]
                want = (21.0 + 8.5 * (n - 1)) if resident else 21.0 * n
                assert abs(first["bytes"] / case.count - want) < 1e-6 * n, (first, want)
            else:
                # the last half-sweep writes z as its FP32 values (4 bytes a
                # cell) instead of an FP64 copy (8), or (linear pairs) leaves z
                # in the level and reads r colour by colour (no z written)
                last = g.profile_solve(d_b, d_x, 1e-11, "fine_rz_half_sweep")
                n = last["launches"]
                want = (14.5 if lin else 18.5) * n if resident else 22.5 * n
                assert abs(last["bytes"] / case.count - want) < 1e-6 * n, (last, want)
            x, info = g.solve(b, 1e-11, 3000)
            g.set_options(precision, warm_start=True)
            xw, infow = g.solve(b2, 1e-11, 3000)
            out.append((x, info, xw, infow))
        (x1, i1, w1, j1), (x0, i0, w0, j0) = out
        if precision == cfdgpu.PRECISION_MIXED and lin:
            for a, c in ((i1, i0), (j1, j0)):
                assert abs(a["iterations"] - c["iterations"]) <= 1, (a, c)
                assert a["relative_residual"] <= 1e-10 + 1e-12, a
            assert rel(x1, x0) <= 1e-9 and rel(w1, w0) <= 1e-9, (rel(x1, x0), rel(w1, w0))
            assert rel(x1, ex) <= 1e-8
            continue
        for a, c in ((i1, i0), (j1, j0)):
            assert a["iterations"] == c["iterations"], (a, c)
            assert a["relative_residual"] == c["relative_residual"], (a, c)
        assert np.array_equal(x1, x0), float(np.max(np.abs(x1 - x0)))
        assert np.array_equal(w1, w0), float(np.max(np.abs(w1 - w0)))
        assert rel(x1, ex) <= 1e-8


@pytest.mark.parametrize("cells", [[64, 48, 40], [63, 20, 9], [128, 40, 12]])
@pytest.mark.parametrize("precision", [cfdgpu.PRECISION_FP64, cfdgpu.PRECISION_MIXED])
def test_two_cell_kernels_solve_like_one_cell(gpu_available, cells, precision):
    """Full solves with the fine half-sweeps and the final r . z half-sweep taking
    two cells per thread (forced; nx a multiple of 4, else one cell) take the
    iterations of the one-cell kernels and give the same field to rounding (only
    the association of the r . z partials differs)."""
    case, b, t = W.urban_block(n=tuple(cells), boxes=4)
    res = []
    for pairs in (1, 0):
        with env(CFD_MG_RB_PAIRS=pairs):
            g = cfdgpu.Grid(case)
        g.set_options(precision)
        print(f"[pairs] {cells} precision={precision} pairs={pairs}", flush=True)
        res.append(g.solve_scaled(b, t, 1e-10, 500))
    (x1, i1), (x0, i0) = res
    assert i1["iterations"] == i0["iterations"], (i1, i0)
    assert i1["relative_residual"] <= 1e-10 and i0["relative_residual"] <= 1e-10
    assert rel(x1, x0) <= 1e-12
