"""Pins the oracle where no published number reaches: closed forms, dense and
pivoted references on tiny grids, and the invariants of the reference's own
unit tests (restated here on the same tiny fixtures).

  * 7-point stencil with 1/dx^2 couplings and 2/dx^2 at outlets
    (tests/numerics.rs:201-222, tests/regressions.rs:410-548);
  * dense IC(0) of the shifted matrix (pressure.rs:530-581);
  * pivoted dense solve with one gauge cell per closed component
    (tests/urban_pressure_reference.rs:80-145);
  * manufactured recovery and projection (tests/multigrid.rs:105-129);
  * second-order convergence (tests/multigrid.rs:131-156);
  * the V-cycle is symmetric and positive definite on the compatible subspace
    (multigrid.rs:845-923, masked.rs:509-616);
  * CFD-E031 / E033 (tests/multigrid.rs:158-174, 266-295).
"""
import math

import numpy as np
import pytest

import oracle
from cfdgpu.case import IC0, MULTIGRID, OPEN, OUTLET, PERIODIC, WALL, INTERIOR, Block, Case


def lid(n, bnd, size=(2.7, 1.9, 3.1), pc=IC0, blocks=()):
    return Case(cells=list(n), size_m=list(size), boundaries=[bnd, PERIODIC, WALL], velocity_m_s=[1.0, 0.0, 0.0],
                preconditioner=pc, blocks=list(blocks))


TINY = [([9, 2, 7], WALL), ([9, 2, 7], OPEN), ([8, 5, 6], PERIODIC), ([2, 2, 2], WALL), ([9, 1, 7], OPEN),
        ([6, 4, 5], OPEN), ([4, 6, 3], PERIODIC)]


def closed_form_matrix(case, fluid):
    """A from the mask alone: 1/dx^2 between face neighbours, 2/dx^2 on outlet faces."""
    n = case.cells
    dx = case.dx
    N = case.count
    A = np.zeros((N, N))
    idx = lambda i, j, k: (k * n[1] + j) * n[0] + i  # noqa: E731
    for k in range(n[2]):
        for j in range(n[1]):
            for i in range(n[0]):
                c = idx(i, j, k)
                if not fluid[c]:
                    continue
                q = [i, j, k]
                for a in range(3):
                    for s in (-1, 1):
                        t = q[a] + s
                        if 0 <= t < n[a] or case.boundaries[a] == PERIODIC:
                            qq = list(q)
                            qq[a] = t % n[a]
                            o = idx(*qq)
                            if o != c and fluid[o]:
                                A[c, c] += 1.0 / dx[a] ** 2
                                A[c, o] -= 1.0 / dx[a] ** 2
                        elif case.boundaries[a] == OPEN:
                            v = case.flow_vector(0.5)[a] * s
                            if v > 1e-12:
                                A[c, c] += 2.0 / dx[a] ** 2
    return A


def dense_from_faces(g):
    f = g.faces()
    A = np.zeros((g.count, g.count))
    for lo, hi, kind, w in zip(f["lo"], f["hi"], f["kind"], f["weight"]):
        if kind == INTERIOR:
            A[lo, lo] += w
            A[hi, hi] += w
            A[lo, hi] -= w
            A[hi, lo] -= w
        elif kind == OUTLET:
            i = lo if lo >= 0 else hi
            A[i, i] += w
    return A


@pytest.mark.parametrize("n,bnd", TINY)
def test_operator_is_the_closed_form_stencil(n, bnd):
    case = lid(n, bnd)
    g = oracle.Grid(case)
    A = closed_form_matrix(case, g.fluid())
    for j in range(g.count):
        e = np.zeros(g.count)
        e[j] = 1.0
        np.testing.assert_allclose(g.apply(e), A[:, j], rtol=1e-12, atol=1e-12 * np.abs(A).max())
        assert np.array_equal(g.apply(e), g.operator_apply(e))


@pytest.mark.parametrize("n,bnd", TINY + [([9, 2, 5], OPEN)])
def test_ic0_matches_dense_incomplete_cholesky(n, bnd):
    blocks = [Block([3.0, 0.0, 0.0], [4.0, 2.0, 5.0])] if n == [9, 2, 5] else []
    case = lid(n, bnd, size=n if blocks else (2.7, 1.9, 3.1), blocks=blocks)
    g = oracle.Grid(case)
    A = dense_from_faces(g)
    N = g.count
    maxdiag = max(1.0, float(np.max(np.diag(A))))
    M = A + 1e-6 * maxdiag * np.eye(N)
    pattern = (np.abs(A) > 0) | np.eye(N, dtype=bool)
    L = np.zeros((N, N))
    for i in range(N):
        for j in range(i):
            if pattern[i, j]:
                L[i, j] = (M[i, j] - L[i, :j] @ L[j, :j]) / L[j, j]
        L[i, i] = math.sqrt(max(M[i, i] - L[i, :i] @ L[i, :i], 1e-12 * maxdiag))
    np.testing.assert_allclose(g.ic0_diagonal(), np.diag(L), rtol=1e-13)
    r = np.sin(0.7 * np.arange(N))
    y = np.linalg.solve(L, r)
    z = np.linalg.solve(L.T, y)
    np.testing.assert_allclose(g.ic0_apply(r), z, rtol=1e-10, atol=1e-10 * np.abs(z).max())


def pivot_reference(g, A, b):
    comp, anch = g.components()
    gauges = [int(np.nonzero(comp == c)[0].min()) for c in range(len(anch)) if not anch[c]]
    active = [i for i in range(g.count) if comp[i] >= 0 and i not in gauges]
    x = np.zeros(g.count)
    x[active] = np.linalg.solve(A[np.ix_(active, active)], b[active])
    return g.center_nullspaces(x)


MASKED = [
    ("barrier-open", Case(cells=[9, 2, 5], size_m=[9, 2, 5], boundaries=[OPEN, WALL, WALL], velocity_m_s=[1.0, 0, 0],
                          blocks=[Block([3.0, 0, 0], [4.0, 2, 5])])),
    ("two-rooms", Case(cells=[9, 3, 5], size_m=[9, 3, 5], boundaries=[WALL, WALL, WALL],
                       blocks=[Block([4.0, 0, 0], [5.0, 3, 5])])),
    ("pocket", Case(cells=[7, 7, 4], size_m=[7, 7, 4], boundaries=[OPEN, WALL, WALL], velocity_m_s=[1.0, 0, 0],
                    blocks=[Block([2.0, 2, 0], [5.0, 3, 4]), Block([2.0, 4, 0], [5.0, 5, 4]), Block([2.0, 3, 0], [3.0, 4, 4]),
                            Block([4.0, 3, 0], [5.0, 4, 4])])),
]


@pytest.mark.parametrize("name,case", MASKED + [(f"lid{n}{b}", lid(n, b)) for n, b in TINY[:4]])
@pytest.mark.parametrize("pc", [IC0, MULTIGRID])
def test_solve_matches_pivoted_dense_reference(name, case, pc):
    case.preconditioner = pc
    g = oracle.Grid(case)
    A = dense_from_faces(g)
    fluid = g.fluid()
    exact = g.center_nullspaces(np.where(fluid, np.sin(0.19 * np.arange(g.count)), 0.0))
    b = A @ exact
    ref = pivot_reference(g, A, b)
    x, info = g.solve(b, 1e-13, 2000)
    assert info["relative_residual"] < 2e-12
    assert np.max(np.abs(x - ref)) < 2e-10
    assert np.max(np.abs(x - exact)) < 2e-10


@pytest.mark.parametrize("n,bnd", TINY)
@pytest.mark.parametrize("pc", [IC0, MULTIGRID])
def test_manufactured_solve_and_projection(n, bnd, pc):
    case = lid(n, bnd, pc=pc)
    g = oracle.Grid(case)
    exact = g.center_nullspaces(np.sin(0.37 * np.arange(g.count)))
    b = -g.divergence(g.gradient(exact))
    x, info = g.solve(b, 1e-12, 2000)
    assert info["relative_residual"] <= 1.1e-11
    assert np.max(np.abs(x - exact)) < 1e-9
    dt = 0.125
    u, p, _ = g.project(g.gradient(exact), np.zeros(g.count), dt, 1e-12, 2000)
    assert np.max(np.abs(u)) < 1e-8
    assert np.max(np.abs(p - exact / dt)) < 1e-8
    assert np.max(np.abs(g.divergence(u))) < 1e-7


def test_second_order_convergence():
    errors = []
    for nx in (12, 24, 48):
        case = Case(cells=[nx, 2, 4], size_m=[1.0, 1.0, 1.0], boundaries=[PERIODIC, PERIODIC, WALL])
        g = oracle.Grid(case)
        i = np.arange(g.count)
        exact = np.sin(2.0 * math.pi * ((i % nx) + 0.5) / nx)
        x, _ = g.solve(4.0 * math.pi**2 * exact, 1e-12, 2000)
        errors.append(math.sqrt(np.mean((x - exact) ** 2)))
    for a, b in zip(errors, errors[1:]):
        assert math.log2(a / b) > 1.99


CYCLE = [([5, 3, 3], WALL), ([5, 3, 3], OPEN), ([5, 1, 4], PERIODIC), ([6, 4, 5], OPEN)]


@pytest.mark.parametrize("name,case", MASKED + [(f"lid{n}{b}", lid(n, b)) for n, b in CYCLE])
def test_vcycle_is_symmetric_positive_on_compatible_subspace(name, case):
    case.preconditioner = MULTIGRID
    g = oracle.Grid(case)
    N = g.count
    comp, anch = g.components()
    M = np.zeros((N, N))
    for j in range(N):
        e = np.zeros(N)
        e[j] = 1.0
        M[:, j] = g.mg_apply(g.center_nullspaces(e))
    scale = np.abs(M).max()
    np.testing.assert_allclose(M, M.T, atol=2e-11 * scale)
    # basis: every fluid cell of an anchored component, differences to the first
    # cell of a closed one
    basis = []
    for c in range(len(anch)):
        members = np.nonzero(comp == c)[0]
        for i in members[(0 if anch[c] else 1):]:
            v = np.zeros(N)
            v[i] = 1.0
            if not anch[c]:
                v[members[0]] = -1.0
            basis.append(v)
    B = np.array(basis).T
    S = B.T @ M @ B
    assert np.linalg.eigvalsh(0.5 * (S + S.T)).min() > 0.0


def test_incompatible_flux_and_iteration_limit():
    g = oracle.Grid(lid([9, 2, 7], WALL))
    b = g.center_nullspaces(np.cos(0.3 * np.arange(g.count))) + 1e-11
    g.solve(b, 1e-12, 2000)
    with pytest.raises(oracle.OracleError) as e:
        g.solve(b + 0.1, 1e-12, 2000)
    assert e.value.code == 31
    with pytest.raises(oracle.OracleError) as e:
        g.solve(b, 1e-12, 1)
    assert e.value.code == 33
