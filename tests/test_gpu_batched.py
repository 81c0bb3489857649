"""Batched serving: the PCG iterations of several solves, each on its own grid,
go out as one launch per stage of the iteration (launch_batch_iteration).

Each solve in a batch owns the blocks its single-grid launch would have, so it
must give exactly the bits that its grid gives alone through solve_device
(the same per-cell arithmetic and the same fixed-shape reductions), whatever
else is in the batch and whenever it joins: in FP64 and in mixed precision,
cold and warm-started, with a zero right-hand side or an iteration limit among
the others, and with requests that cannot be batched (another preconditioner,
nx not a multiple of 4) falling back to their own streams.
"""
import dataclasses

import numpy as np
import pytest

import cfdgpu
from cfdgpu import workloads as W
from cfdgpu.case import IC0, MULTIGRID, Case, Block, OPEN, SLIP, GROUND_OPEN_TOP

pytestmark = pytest.mark.gpu

PRECISIONS = [cfdgpu.PRECISION_FP64, cfdgpu.PRECISION_MIXED]


def _pool(n, seed=7):
    return [W.served_request(seed, k) for k in range(n)]


def _solo(g, db, dt, tol, precision, max_iterations=2000, warm=False):
    dx = cfdgpu.DeviceArray(g.count)
    g.set_options(precision, warm_start=warm)
    info = g.solve_device(db, dx, tol, max_iterations, d_scale=dt)
    return dx.numpy(), info["iterations"]


@pytest.mark.parametrize("precision", PRECISIONS, ids=["fp64", "mixed"])
def test_batched_solves_equal_solo_solves_bitwise(gpu_available, precision):
    reqs = _pool(7)
    server = cfdgpu.Server(max_active=8)
    grids = [server.admit(c) for c, _, _, _ in reqs]
    bufs = [(cfdgpu.DeviceArray.from_host(b), cfdgpu.DeviceArray.from_host(t)) for _, b, t, _ in reqs]
    solo = [_solo(g, db, dt, tol, precision) for g, (db, dt), (_, _, _, tol) in zip(grids, bufs, reqs)]
    outs = [cfdgpu.DeviceArray(g.count) for g in grids]
    tickets = [server.submit(g, db, dx, tol, 2000, precision, d_scale=dt)
               for g, (db, dt), dx, (_, _, _, tol) in zip(grids, bufs, outs, reqs)]
    server.drain()
    for k, tk in enumerate(tickets):
        info, lat, q = server.result(tk)
        assert info["iterations"] == solo[k][1], (k, info["iterations"], solo[k][1])
        assert info["relative_residual"] <= 10 * reqs[k][3] + 1e-12
        x = outs[k].numpy()
        assert np.array_equal(x, solo[k][0]), (k, float(np.max(np.abs(x - solo[k][0]))))
    server.close()


def test_requests_joining_a_running_batch(gpu_available):
    """Requests submitted one by one while the batch iterates (both precisions,
    the tolerances of the served mix): each joins at the next iteration and
    ends with its solo bits."""
    reqs = _pool(6, seed=19)
    server = cfdgpu.Server(max_active=8)
    grids = [server.admit(c) for c, _, _, _ in reqs]
    bufs = [(cfdgpu.DeviceArray.from_host(b), cfdgpu.DeviceArray.from_host(t)) for _, b, t, _ in reqs]
    precs = [PRECISIONS[k % 2] for k in range(len(reqs))]
    solo = [_solo(g, db, dt, tol, p) for g, (db, dt), (_, _, _, tol), p in zip(grids, bufs, reqs, precs)]
    outs = [cfdgpu.DeviceArray(g.count) for g in grids]
    tickets = []
    for k, (g, (db, dt), dx, (_, _, _, tol)) in enumerate(zip(grids, bufs, outs, reqs)):
        tickets.append(server.submit(g, db, dx, tol, 2000, precs[k], d_scale=dt))
        for _ in range(3 + k):
            server.progress()
    server.drain()
    for k, tk in enumerate(tickets):
        info, _, _ = server.result(tk)
        assert info["iterations"] == solo[k][1]
        assert np.array_equal(outs[k].numpy(), solo[k][0])
    server.close()


def test_zero_rhs_and_iteration_limit_inside_a_batch(gpu_available):
    reqs = _pool(4, seed=23)
    server = cfdgpu.Server(max_active=8)
    grids = [server.admit(c) for c, _, _, _ in reqs]
    bufs = [(cfdgpu.DeviceArray.from_host(b), cfdgpu.DeviceArray.from_host(t)) for _, b, t, _ in reqs]
    solo = [_solo(g, db, dt, tol, cfdgpu.PRECISION_FP64) for g, (db, dt), (_, _, _, tol) in zip(grids, bufs, reqs)]
    outs = [cfdgpu.DeviceArray(g.count) for g in grids]
    zero = cfdgpu.DeviceArray.from_host(np.zeros(grids[0].count))
    t_zero = server.submit(grids[0], zero, outs[0], 1e-8, 2000, cfdgpu.PRECISION_FP64)
    t_limit = server.submit(grids[1], bufs[1][0], outs[1], 1e-12, 2, cfdgpu.PRECISION_MIXED, d_scale=bufs[1][1])
    rest = [server.submit(grids[k], bufs[k][0], outs[k], reqs[k][3], 2000, cfdgpu.PRECISION_FP64,
                          d_scale=bufs[k][1]) for k in (2, 3)]
    server.drain()
    info, _, _ = server.result(t_zero)
    assert info["iterations"] == 0 and not np.any(outs[0].numpy())
    with pytest.raises(cfdgpu.CfdError) as e:
        server.result(t_limit)
    assert e.value.code == 33
    for k, tk in zip((2, 3), rest):
        info, _, _ = server.result(tk)
        assert info["iterations"] == solo[k][1]
        assert np.array_equal(outs[k].numpy(), solo[k][0])
    server.close()


def test_warm_started_chain_in_a_batch(gpu_available):
    """Two grids of the same case, warm start on: the chain b1, b2 solved through
    solve_device on one and through the batching server on the other."""
    case, b, t, tol = W.served_request(29, 1)
    rng = np.random.default_rng(3)
    b2 = b + 0.02 * np.sqrt(np.mean(b**2)) * rng.standard_normal(b.size) * (b != 0)
    server = cfdgpu.Server(max_active=4)
    g_solo, g_batch = server.admit(case), server.admit(case)
    dt = cfdgpu.DeviceArray.from_host(t)
    dbs = [cfdgpu.DeviceArray.from_host(v) for v in (b, b2)]
    for precision in PRECISIONS:
        g_batch.set_options(precision, warm_start=True)
        xs = [_solo(g_solo, db, dt, 1e-8, precision, warm=True) for db in dbs]
        for k, db in enumerate(dbs):
            dx = cfdgpu.DeviceArray(g_batch.count)
            tk = server.submit(g_batch, db, dx, 1e-8, 2000, precision, d_scale=dt)
            server.wait(tk)
            info, _, _ = server.result(tk)
            assert info["iterations"] == xs[k][1]
            assert np.array_equal(dx.numpy(), xs[k][0])
    server.close()


def test_unbatchable_requests_fall_back_beside_a_batch(gpu_available):
    """An IC(0) grid and a grid with nx = 66 (not a multiple of 4) run on their
    own streams while a batch runs; every result equals its solo result."""
    reqs = _pool(3, seed=31)
    odd = Case(cells=[66, 48, 32], size_m=[132.0, 96.0, 64.0], boundaries=[OPEN, SLIP, GROUND_OPEN_TOP],
               blocks=[Block([40.0, 30.0, 0.0], [60.0, 50.0, 30.0])], preconditioner=MULTIGRID)
    ic = dataclasses.replace(reqs[0][0], preconditioner=IC0)
    cases = [c for c, _, _, _ in reqs] + [odd, ic]
    server = cfdgpu.Server(max_active=8)
    grids = [server.admit(c) for c in cases]
    rng = np.random.default_rng(5)
    rhs = [b for _, b, _, _ in reqs]
    for g, c in zip(grids[len(reqs):], (odd, ic)):
        # compatible with every closed component (the odd grid has no wind: its domain is closed)
        rhs.append(g.center_nullspaces(np.where(W.fluid_mask(c), rng.standard_normal(c.count), 0.0)))
    tols = [tol for _, _, _, tol in reqs] + [1e-8, 1e-8]
    dbs = [cfdgpu.DeviceArray.from_host(v) for v in rhs]
    solo = [_solo(g, db, None, tol, cfdgpu.PRECISION_FP64, 3000) for g, db, tol in zip(grids, dbs, tols)]
    outs = [cfdgpu.DeviceArray(g.count) for g in grids]
    tickets = [server.submit(g, db, dx, tol, 3000, cfdgpu.PRECISION_FP64) for g, db, dx, tol in
               zip(grids, dbs, outs, tols)]
    server.drain()
    for k, tk in enumerate(tickets):
        info, _, _ = server.result(tk)
        assert info["iterations"] == solo[k][1]
        assert np.array_equal(outs[k].numpy(), solo[k][0])
    server.close()


def test_campaign_perturbation_is_seeded_and_confined_to_fluid(gpu_available):
    """cfd_grid_perturb (the campaign's chained right-hand sides): solid cells
    keep their values, the fluid cells move by rel * rms(src) times a standard
    normal field that the seed fixes (same seed: the same bits; another seed:
    another field), src == dst allowed."""
    case, b, t, tol = W.served_request(37, 2)
    g = cfdgpu.Grid(case)
    fluid = W.fluid_mask(case)
    src = cfdgpu.DeviceArray.from_host(b)
    d1, d2, d3 = (cfdgpu.DeviceArray(g.count) for _ in range(3))
    g.perturb(src, d1, 0.02, 7)
    g.perturb(src, d2, 0.02, 7)
    g.perturb(src, d3, 0.02, 8)
    x1, x2, x3 = d1.numpy(), d2.numpy(), d3.numpy()
    assert np.array_equal(x1, x2) and not np.array_equal(x1, x3)
    assert np.array_equal(x1[~fluid], b[~fluid])
    rms = np.sqrt(np.mean(b[fluid] ** 2))
    xi = (x1[fluid] - b[fluid]) / (0.02 * rms)
    n = xi.size
    assert abs(xi.mean()) < 5 / np.sqrt(n) and abs(xi.std() - 1.0) < 5 / np.sqrt(n), (xi.mean(), xi.std())
    # a fourth moment of 3 (normal, not uniform)
    assert abs(np.mean(xi ** 4) - 3.0) < 0.1
    g.perturb(src, src, 0.02, 7)
    assert np.array_equal(src.numpy(), x1)
