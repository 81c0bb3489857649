"""Serving: continuous batching of pressure solves on resident grids.

Each result must be the same alone and in the mix (the reductions have a fixed
shape, so bitwise), a request arriving while others run must start and finish,
and the HBM admission must refuse a grid that does not fit.
"""
import numpy as np
import pytest

import cfdgpu
from cfdgpu import workloads as W

pytestmark = pytest.mark.gpu


def _requests(n, seed=7):
    out = []
    for k in range(n):
        case, b, t, tol = W.served_request(seed, k)
        out.append((case, b, t, tol))
    return out


def test_results_alone_equal_results_in_the_mix(gpu_available):
    reqs = _requests(6)
    server = cfdgpu.Server(max_active=4)
    grids, bufs = [], []
    for case, b, t, tol in reqs:
        g = server.admit(case)
        grids.append(g)
        bufs.append((cfdgpu.DeviceArray.from_host(b), cfdgpu.DeviceArray.from_host(t), cfdgpu.DeviceArray(g.count)))
    alone = []
    for (case, b, t, tol), g, (db, dt, dx) in zip(reqs, grids, bufs):
        tk = server.submit(g, db, dx, tol, 2000, cfdgpu.PRECISION_MIXED, d_scale=dt)
        server.wait(tk)
        info, lat, q = server.result(tk)
        alone.append((dx.numpy(), info["iterations"]))
    tickets = []
    for rep in range(2):  # every grid twice: the second waits for its grid to be free
        for (case, b, t, tol), g, (db, dt, dx) in zip(reqs, grids, bufs):
            tickets.append((server.submit(g, db, dx, tol, 2000, cfdgpu.PRECISION_MIXED, d_scale=dt), g))
    server.drain()
    for k, (tk, g) in enumerate(tickets):
        info, lat, q = server.result(tk)
        i = k % len(reqs)
        assert info["iterations"] == alone[i][1]
        assert info["relative_residual"] <= 10 * reqs[i][3] + 1e-12
    for i, (db, dt, dx) in enumerate(bufs):
        assert np.array_equal(dx.numpy(), alone[i][0])
    server.close()


def test_mixed_precision_served_matches_plain_fp64(gpu_available):
    case, b, t, tol = W.served_request(11, 0)
    server = cfdgpu.Server(max_active=2)
    g = server.admit(case)
    db, dt, d1, d2 = (cfdgpu.DeviceArray.from_host(b), cfdgpu.DeviceArray.from_host(t), cfdgpu.DeviceArray(g.count),
                      cfdgpu.DeviceArray(g.count))
    t1 = server.submit(g, db, d1, 1e-10, 2000, cfdgpu.PRECISION_FP64, d_scale=dt)
    t2 = server.submit(g, db, d2, 1e-10, 2000, cfdgpu.PRECISION_MIXED, d_scale=dt)
    server.drain()
    i1, _, _ = server.result(t1)
    i2, _, _ = server.result(t2)
    x1, x2 = d1.numpy(), d2.numpy()
    assert i1["relative_residual"] <= 1.01e-9 and i2["relative_residual"] <= 1.01e-9
    assert np.linalg.norm(x1 - x2) / np.linalg.norm(x1) < 1e-7
    server.close()


def test_hbm_admission_refuses_what_does_not_fit(gpu_available):
    case, b, t, tol = W.served_request(3, 2)
    server = cfdgpu.Server(max_active=2, hbm_budget_bytes=1 << 20)
    with pytest.raises(cfdgpu.CfdError) as e:
        server.admit(case)
    assert e.value.code == 203 and "admission" in str(e.value)
    big = cfdgpu.Server(max_active=2, hbm_budget_bytes=1 << 40)
    g = big.admit(case)
    used, budget = big.usage()
    assert used > 0 and used == cfdgpu.lib().cfd_grid_footprint(g.h)
    big.release(g)
    assert big.usage()[0] == 0
    big.close()
    server.close()


def test_warm_start_reuses_the_resident_solution(gpu_available):
    case, b, t, tol = W.served_request(5, 1)
    g = cfdgpu.Grid(case)
    g.set_options(cfdgpu.PRECISION_MIXED, warm_start=True)
    db, dt, dx = cfdgpu.DeviceArray.from_host(b), cfdgpu.DeviceArray.from_host(t), cfdgpu.DeviceArray(g.count)
    cold = g.solve_device(db, dx, 1e-8, 2000, d_scale=dt)
    rng = np.random.default_rng(0)
    b2 = b + 0.02 * np.sqrt(np.mean(b**2)) * rng.standard_normal(b.size) * (b != 0)
    db2 = cfdgpu.DeviceArray.from_host(b2)
    warm = g.solve_device(db2, dx, 1e-8, 2000, d_scale=dt)
    assert warm["relative_residual"] <= 1.01e-7
    assert warm["iterations"] <= cold["iterations"]


def test_admitted_grid_takes_options_and_profiles(gpu_available):
    """A grid admitted by the server is a full Grid: options, direct solves with a
    profiling control, and the profile checker's rules."""
    from profile_rules import check

    case, b, t, tol = W.served_request(3, 2)
    server = cfdgpu.Server(max_active=2)
    g = server.admit(case)
    g.set_options(cfdgpu.PRECISION_FP64, cycle=cfdgpu.CYCLE_RB)
    ctl = cfdgpu.Control(profile=True)
    x, info = g.solve_scaled(b, t, tol, 2000, control=ctl)
    assert info["relative_residual"] <= 10 * tol + 1e-12
    check(ctl.profile(), "multigrid" if case.preconditioner == cfdgpu.MULTIGRID else "ic0")
    server.release(g)
    server.close()
