"""Queued HBM admission (cfd_server_admit_async / cfd_server_admitted).

Grids that do not fit the server's budget wait in ticket order and are built
by the server's builder thread when a release makes room; a queued grid is the
grid a direct admission builds (the same solve, bit for bit); a grid larger
than the whole budget and a case that cannot be built end with their errors;
a server closed with admissions still queued or built frees everything.
"""
import dataclasses
import time

import numpy as np
import pytest

import cfdgpu
from cfdgpu import workloads as W

pytestmark = pytest.mark.gpu


def _same_shape(n, shape=(64, 64, 64), seed=41):
    out, k = [], 0
    while len(out) < n:
        r = W.served_request(seed, k)
        k += 1
        if tuple(int(v) for v in r[0].cells) == shape:
            out.append(r)
    return out


def _wait(server, ticket, timeout=120.0):
    end = time.time() + timeout
    while True:
        g = server.admitted(ticket)
        if g is not None:
            return g
        assert time.time() < end, "queued admission not done"
        time.sleep(0.005)


def _solve(g, b, t):
    db, dt, dx = cfdgpu.DeviceArray.from_host(b), cfdgpu.DeviceArray.from_host(t), cfdgpu.DeviceArray(g.count)
    g.set_options(cfdgpu.PRECISION_FP64)
    info = g.solve_device(db, dx, 1e-8, 2000, d_scale=dt)
    return dx.numpy(), info["iterations"]


def test_queued_admissions_wait_for_room(gpu_available):
    reqs = _same_shape(4)
    twin = cfdgpu.Server(max_active=8, lanes=3)
    twins = [twin.admit(c) for c, _, _, _ in reqs]
    fp = [cfdgpu.lib().cfd_grid_footprint(g.h) for g in twins]
    want = [_solve(g, b, t) for g, (_, b, t, _) in zip(twins, reqs)]
    budget = int(2.5 * max(fp))
    server = cfdgpu.Server(max_active=8, hbm_budget_bytes=budget, lanes=3)
    tks = [server.admit_async(c) for c, _, _, _ in reqs]
    g0 = _wait(server, tks[0])
    g1 = _wait(server, tks[1])
    time.sleep(0.5)
    # two grids in, the third waits for room (and the fourth behind it)
    assert server.admitted(tks[2]) is None and server.admitted(tks[3]) is None
    used, cap = server.usage()
    assert cap == budget and used == fp[0] + fp[1]
    for g, (_, b, t, _), w in zip((g0, g1), reqs, want):
        x, its = _solve(g, b, t)
        assert its == w[1] and np.array_equal(x, w[0])
    server.release(g0)
    g2 = _wait(server, tks[2])
    time.sleep(0.3)
    assert server.admitted(tks[3]) is None
    assert server.usage()[0] == fp[1] + fp[2]
    server.release(g1)
    g3 = _wait(server, tks[3])
    assert server.usage()[0] == fp[2] + fp[3]
    for g, k in ((g2, 2), (g3, 3)):
        _, b, t, _ = reqs[k]
        x, its = _solve(g, b, t)
        assert its == want[k][1] and np.array_equal(x, want[k][0])
    # a queued grid serves requests like any other
    _, b, t, tol = reqs[3]
    db, dt, dx = cfdgpu.DeviceArray.from_host(b), cfdgpu.DeviceArray.from_host(t), cfdgpu.DeviceArray(g3.count)
    tk = server.submit(g3, db, dx, 1e-8, 2000, cfdgpu.PRECISION_FP64, d_scale=dt)
    server.wait(tk)
    assert server.result(tk)[0]["iterations"] == want[3][1] and np.array_equal(dx.numpy(), want[3][0])
    server.close()
    twin.close()


def test_queued_admission_errors(gpu_available):
    case, _, _, _ = W.served_request(3, 2)
    small = cfdgpu.Server(max_active=2, hbm_budget_bytes=1 << 20)
    tk = small.admit_async(case)
    with pytest.raises(cfdgpu.CfdError) as e:
        _wait(small, tk)
    assert e.value.code == 203 and "budget" in str(e.value)
    assert small.usage()[0] == 0
    bad = dataclasses.replace(case, preconditioner=7)
    tk = small.admit_async(bad)
    with pytest.raises(cfdgpu.CfdError) as e:
        _wait(small, tk)
    assert e.value.code == 200 and "preconditioner" in str(e.value)
    with pytest.raises(cfdgpu.CfdError):
        small.admitted(tk)  # forgotten once done
    small.close()


def test_server_closed_with_admissions_outstanding(gpu_available):
    reqs = _same_shape(3, seed=43)
    server = cfdgpu.Server(max_active=4, lanes=3)
    tks = [server.admit_async(c) for c, _, _, _ in reqs]
    _wait(server, tks[0])
    server.close()
