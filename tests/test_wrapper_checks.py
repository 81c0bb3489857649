"""The Python host refuses buffers of the wrong length or type before anything
reaches the library: every entry point reads and writes exactly one value per
cell (or per face) through the pointers it is given. Runs without a device: the
checks come before the first library call."""
import numpy as np
import pytest

import cfdgpu


def fake_grid(count=24, faces=80):
    g = cfdgpu.Grid.__new__(cfdgpu.Grid)
    g.h = None
    g.count = count
    g.face_count = faces
    g.tolerance = 1e-10
    g.max_iterations = 100
    return g


def fake_device(n, dtype=np.float64):
    d = cfdgpu.DeviceArray.__new__(cfdgpu.DeviceArray)
    d.n = n
    d.dtype = np.dtype(dtype)
    d.nbytes = n * d.dtype.itemsize
    d.ptr = object()  # never dereferenced: the check fails first
    return d


@pytest.mark.parametrize("call", ["apply", "precondition", "center_nullspaces", "solve", "solve_scaled"])
def test_host_vectors_of_the_wrong_length_are_refused(call):
    g = fake_grid()
    for bad in (np.zeros(23), np.zeros(25), np.zeros((4, 6))):
        with pytest.raises(ValueError):
            getattr(g, call)(bad)
    if call == "solve_scaled":
        with pytest.raises(ValueError):
            g.solve_scaled(np.zeros(24), np.zeros(12))


def test_face_and_pressure_vectors_are_checked():
    g = fake_grid()
    with pytest.raises(ValueError):
        g.projection_rhs(np.zeros(79), 1.0)
    with pytest.raises(ValueError):
        g.project(np.zeros(80), np.zeros(23), 1.0)
    with pytest.raises(ValueError):
        g.project(np.zeros(81), np.zeros(24), 1.0)


def test_device_arrays_of_the_wrong_length_or_type_are_refused():
    g = fake_grid()
    good = fake_device(24)
    for bad in (fake_device(23), fake_device(24, np.float32), fake_device(48, np.float32), np.zeros(24)):
        with pytest.raises(ValueError):
            g.solve_device(bad, good)
        with pytest.raises(ValueError):
            g.solve_device(good, bad)
        with pytest.raises(ValueError):
            g.solve_device(good, good, d_scale=bad)
    freed = fake_device(24)
    freed.ptr = None
    with pytest.raises(ValueError):
        g.solve_device(freed, good)


def test_server_submit_checks_against_the_grid():
    s = cfdgpu.Server.__new__(cfdgpu.Server)
    s.h = None
    g = fake_grid()
    with pytest.raises(ValueError):
        s.submit(g, fake_device(23), fake_device(24), 1e-8)
    with pytest.raises(ValueError):
        s.submit(g, fake_device(24), fake_device(24, np.float32), 1e-8)
