"""The grid context from a caller's built Grid (cfd_grid_create_desc): the
pressure path reads the Grid's faces (Operator::build, pressure.rs:255-279),
whose area and distance a caller may edit after Grid::build (pressure.rs:244).

Checked without a device (the descriptor is read and refused before any device
call): a face out of Grid::faces order, a face list of the wrong length, an
interior face without two fluid cells, an outlet inside the domain and
components that differ from the faces are CFD_EINVAL; a pressure weight that
differs within an axis and face kind is CFD_EUNSUPPORTED naming the face (the
device operator holds one interior and one outlet weight per axis). On the
device (tests/test_gpu_descriptor.py) a descriptor grid solves like the case's
grid, and the reference's weight-edit test is restated."""
import re

import numpy as np
import pytest

import cfdgpu
import oracle
from cfdgpu.case import MULTIGRID, OPEN, PERIODIC, WALL, Block, Case

EINVAL, EUNSUPPORTED = 200, 201


def outlet_case(cells=(9, 2, 7)):
    """tests/multigrid.rs:12-27: the lid cavity with an open x axis and the
    laminar initial velocity (1, 0, 0), so the x-high face is an outlet."""
    return Case(cells=list(cells), size_m=[1.0, 1.0, 1.0], boundaries=[OPEN, PERIODIC, WALL],
                velocity_m_s=(1.0, 0.0, 0.0), preconditioner=MULTIGRID)


def descriptor(case, with_components=True):
    """The Grid of `case` as the reference builds it (the oracle's faces, with the
    area and distance of grid.rs:341-429: the cell face area, the cell width for
    interior faces and half of it otherwise)."""
    o = oracle.Grid(case)
    f = o.faces()
    dx = case.dx
    volume = dx[0] * dx[1] * dx[2]
    faces = np.zeros(o.face_count, cfdgpu.FACE_DTYPE)
    for k in ("axis", "kind", "lo", "hi"):
        faces[k] = f[k]
    d = np.asarray(dx)[f["axis"]]
    faces["area"] = volume / d
    faces["distance"] = np.where(f["kind"] == cfdgpu.FACE_INTERIOR, d, 0.5 * d)
    coupled = (f["kind"] == cfdgpu.FACE_INTERIOR) | (f["kind"] == cfdgpu.FACE_OUTLET)
    # the weights the descriptor gives are the oracle's, bit for bit
    w = faces["area"] / (faces["distance"] * volume)
    assert np.array_equal(w[coupled], f["weight"][coupled])
    comp, anch = o.components() if with_components else (None, None)
    return cfdgpu.GridDescriptor(case.cells, dx, volume, [b == PERIODIC for b in case.boundaries], o.fluid(), faces,
                                 MULTIGRID, comp, None if anch is None else anch.astype(np.uint8))


def refused(desc):
    with pytest.raises(cfdgpu.CfdError) as e:
        cfdgpu.Grid.from_descriptor(desc)
    return e.value


def named_faces(e):
    m = re.search(r"faces (\d+) and (\d+)", str(e))
    return {int(m.group(1)), int(m.group(2))} if m else set()


def first(desc, kind, axis=None):
    f = desc.faces
    sel = f["kind"] == kind
    if axis is not None:
        sel &= f["axis"] == axis
    return int(np.flatnonzero(sel)[0])


def test_edited_interior_weight_is_refused_naming_the_face():
    d = descriptor(outlet_case())
    i = first(d, cfdgpu.FACE_INTERIOR)
    d.faces["distance"][i] *= 1.0 + 1e-15  # a few-ulp weight change (tests/multigrid.rs:254)
    e = refused(d)
    assert e.code == EUNSUPPORTED and i in named_faces(e), str(e)


def test_edited_outlet_weight_is_refused():
    d = descriptor(outlet_case())
    i = first(d, cfdgpu.FACE_OUTLET)
    d.faces["area"][i] *= 2.0  # tests/multigrid.rs:257
    e = refused(d)
    assert e.code == EUNSUPPORTED and i in named_faces(e) and "outlet" in str(e), str(e)


def test_faces_out_of_order_or_missing_are_refused():
    d = descriptor(outlet_case())
    nx = int(np.count_nonzero(d.faces["axis"] == 0))
    swapped = d.faces.copy()
    swapped[[nx, nx + 1]] = swapped[[nx + 1, nx]]  # two y faces
    bad = descriptor(outlet_case())
    bad.faces = swapped
    e = refused(bad)
    assert e.code == EINVAL and "not the Grid's" in str(e)
    short = descriptor(outlet_case())
    short.faces = short.faces[:-1]
    assert refused(short).code == EINVAL


def test_inconsistent_kinds_and_components_are_refused():
    d = descriptor(outlet_case())
    w = first(d, cfdgpu.FACE_WALL, axis=2)  # a domain face of a fluid cell
    d.faces["kind"][w] = cfdgpu.FACE_INTERIOR
    assert refused(d).code == EINVAL
    d = descriptor(outlet_case())
    i = first(d, cfdgpu.FACE_INTERIOR)
    d.faces["kind"][i] = cfdgpu.FACE_OUTLET  # an outlet inside the domain
    assert refused(d).code == EINVAL
    case = Case(cells=[8, 6, 6], size_m=[1.0, 1.0, 1.0], boundaries=[WALL, WALL, WALL],
                blocks=[Block([0.5, 0.0, 0.0], [0.625, 1.0, 1.0])], preconditioner=MULTIGRID)
    d = descriptor(case)
    assert d.anchored.size == 2  # the full-height wall splits the box
    d.component = np.where(d.component >= 0, 1 - d.component, -1)  # the two components renumbered
    e = refused(d)
    assert e.code == EINVAL and "components" in str(e)


def test_edited_wall_faces_are_not_read():
    """Walls carry no pressure weight (tests/multigrid.rs:249): an edited wall
    face leaves the descriptor's topology and weights as they were, so it is
    read up to the device set-up (without a device that set-up is what fails)."""
    d = descriptor(outlet_case())
    d.faces["area"][first(d, cfdgpu.FACE_WALL)] *= 2.0
    try:
        g = cfdgpu.Grid.from_descriptor(d)
    except cfdgpu.CfdError as e:
        assert e.code not in (EINVAL, EUNSUPPORTED), str(e)
    else:
        g.close()
