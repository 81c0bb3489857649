"""Case description of the pressure path: the parts of a CFD-craft case that fix
the pressure operator, and the C structure that carries them over the C-ABI.

Only the keys that reach the pressure solve are read (crates/cfd-craft/src/case.rs):
``domain.size_m``, ``grid.cells``, ``boundaries`` (case.rs:175-230), ``flow``
(its ``vector`` classifies open faces, case.rs:345-367), ``geometry`` boxes and
polygon prisms (case.rs:415-459), and ``numerics.pressure_preconditioner``,
``pressure_relative_tolerance`` and ``pressure_max_iterations``
(case.rs:958-975). Everything else in a case (transport, run, output, ...)
belongs to the rest of the solver and is ignored here.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

# Boundary codes; the order is part of the C-ABI (include/cfdgpu.h).
PERIODIC, OPEN, WALL, SLIP, GROUND_OPEN_TOP = 0, 1, 2, 3, 4
BOUNDARY_NAMES = {
    "periodic": PERIODIC,
    "open": OPEN,
    "wall": WALL,
    "slip": SLIP,
    "ground_open_top": GROUND_OPEN_TOP,
}
FLOW_VECTOR, FLOW_NEUTRAL_RANS = 0, 1
IC0, MULTIGRID = 0, 1
PRECONDITIONERS = {"ic0": IC0, "multigrid": MULTIGRID}

# FaceKind order of grid.rs:60-81.
INTERIOR, WALL_FACE, SLIP_FACE, INLET, OUTLET, BLOCKED = 0, 1, 2, 3, 4, 5


class CBlock(ctypes.Structure):
    _fields_ = [
        ("min_m", ctypes.c_double * 3),
        ("max_m", ctypes.c_double * 3),
        ("n_vertices", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("footprint_xy", ctypes.POINTER(ctypes.c_double)),
    ]


class CCase(ctypes.Structure):
    _fields_ = [
        ("cells", ctypes.c_int64 * 3),
        ("size_m", ctypes.c_double * 3),
        ("boundary", ctypes.c_int32 * 3),
        ("flow_kind", ctypes.c_int32),
        ("velocity_m_s", ctypes.c_double * 3),
        ("speed_m_s", ctypes.c_double),
        ("reference_height_m", ctypes.c_double),
        ("roughness_m", ctypes.c_double),
        ("direction_deg_from", ctypes.c_double),
        ("preconditioner", ctypes.c_int32),
        ("n_blocks", ctypes.c_int32),
        ("blocks", ctypes.POINTER(CBlock)),
    ]


@dataclass
class Block:
    """An axis-aligned box, or a polygon prism when ``footprint`` is given
    (case.rs:415-459): cells whose centers lie inside are solid."""

    min_m: Sequence[float]
    max_m: Sequence[float]
    footprint: Optional[Sequence[Sequence[float]]] = None
    id: str = ""


@dataclass
class Case:
    cells: Sequence[int]
    size_m: Sequence[float]
    boundaries: Sequence[int]
    flow_kind: int = FLOW_VECTOR
    velocity_m_s: Sequence[float] = (0.0, 0.0, 0.0)
    speed_m_s: float = 0.0
    reference_height_m: float = 10.0
    roughness_m: float = 0.1
    direction_deg_from: float = 270.0
    blocks: List[Block] = field(default_factory=list)
    preconditioner: int = IC0
    tolerance: float = 1e-10
    max_iterations: int = 1200

    @property
    def count(self) -> int:
        return int(self.cells[0]) * int(self.cells[1]) * int(self.cells[2])

    @property
    def dx(self):
        return [self.size_m[a] / self.cells[a] for a in range(3)]

    def face_dims(self):
        dims = []
        for a in range(3):
            d = [int(v) for v in self.cells]
            if self.boundaries[a] != PERIODIC:
                d[a] += 1
            dims.append(d)
        return dims

    def face_count(self) -> int:
        return sum(math.prod(d) for d in self.face_dims())

    def flow_vector(self, z: float):
        """Flow::vector (case.rs:345-367)."""
        if self.flow_kind == FLOW_NEUTRAL_RANS:
            z0 = self.roughness_m
            u = self.speed_m_s * math.log((max(z, 0.0) + z0) / z0) / math.log((self.reference_height_m + z0) / z0)
            t = self.direction_deg_from * (math.pi / 180.0)
            return [-u * math.sin(t), -u * math.cos(t), 0.0]
        return list(self.velocity_m_s)

    @classmethod
    def from_json(cls, c: dict) -> "Case":
        """Reads the pressure-relevant keys of a `cfd-craft.case/0.1` document."""
        flow = c["flow"]
        kind = flow["kind"]
        kw = {}
        if kind == "neutral_rans":
            kw = dict(
                flow_kind=FLOW_NEUTRAL_RANS,
                speed_m_s=float(flow["speed_m_s"]),
                reference_height_m=float(flow["reference_height_m"]),
                roughness_m=float(flow["roughness_m"]),
                direction_deg_from=float(flow["direction_deg_from"]),
            )
        elif kind == "laminar":
            kw = dict(velocity_m_s=[float(v) for v in flow["initial_velocity_m_s"]])
        elif kind == "prescribed_uniform":
            kw = dict(velocity_m_s=[float(v) for v in flow["velocity_m_s"]])
        else:
            raise ValueError(f"unknown flow kind {kind!r}")
        blocks = [
            Block(min_m=list(b["min_m"]), max_m=list(b["max_m"]), footprint=b.get("footprint_m"), id=b.get("id", ""))
            for b in c.get("geometry", [])
        ]
        num = c.get("numerics", {})
        b = c["boundaries"]
        return cls(
            cells=[int(v) for v in c["grid"]["cells"]],
            size_m=[float(v) for v in c["domain"]["size_m"]],
            boundaries=[BOUNDARY_NAMES[b["x"]], BOUNDARY_NAMES[b["y"]], BOUNDARY_NAMES[b["z"]]],
            blocks=blocks,
            preconditioner=PRECONDITIONERS[num.get("pressure_preconditioner", "ic0")],
            tolerance=float(num.get("pressure_relative_tolerance", 1e-10)),
            max_iterations=int(num.get("pressure_max_iterations", 1200)),
            **kw,
        )

    def to_c(self):
        """Returns (CCase, keepalive): the structure and the buffers it points into."""
        keep = []
        blocks = (CBlock * max(1, len(self.blocks)))()
        for i, b in enumerate(self.blocks):
            blocks[i].min_m[:] = [float(v) for v in b.min_m]
            blocks[i].max_m[:] = [float(v) for v in b.max_m]
            if b.footprint:
                flat = (ctypes.c_double * (2 * len(b.footprint)))(*[float(v) for p in b.footprint for v in p])
                keep.append(flat)
                blocks[i].n_vertices = len(b.footprint)
                blocks[i].footprint_xy = ctypes.cast(flat, ctypes.POINTER(ctypes.c_double))
            else:
                blocks[i].n_vertices = 0
        keep.append(blocks)
        c = CCase()
        c.cells[:] = [int(v) for v in self.cells]
        c.size_m[:] = [float(v) for v in self.size_m]
        c.boundary[:] = [int(v) for v in self.boundaries]
        c.flow_kind = int(self.flow_kind)
        c.velocity_m_s[:] = [float(v) for v in self.velocity_m_s]
        c.speed_m_s = float(self.speed_m_s)
        c.reference_height_m = float(self.reference_height_m)
        c.roughness_m = float(self.roughness_m)
        c.direction_deg_from = float(self.direction_deg_from)
        c.preconditioner = int(self.preconditioner)
        c.n_blocks = len(self.blocks)
        c.blocks = ctypes.cast(blocks, ctypes.POINTER(CBlock))
        return c, keep
