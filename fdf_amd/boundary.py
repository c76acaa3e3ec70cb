"""Boundary-condition configuration of the Eulerian field preparation.

Mirrors the reference's per-face configuration (BoundaryFaceConfig,
common.h:253-262): a mathematical type, a handler and the handler's
parameters for each of the six faces -Xi +Xi -Eta +Eta -Zeta +Zeta.  Files in
the reference's bcs.dat format are parsed by the library
(picgpu_parse_boundary_file, io.c:2515-2680).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

from . import _lib as L

TYPES = {"WALL": L.BCTYPE_WALL, "SYMMETRY": L.BCTYPE_SYMMETRY, "INLET": L.BCTYPE_INLET,
         "OUTLET": L.BCTYPE_OUTLET, "NOGRAD": L.BCTYPE_NOGRAD}
HANDLERS = {"noslip": L.BC_WALL_NOSLIP, "constant_velocity": L.BC_INLET_CONSTANT_VELOCITY,
            "conservation": L.BC_OUTLET_CONSERVATION, "allcopy": L.BC_NOGRAD_COPY_GHOST,
            "parabolic": L.BC_INLET_PARABOLIC}


@dataclass
class BoundaryConfig:
    """type[f], handler[f], vel[f] = (vx, vy, vz), vmax[f] for faces f = 0..5."""

    type: list = field(default_factory=lambda: [L.BCTYPE_WALL] * 6)
    handler: list = field(default_factory=lambda: [L.BC_WALL_NOSLIP] * 6)
    vel: list = field(default_factory=lambda: [[0.0, 0.0, 0.0] for _ in range(6)])
    vmax: list = field(default_factory=lambda: [1.0] * 6)

    def set_face(self, face, type_name: str, handler_name: str, vx=0.0, vy=0.0, vz=0.0, vmax=1.0):
        f = L.FACE_NAMES.index(face) if isinstance(face, str) else int(face)
        self.type[f] = TYPES[type_name.upper()]
        self.handler[f] = HANDLERS[handler_name.lower()]
        self.vel[f] = [float(vx), float(vy), float(vz)]
        self.vmax[f] = float(vmax)
        return self

    def to_c(self) -> L.PicgpuBoundaryConfig:
        b = L.PicgpuBoundaryConfig()
        for f in range(6):
            b.type[f] = int(self.type[f])
            b.handler[f] = int(self.handler[f])
            b.vmax[f] = float(self.vmax[f])
            for c in range(3):
                b.velocity[f][c] = float(self.vel[f][c])
        return b

    @classmethod
    def from_c(cls, b: L.PicgpuBoundaryConfig) -> "BoundaryConfig":
        return cls(type=list(b.type), handler=list(b.handler), vel=[list(b.velocity[f]) for f in range(6)],
                   vmax=list(b.vmax))

    @classmethod
    def from_file(cls, path: str) -> "BoundaryConfig":
        lib = L.load()
        b = L.PicgpuBoundaryConfig()
        L.check(lib.picgpu_parse_boundary_file(str(path).encode(), C.byref(b)))
        return cls.from_c(b)

    @classmethod
    def default(cls) -> "BoundaryConfig":
        lib = L.load()
        b = L.PicgpuBoundaryConfig()
        lib.picgpu_default_boundary_config(C.byref(b))
        return cls.from_c(b)


def channel(inlet_face: int = 4, handler: str = "constant_velocity", speed: float = 1.0,
            outlet: str = "allcopy") -> BoundaryConfig:
    """Walls on four faces, an inlet on `inlet_face` along its axis and a
    zero-gradient outlet on the opposite face."""
    cfg = BoundaryConfig()
    axis = inlet_face // 2
    v = [0.0, 0.0, 0.0]
    v[axis] = speed if inlet_face % 2 == 0 else -speed
    cfg.set_face(inlet_face, "INLET", handler, *v, vmax=speed)
    if outlet == "allcopy":
        cfg.set_face(inlet_face ^ 1, "NOGRAD", "allcopy")
    return cfg
