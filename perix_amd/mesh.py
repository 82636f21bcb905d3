"""Structured PD point cloud (host side, computed by libpxg's C++ code)."""
from dataclasses import dataclass, field

from . import pxg

WALLS = {"left": 0, "right": 1, "bottom": 2, "top": 3, "back": 4, "front": 5}


@dataclass
class StructuredGrid:
    """A 'perix' Mesh block (ReadMeshBlock.cpp:38-58): quad4 in 2D, hex8 in 3D."""
    dim: int
    nx: int
    ny: int
    nz: int = 1
    lo: tuple = (0.0, 0.0, 0.0)
    hi: tuple = (1.0, 1.0, 1.0)


@dataclass
class Cloud:
    grid: StructuredGrid
    coords: object
    vols: object
    mirror: object
    wall: object
    dxyz: object
    nbulk: int
    n: int = field(init=False)

    def __post_init__(self):
        self.n = len(self.vols)

    def group(self, name):
        """1-based ids of the '<name>nodes_ghost' group."""
        import numpy as np
        return (np.nonzero(self.wall == WALLS[name])[0] + 1).astype(np.int32)


def build_structured_cloud(g):
    coords, vols, mirror, wall, dxyz, nbulk = pxg.structured_cloud(g)
    return Cloud(g, coords, vols, mirror, wall, dxyz, nbulk)


def horizon_radius(g, factor):
    return pxg.horizon(g, factor)
