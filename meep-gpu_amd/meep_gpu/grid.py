"""The computational grid: cell counts, dx/dt and which axes are resolved.

Mirrors reference ``meep_gpu.grid.Grid`` (grid.py:293) for the configurations this engine
steps: a Cartesian cell in 3-D, in MEEP's 2-D (z translationally invariant, one stored cell,
reference grid.py:66-76) or in its 1-D (x and y invariant, the cell along z; grid.py:1217-1223).
Every axis wraps unless the driver marks it metallic.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Tuple


def meep_cell_count(length: float, resolution: float) -> int:
    """MEEP's ``int(size*a + 0.5)`` (reference grid.py:288): truncation of x + 1/2."""
    return int(length * resolution + 0.5)


@dataclass
class Grid:
    resolution: float
    cell_size: Tuple[float, float, float]
    courant: float = 0.5
    dimensions: int = 3
    shape: Tuple[int, int, int] = field(init=False)
    dx: float = field(init=False)
    dt: float = field(init=False)

    def __post_init__(self):
        if self.resolution <= 0 or self.courant <= 0:
            raise ValueError("resolution and courant must be positive")
        if self.dimensions not in (1, 2, 3):
            raise ValueError(f"dimensions must be 1, 2 or 3, got {self.dimensions!r}")
        self.dx = 1.0 / self.resolution
        self.dt = self.courant / self.resolution
        counts = []
        for axis in range(3):
            if self.is_invariant(axis):
                if self.cell_size[axis] not in (0, 0.0):
                    raise ValueError(f"a {self.dimensions}-D cell has zero extent along its "
                                     f"invariant axes ({'xyz'[axis]} here)")
                counts.append(1)
            else:
                counts.append(meep_cell_count(self.cell_size[axis], self.resolution))
        if min(counts) < 1:
            raise ValueError(f"cell {self.cell_size!r} at resolution {self.resolution} is empty")
        self.shape = tuple(counts)

    def is_invariant(self, axis: int) -> bool:
        """MEEP's assignment: z at 2-D, x and y at 1-D."""
        return (self.dimensions == 2 and axis == 2) or (self.dimensions == 1 and axis < 2)

    @property
    def invariant(self) -> Tuple[bool, bool, bool]:
        return tuple(self.is_invariant(a) for a in range(3))

    @property
    def total_cells(self) -> int:
        nx, ny, nz = self.shape
        return nx * ny * nz

    @property
    def dtdx(self) -> float:
        return self.dt / self.dx
