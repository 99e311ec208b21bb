"""The computational grid: cell counts, dx/dt and which axes are resolved.

Mirrors reference ``meep_gpu.grid.Grid`` (grid.py:293) for the configurations this engine
steps: a Cartesian cell in 3-D, in MEEP's 2-D (z translationally invariant, one stored cell,
reference grid.py:66-76) or in its 1-D (x and y invariant, the cell along z; grid.py:1217-1223).
Every axis wraps unless the driver marks it metallic.

Mirror symmetry (reference grid.py:404-575, MEEP's grid_volume::halve in vec.cpp): a folded
axis of N full cells stores MEEP's halved ``N - N//2 + 1`` cells with cell 0 one cell below the
mirror plane, and one more over periodic boundaries -- the slot past the window top that images
row ``N - stored + 2`` about the second mirror a periodic fold implies (stepping.py:2361-2390).
Over metallic boundaries the far wall is not stored.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional, Sequence, Tuple


def meep_cell_count(length: float, resolution: float) -> int:
    """MEEP's ``int(size*a + 0.5)`` (reference grid.py:288): truncation of x + 1/2."""
    return int(length * resolution + 0.5)


@dataclass
class Grid:
    resolution: float
    cell_size: Tuple[float, float, float]
    courant: float = 0.5
    dimensions: int = 3
    #: mirror planes: ``(axis, phase)`` with axis "X"/"Y"/"Z" (or 0..2) and phase +1 (even,
    #: MEEP's default) or -1 (odd); a bare axis name is an even mirror
    symmetry: Sequence = ()
    #: per axis: a metallic boundary (only a fold's storage depends on it here)
    metallic: Tuple[bool, bool, bool] = (False, False, False)
    shape: Tuple[int, int, int] = field(init=False)
    shape_full: Tuple[int, int, int] = field(init=False)
    #: per axis None or (phase, kind, far row): kind 1 over metallic boundaries, 2 over
    #: periodic ones, whose last slot images stored row `far row`
    folds: Tuple[Optional[Tuple[int, int, Optional[int]]], ...] = field(init=False)
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
        self.shape_full = tuple(counts)
        self.metallic = tuple(bool(m) for m in self.metallic)
        folds = [None, None, None]
        for entry in self.symmetry:
            axis, phase = (entry, 1) if isinstance(entry, (str, int)) else entry
            a = "XYZ".index(axis.upper()) if isinstance(axis, str) else int(axis)
            if a not in (0, 1, 2):
                raise ValueError(f"a mirror plane is normal to X, Y or Z, got {axis!r}")
            if phase not in (1, -1):
                raise ValueError(f"a mirror's phase is +1 (even) or -1 (odd), got {phase!r}")
            if folds[a] is not None:
                raise ValueError(f"two mirror planes normal to {'XYZ'[a]}")
            if self.is_invariant(a):
                raise ValueError(f"axis {'xyz'[a]} is invariant: it has no mirror plane")
            full = counts[a]
            periodic = not self.metallic[a]
            stored = full - full // 2 + 1 + (1 if periodic else 0)
            if stored < 4:
                raise ValueError(f"a folded axis needs 4 stored cells, {'xyz'[a]} has {stored}")
            counts[a] = stored
            folds[a] = (int(phase), 2 if periodic else 1, full - stored + 2 if periodic else None)
        self.folds = tuple(folds)
        self.shape = tuple(counts)

    def owned(self, axis: int) -> int:
        """MEEP's window on the axis in cells: the stored count, less the far ghost slot of a
        periodic fold (reference Grid.owned_cells)."""
        fold = self.folds[axis]
        return self.shape[axis] - (1 if fold is not None and fold[1] == 2 else 0)

    def fold_site(self, axis: int, doubled: int, parity: int):
        """Where MEEP serves a requested site of a folded axis (``loop_in_chunks``' symmetry
        images, the reference's ``folded_axis_sites`` dft.py:1004-1100): the site at doubled
        full-cell coordinate ``doubled`` of a component with Yee shift ``parity`` on the axis
        is first brought into the full cell's owned window by a lattice vector (periodic
        boundaries; past a metallic face nothing serves it), then, below the plane (under
        ``-parity``), reflected to ``-doubled``; on a periodic fold a site past the stored
        window is reflected again through the second mirror at doubled ``N``. Returns
        ``(stored row, reflections)`` -- row ``j`` sits at doubled ``2 j - 2 + parity`` -- or
        None. Each reflection multiplies the value by the component's parity."""
        fold = self.folds[axis]
        full = self.shape_full[axis]
        lattice = 2 * full
        little = -(full - full % 2) + 2 - parity
        big = little + lattice - 2
        if doubled < little or doubled > big:
            if fold[1] == 1:  # metallic: no lattice vector; the wall holds zero
                return None
            doubled += lattice if doubled < little else -lattice
            if doubled < little or doubled > big:
                return None
        flips = 0
        if doubled < -parity:
            doubled, flips = -doubled, 1
        top = self.owned(axis) - 1  # a periodic fold's far ghost row is not owned
        row = (doubled + 2 - parity) // 2
        if row > top:
            if fold[1] == 1:
                return None
            doubled, flips = lattice - doubled, flips + 1
            row = (doubled + 2 - parity) // 2
            if row < 0 or row > top or doubled < -parity:
                return None
        return row, flips

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
