"""Uniaxial PML coefficient table and the one-rank chunk membership the kernels select forms by.

Mirrors reference ``meep_gpu.pml.PML`` (pml.py:324) and ``chunk_map.one_rank_line``
(chunk_map.py:180) for wrapping axes. Per axis and per Yee sub-lattice (integer positions,
offset 0; half-integer, offset 1/2) the table holds float32 vectors

    kms = 1 - sig,   sinv = 1/(1 + sig),   kps = 1 + sig     (kappa = 1)

with ``sig = (dt/2) * prefac * u**order``. The grading follows MEEP's ``pml_x``
(structure.cpp:625-628 via pml.py:469-598): distance from each wall counted in half cells,
with the extent snapped to ``int(2*cells + 0.5)`` half cells and ``u`` normalised by the raw
request. On a wrapping axis the integer sample stored at index 0 is MEEP's owned lattice
point at 2N, graded from the upper wall; on a metallic axis, which does not wrap (reference
grid.py:1267-1289), it is the low wall itself at 0 (pml.py:578-582).
``prefac = -ln(R) / (4 * cells*dx / (order+1))``, per face (pml.py:436).

The membership line of an axis says whether the one-rank chunk owning a stored row carries
that axis's layer (MEEP picks the PML form per chunk, step_db.cpp:56-59): a face of ``t``
cells owns ``int(t + 1.5)`` rows from its wall, rows ``1..span`` on the integer lattice at the
low wall and ``0..span-1`` on the half lattice; the top ``span`` on the half lattice at the
high wall and the top ``span-1`` plus row 0 (the wrap of the wall plane) on the integer one;
a metallic axis has no such wrap, and its integer row 0 (the wall) is owned by no chunk.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, Optional, Sequence, Tuple

import numpy

from .grid import Grid

Faces = Tuple[Tuple[float, float], Tuple[float, float], Tuple[float, float]]
_AXES = ("x", "y", "z")


def normalize_faces(spec, grid: Grid) -> Faces:
    """Scalar (every resolved face), ``{axis: t | (low, high)}`` or per-axis sequence."""
    def pair(value):
        if isinstance(value, (tuple, list)):
            if len(value) != 2:
                raise ValueError(f"a face pair needs (low, high), got {value!r}")
            return (value[0], value[1])
        return (value, value)

    faces = [(0, 0), (0, 0), (0, 0)]
    if isinstance(spec, dict):
        for key, value in spec.items():
            name = str(key).lower()
            if name not in _AXES:
                raise ValueError(f"unknown PML axis {key!r}")
            faces[_AXES.index(name)] = pair(value)
    elif isinstance(spec, (tuple, list)):
        if len(spec) != 3:
            raise ValueError("a positional PML spec has one entry per axis")
        faces = [pair(v) for v in spec]
    else:
        faces = [pair(spec) if not grid.is_invariant(a) else (0, 0) for a in range(3)]
        # the scalar means every face that is a boundary: a mirror plane is not (pml.py:387-415)
        for a in range(3):
            if grid.folds[a] is not None:
                faces[a] = (0, faces[a][1])
    for a in range(3):
        lo, hi = faces[a]
        if grid.folds[a] is not None and lo:
            raise ValueError(f"the {_AXES[a]} axis is mirror-symmetric: its low face is the "
                             f"mirror plane, which carries no absorber")
        if lo < 0 or hi < 0 or not (math.isfinite(lo) and math.isfinite(hi)):
            raise ValueError(f"PML thickness on {_AXES[a]} must be finite and >= 0")
        if grid.is_invariant(a) and (lo or hi):
            raise ValueError("an invariant axis carries no absorber")
        if lo + hi > grid.shape[a]:
            raise ValueError(f"PML ({lo}, {hi}) does not fit on the {_AXES[a]} axis "
                             f"({grid.shape[a]} cells)")
    return tuple(tuple(f) for f in faces)


def half_cell_extent(cells: float) -> int:
    return int(2.0 * float(cells) + 0.5)


@dataclass
class PML:
    grid: Grid
    thickness: object
    order: int = 2
    R_asymptotic: float = 1e-15
    #: Replaces -ln(R) when set (a fixture keeps libm out of its table this way).
    ln_reflection: Optional[float] = None
    #: Per axis: periodic (True) or terminated by a metallic wall (False).
    wraps: Tuple[bool, bool, bool] = (True, True, True)
    faces: Faces = field(init=False)
    vectors: Dict[Tuple[int, int], Tuple[numpy.ndarray, numpy.ndarray, numpy.ndarray]] = \
        field(init=False, repr=False)
    members: Dict[Tuple[int, int], numpy.ndarray] = field(init=False, repr=False)

    def __post_init__(self):
        if self.order < 1:
            raise ValueError("grading order must be >= 1")
        if not 0.0 < self.R_asymptotic < 1.0:
            raise ValueError("R_asymptotic must lie in (0, 1)")
        self.faces = normalize_faces(self.thickness, self.grid)
        self.wraps = tuple(bool(w) for w in self.wraps)
        if len(self.wraps) != 3:
            raise ValueError("wraps needs one flag per axis")
        # a folded axis does not wrap; its upper wall is MEEP's window top, 2 * owned half
        # cells, one cell below the stored end on a periodic fold (pml.py:600-611)
        self.wraps = tuple(w and self.grid.folds[a] is None for a, w in enumerate(self.wraps))
        self.vectors = {}
        self.members = {}
        for half in (0, 1):
            for axis in range(3):
                n = self.grid.shape[axis]
                sig = self._sigma(n, self.faces[axis], half, self.wraps[axis],
                                  2 * self.grid.owned(axis))
                one = numpy.ones(n, dtype=numpy.float32)
                total = one + sig
                self.vectors[(half, axis)] = ((one - sig).astype(numpy.float32),
                                              (numpy.float32(1.0) / total).astype(numpy.float32),
                                              total.astype(numpy.float32))
                self.members[(half, axis)] = self._line(n, self.faces[axis], half,
                                                        self.wraps[axis])

    @property
    def is_active(self) -> bool:
        return any(lo > 0 or hi > 0 for lo, hi in self.faces)

    def prefac(self, cells: float) -> float:
        depth = cells * self.grid.dx
        if depth <= 0:
            return 0.0
        ln_r = self.ln_reflection if self.ln_reflection is not None \
            else float(-numpy.log(self.R_asymptotic))
        return float(ln_r / (4.0 * depth * (1.0 / (self.order + 1))))

    def _sigma(self, n: int, faces: Tuple[float, float], half: int,
               wraps: bool = True, top: Optional[int] = None) -> numpy.ndarray:
        out = numpy.zeros(n, dtype=numpy.float32)
        # half-cell index of every stored row; on a wrapping axis index 0 of the integer
        # lattice is the image of 2n
        pos = 2 * numpy.arange(n) + half
        if wraps:
            pos = (pos - 1) % (2 * n) + 1
        for wall, cells in ((0, faces[0]), (2 * n if top is None else top, faces[1])):
            if cells <= 0:
                continue
            reach = half_cell_extent(cells)
            gap = numpy.abs(pos - wall)
            frac = 1.0 + ((reach - 2.0 * cells) - gap) / (2.0 * cells)
            hit = gap < reach
            out[hit] = ((0.5 * self.grid.dt) * self.prefac(cells)) * frac[hit] ** self.order
        return out

    @staticmethod
    def _line(n: int, faces: Tuple[float, float], half: int, wraps: bool = True) -> numpy.ndarray:
        line = numpy.zeros(n, dtype=bool)
        low, high = faces
        if low > 0:
            span = min(n, int(low + 1.5))
            if half:
                line[:span] = True
            else:
                line[1:min(n, span + 1)] = True
        if high > 0:
            span = min(n, int(high + 1.5))
            if half:
                line[n - span:] = True
            else:
                line[max(0, n - span + 1):] = True
                if wraps:
                    line[0] = True
        return line

    def member_fraction(self, axis: int, half: int) -> float:
        return float(self.members[(half, axis)].mean())
