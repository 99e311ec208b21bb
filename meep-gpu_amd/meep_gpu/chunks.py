"""The one-rank chunk division and the per-chunk facts a component steps with.

MEEP divides a one-process simulation into structure chunks along the PML's effort volumes
(structure.cpp:509-524): per axis a low layer piece, the interior and a high layer piece, each
layer piece spanning ``int(cells + 1.5)`` cells from its wall, and every product of pieces is a
chunk. Two facts are per chunk, not per cell:

* whether the chunk keeps a conductivity array for a component: ``set_conductivity`` samples
  the chunk's owned samples plus one and deletes the array when all are exactly zero
  (structure.cpp:883-900), and a chunk without one steps the lossless forms (MEEP
  step_generic.cpp), which under a PML round differently from the conductive forms with
  sigma 0;
* which chunk owns each stored cell (MEEP's ``owns``: ``lo < p <= hi`` in doubled
  coordinates, vec.cpp:445-462).

Restated from the reference's ``ChunkLayout.one_rank`` (chunk_map.py:384-428),
``_resolve``/``owned`` (:497-523), ``_resolved_rows`` (:613-647), ``owner_volume`` (:649-673),
``loop_rows`` (:675-718) and ``conductive_chunks``/``conductive_mask`` (:720-755), for the
periodic and metallic Cartesian grids this engine steps (no folded or cylindrical axis).
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import List, Sequence, Tuple

import numpy

from .grid import Grid

#: Yee shifts (1 = half-integer position) along x, y, z per component (MEEP's iyee_shift).
IYEE = {"Ex": (1, 0, 0), "Ey": (0, 1, 0), "Ez": (0, 0, 1),
        "Dx": (1, 0, 0), "Dy": (0, 1, 0), "Dz": (0, 0, 1),
        "Hx": (0, 1, 1), "Hy": (1, 0, 1), "Hz": (1, 1, 0),
        "Bx": (0, 1, 1), "By": (1, 0, 1), "Bz": (1, 1, 0)}


@dataclass(frozen=True)
class Frame:
    """One axis on MEEP's doubled lattice: stored index i at Yee shift s sits at doubled
    ``origin + 2 i + s``; the window is ``(origin, origin + 2 owned]``."""
    origin: int
    stored: int
    owned: int
    wraps: bool
    invariant: bool

    @property
    def window(self) -> Tuple[int, int]:
        return self.origin, self.origin + 2 * self.owned


def grid_frames(grid: Grid, wraps: Sequence[bool] = (True, True, True)) -> Tuple[Frame, ...]:
    """A centred axis starts at ``-(n - n % 2)`` (MEEP's ``icenter``, grid.py:1095-1130)."""
    frames = []
    for axis in range(3):
        n = grid.shape[axis]
        if grid.is_invariant(axis):
            frames.append(Frame(0, 1, 1, True, True))
        else:
            frames.append(Frame(-(n - n % 2), n, n, bool(wraps[axis]), False))
    return tuple(frames)


@dataclass
class ChunkDivision:
    frames: Tuple[Frame, ...]
    boxes: List[Tuple[Tuple[int, int, int], Tuple[int, int, int]]]
    #: per chunk and axis: does the chunk carry that axis's PML layer (MEEP's sigsize > 1)
    flags: List[Tuple[bool, bool, bool]] = None

    @classmethod
    def one_rank(cls, grid: Grid, faces, wraps: Sequence[bool] = (True, True, True)):
        """The {low, interior, high} product per axis; ``faces`` per axis (low, high) cells."""
        frames = grid_frames(grid, wraps)
        pieces_by_axis = []
        for axis, frame in enumerate(frames):
            if frame.invariant:
                pieces_by_axis.append([(frame.origin, frame.origin, False)])
                continue
            lo_wall, hi_wall = frame.window
            low, high = faces[axis]
            low_hi = lo_wall + 2 * min(frame.owned, int(low + 1.5)) if low > 0 else lo_wall
            high_lo = hi_wall - 2 * min(frame.owned, int(high + 1.5)) if high > 0 else hi_wall
            if low_hi <= high_lo:
                pieces = [(lo_wall, low_hi, True), (low_hi, high_lo, False),
                          (high_lo, hi_wall, True)]
            else:  # overlapping effort volumes: the overlap is a chunk of its own
                pieces = [(lo_wall, high_lo, True), (high_lo, low_hi, True),
                          (low_hi, hi_wall, True)]
            pieces_by_axis.append([p for p in pieces if p[0] < p[1]])
        boxes, flags = [], []
        for px in pieces_by_axis[0]:
            for py in pieces_by_axis[1]:
                for pz in pieces_by_axis[2]:
                    boxes.append(((px[0], py[0], pz[0]), (px[1], py[1], pz[1])))
                    flags.append((px[2], py[2], pz[2]))
        return cls(frames, boxes, flags)

    def chunk_owns(self, chunk: numpy.ndarray, coords) -> numpy.ndarray:
        """Does chunk ``chunk`` (an index per point, -1 none) own each point of doubled
        ``coords`` (one broadcastable array per axis)? ``lo < p <= hi`` on resolved axes."""
        index = numpy.asarray(chunk, dtype=numpy.int64)
        lows = numpy.asarray([lo for lo, _ in self.boxes], dtype=numpy.int64)
        highs = numpy.asarray([hi for _, hi in self.boxes], dtype=numpy.int64)
        safe = numpy.where(index >= 0, index, 0)
        owned = index >= 0
        for axis, frame in enumerate(self.frames):
            if frame.invariant:
                continue
            p = numpy.asarray(coords[axis], dtype=numpy.int64)
            owned = owned & (p > lows[safe, axis]) & (p <= highs[safe, axis])
        return owned

    def chunk_of(self, coords) -> numpy.ndarray:
        """The chunk owning each point of doubled ``coords``, or -1 where none does."""
        arrays = [numpy.asarray(coords[a], dtype=numpy.int64)
                  for a in range(3) if not self.frames[a].invariant]
        shape = numpy.broadcast_shapes(*[a.shape for a in arrays]) if arrays else ()
        out = numpy.full(shape, -1, dtype=numpy.int64)
        for k in range(len(self.boxes)):
            hit = self.chunk_owns(numpy.full(shape, k, dtype=numpy.int64), coords)
            out = numpy.where(hit, k, out)
        return out

    def _resolved(self, axis: int, shift: int) -> numpy.ndarray:
        """Each stored row's doubled coordinate, moved into the window on a wrapping axis; a
        row MEEP does not own (a wall) takes the nearest owned row's."""
        frame = self.frames[axis]
        if frame.invariant:
            return numpy.zeros(frame.stored, dtype=numpy.int64)
        points = frame.origin + 2 * numpy.arange(frame.stored, dtype=numpy.int64) + shift
        lo, hi = frame.window
        if frame.wraps:
            points = numpy.where(points <= lo, points + (hi - lo), points)
            points = numpy.where(points > hi, points - (hi - lo), points)
        owned = (points > lo) & (points <= hi)
        owned_at = numpy.flatnonzero(owned)
        for row in numpy.flatnonzero(~owned):
            points[row] = points[owned_at[numpy.argmin(numpy.abs(owned_at - row))]]
        return points

    def owner_volume(self, iyee: Sequence[int]) -> numpy.ndarray:
        """Per stored cell of one component lattice, the index of the chunk owning it."""
        coords = [self._resolved(axis, int(iyee[axis])) for axis in range(3)]
        grids = numpy.meshgrid(*coords, indexing="ij")
        owner = numpy.full(grids[0].shape, -1, dtype=numpy.int64)
        for k, (lo, hi) in enumerate(self.boxes):
            hit = owner < 0
            for axis, frame in enumerate(self.frames):
                if not frame.invariant:
                    hit &= (grids[axis] > lo[axis]) & (grids[axis] <= hi[axis])
            owner[hit] = k
        if numpy.any(owner < 0):
            raise ValueError("a cell's coordinate is owned by no chunk")
        return owner

    def loop_rows(self, axis: int, shift: int, chunk: int) -> numpy.ndarray:
        """The stored rows of the samples MEEP's LOOP_OVER_VOL visits in one chunk: its owned
        samples plus one (below on the integer lattice, above on the half one); a sample past
        a wrapping window's top reads its periodic image, past a wall the last stored row."""
        frame = self.frames[axis]
        if frame.invariant:
            return numpy.arange(frame.stored, dtype=numpy.int64)
        lo, hi = self.boxes[chunk][0][axis], self.boxes[chunk][1][axis]
        rows = (numpy.arange(lo + shift, hi + shift + 1, 2, dtype=numpy.int64)
                - frame.origin - shift) // 2
        if frame.wraps:
            return rows % frame.owned
        return numpy.clip(rows, 0, frame.stored - 1)

    def conductive_mask(self, sigma: numpy.ndarray, iyee: Sequence[int]) -> numpy.ndarray:
        """Per stored cell: does the chunk owning it keep this component's conductivity (some
        sample it loops over nonzero)?"""
        sigma = numpy.asarray(sigma)
        present = numpy.array([
            bool(numpy.any(sigma[numpy.ix_(*[self.loop_rows(a, int(iyee[a]), k)
                                              for a in range(3)])] != 0))
            for k in range(len(self.boxes))], dtype=bool)
        return present[self.owner_volume(iyee)]
