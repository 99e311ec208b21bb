"""MEEP's one-rank chunk division as per-axis cuts, and the per-chunk facts it decides.

MEEP splits a one-process run into structure chunks at the PML's effort volumes
(structure.cpp:509-524): along each axis a piece at each face that carries a layer, spanning
``int(thickness + 1.5)`` cells from the wall, and the rest; the chunks are the products of
the pieces. Everything this engine needs from the division is per axis, so it is held as
per-axis cut rows rather than as boxes:

* **Ownership** (MEEP's ``owns``, ``lo < p <= hi`` in doubled coordinates, vec.cpp:445-462).
  For a piece ``[a, b)`` in cells, a sample on the half-integer lattice (Yee shift 1) at row
  ``i`` sits at doubled ``2 i + 1`` from the axis origin, so the piece owns rows ``a .. b-1``;
  one on the integer lattice sits at ``2 i`` and the piece owns rows ``a+1 .. b``: row ``b``
  of the last piece is row 0's periodic image, and on a metallic axis row 0 is the wall,
  which no piece owns (it is counted with its neighbour, row 1).
* **The conductivity keep test** (structure.cpp:883-900): a chunk keeps a component's array
  when a sample its loop visits is nonzero, and the loop runs over the owned samples plus
  the one on the other side of the chunk's low face: rows ``a .. b`` inclusive on either
  lattice (the periodic image past the top, the last row past a wall). A chunk without the
  array steps the lossless forms (MEEP step_generic.cpp), which under a PML round
  differently from the conductive forms with sigma 0.
* **The first row of a piece**, where a centred read of an integer-lattice component takes a
  word the reading chunk does not own (``readback``).

The grid is centred (an axis of ``n`` cells starts at doubled ``-(n - n % 2)``, MEEP's
``center_origin``); periodic and metallic Cartesian axes only (no folded or cylindrical axis).
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
class AxisCuts:
    """One axis: ``n`` stored rows cut into pieces ``[cuts[k], cuts[k+1])`` of cells, each
    carrying the axis's layer or not."""
    n: int
    cuts: Tuple[int, ...]
    layer: Tuple[bool, ...]
    wraps: bool = True
    invariant: bool = False

    @property
    def pieces(self) -> int:
        return len(self.layer)

    @property
    def origin(self) -> int:
        """The doubled coordinate of the axis's low face (the centred grid's)."""
        return 0 if self.invariant else -(self.n - self.n % 2)

    @classmethod
    def flat(cls) -> "AxisCuts":
        """An axis the dimension does not carry: one row, one piece."""
        return cls(1, (0, 1), (False,), True, True)

    @classmethod
    def effort(cls, n: int, low: float, high: float, wraps: bool) -> "AxisCuts":
        """MEEP's effort-volume cut of one axis with layers ``low``/``high`` cells thick."""
        lo = min(n, int(low + 1.5)) if low > 0 else 0
        hi = n - min(n, int(high + 1.5)) if high > 0 else n
        if lo <= hi:
            cuts, layer = [0, lo, hi, n], [True, False, True]
        else:  # the two layer pieces overlap: the overlap is a piece of its own, all layer
            cuts, layer = [0, hi, lo, n], [True, True, True]
        keep = [k for k in range(3) if cuts[k] < cuts[k + 1]]
        return cls(n, tuple([cuts[keep[0]]] + [cuts[k + 1] for k in keep]),
                   tuple(layer[k] for k in keep), bool(wraps))

    def owner(self, shift: int) -> numpy.ndarray:
        """Per stored row, the piece owning the sample with Yee shift ``shift`` there."""
        if self.invariant:
            return numpy.zeros(self.n, dtype=numpy.int64)
        rows = numpy.arange(self.n)
        edges = numpy.asarray(self.cuts[1:-1], dtype=numpy.int64)
        if shift:
            return numpy.searchsorted(edges, rows, side="right").astype(numpy.int64)
        # rows a+1 .. b belong to piece [a, b): row r to the piece whose interval holds r - 1
        out = numpy.searchsorted(edges, rows - 1, side="right").astype(numpy.int64)
        out[0] = self.pieces - 1 if self.wraps else 0  # the top's image, or the wall
        return out

    def visited(self, piece: int) -> numpy.ndarray:
        """The stored rows one piece's conductivity loop visits: ``a .. b`` inclusive."""
        if self.invariant:
            return numpy.zeros(1, dtype=numpy.int64)
        rows = numpy.arange(self.cuts[piece], self.cuts[piece + 1] + 1, dtype=numpy.int64)
        return rows % self.n if self.wraps else numpy.minimum(rows, self.n - 1)

    def first_rows(self) -> numpy.ndarray:
        """Per stored row: is it the first row of its piece (on the half-integer lattice)?"""
        out = numpy.zeros(self.n, dtype=bool)
        if not self.invariant:
            out[numpy.asarray(self.cuts[:-1], dtype=numpy.int64)] = True
        return out

    def layer_rows(self) -> numpy.ndarray:
        """Per stored row (half-integer lattice): does its piece carry the layer?"""
        return numpy.asarray(self.layer, dtype=bool)[self.owner(1)]


@dataclass(frozen=True)
class ChunkDivision:
    """The product of three axes' pieces; chunk ``(px, py, pz)`` is number
    ``(px * Py + py) * Pz + pz``."""
    axes: Tuple[AxisCuts, AxisCuts, AxisCuts]

    @classmethod
    def one_rank(cls, grid: Grid, faces, wraps: Sequence[bool] = (True, True, True)):
        """``faces`` per axis: (low, high) layer thickness in cells. On a mirror-folded axis the
        cut is the same rule over MEEP's owned window of the halved axis (its low face, the
        mirror plane, carries no layer), and a periodic fold's far ghost row, past the window,
        joins the last piece."""
        axes = []
        for a in range(3):
            if grid.is_invariant(a):
                axes.append(AxisCuts.flat())
                continue
            folds = getattr(grid, "folds", (None, None, None))
            if folds[a] is None:
                axes.append(AxisCuts.effort(grid.shape[a], *faces[a], wraps[a]))
                continue
            owned = grid.owned(a)
            cut = AxisCuts.effort(owned, *faces[a], False)
            axes.append(AxisCuts(grid.shape[a], cut.cuts[:-1] + (grid.shape[a],), cut.layer,
                                 False))
        return cls(tuple(axes))

    @property
    def count(self) -> int:
        return self.axes[0].pieces * self.axes[1].pieces * self.axes[2].pieces

    def _number(self, px, py, pz):
        return (px * self.axes[1].pieces + py) * self.axes[2].pieces + pz

    def owner_volume(self, iyee: Sequence[int]) -> numpy.ndarray:
        """Per stored cell of one component lattice, the number of the chunk owning it."""
        px, py, pz = (self.axes[a].owner(int(iyee[a])) for a in range(3))
        return self._number(px[:, None, None], py[None, :, None], pz[None, None, :])

    def keeps(self, sigma: numpy.ndarray, iyee: Sequence[int]) -> numpy.ndarray:
        """Per chunk: does it keep this component's conductivity (a visited sample nonzero)?"""
        sigma = numpy.asarray(sigma)
        out = numpy.zeros(self.count, dtype=bool)
        for px in range(self.axes[0].pieces):
            rx = self.axes[0].visited(px)
            for py in range(self.axes[1].pieces):
                ry = self.axes[1].visited(py)
                for pz in range(self.axes[2].pieces):
                    rz = self.axes[2].visited(pz)
                    out[self._number(px, py, pz)] = bool(
                        numpy.any(sigma[numpy.ix_(rx, ry, rz)] != 0))
        return out

    def conductive_mask(self, sigma: numpy.ndarray, iyee: Sequence[int]) -> numpy.ndarray:
        """Per stored cell: does the chunk owning it keep this component's conductivity?"""
        return self.keeps(sigma, iyee)[self.owner_volume(iyee)]

    def reader_masks(self, component: str, axes: Sequence[int], shape, first_step: bool):
        """Per word of a B component's centred sum, words ordered (f, n1[, n2, n12]): (take
        the H word, take a zero).

        The chunk that reads a cell centre is the one owning the centre, row by row the piece
        holding the centre's row on each axis. Along an integer-lattice axis the word above
        the centre is always that chunk's own; the word below is another chunk's exactly at
        the first row of the reader's piece (its low face). A word the reader does not own is
        what MEEP's boundary exchange left in the reader's array there: where the reader holds
        H as B's alias (no layer on the component's own axis) that is the owner's H word, and
        where the reader keeps an H of its own it is +0 when the run has taken exactly one step
        (``first_step``). MEEP array_slice.cpp over the chunk arrays, update_eh.cpp:162-172;
        the reference states the same rule as HAliasRule.masks (readback.py:491-546)."""
        own = "xyz".index(component[1])

        def along(axis, rows):
            view = [1, 1, 1]
            view[axis] = shape[axis]
            return numpy.asarray(rows).reshape(view)

        aliasing = ~along(own, self.axes[own].layer_rows())
        below = [along(axis, self.axes[axis].first_rows()) for axis in axes]
        take_h, take_zero = [], []
        for corner in range(1 << len(axes)):  # bit q set: the word above along axes[q]
            notowned = numpy.zeros((1, 1, 1), dtype=bool)
            for q in range(len(axes)):
                if not corner >> q & 1:
                    notowned = notowned | below[q]
            take_h.append(numpy.broadcast_to(aliasing & notowned, shape))
            stale = notowned & ~aliasing if first_step else numpy.zeros((1, 1, 1), dtype=bool)
            take_zero.append(numpy.broadcast_to(stale, shape))
        return take_h, take_zero


@dataclass(frozen=True)
class BoxDivision:
    """A lifted division held as MEEP's chunk boxes ``(lo, hi]`` in doubled coordinates with
    their per-axis layer flags, over per-axis frames ``(origin, stored, owned, wraps,
    invariant)``: any division, the halved grid of a mirror fold included, for the one
    question the read-out asks of it (``reader_masks``)."""
    frames: Tuple[Tuple[int, int, int, bool, bool], ...]
    boxes: Tuple[Tuple[Tuple[int, int, int], Tuple[int, int, int]], ...]
    flags: Tuple[Tuple[bool, bool, bool], ...]

    def _contains(self, box: int, coords) -> numpy.ndarray:
        lo, hi = self.boxes[box]
        inside = numpy.ones((1, 1, 1), dtype=bool)
        for a in range(3):
            if self.frames[a][4]:  # an invariant axis bounds nothing
                continue
            inside = inside & (coords[a] > lo[a]) & (coords[a] <= hi[a])
        return inside

    def reader_masks(self, component: str, axes: Sequence[int], shape, first_step: bool):
        """The rule of ``ChunkDivision.reader_masks``, answered by box containment: the
        reader is the box holding the centre (doubled ``origin + 2i + 1``), a word at
        ``centre -/+ 1`` along the centred axes is the reader's when its box holds it."""
        own = "xyz".index(component[1])
        centre = []
        for a in range(3):
            view = [1, 1, 1]
            view[a] = shape[a]
            centre.append((self.frames[a][0] + 2 * numpy.arange(shape[a]) + 1).reshape(view))
        reader = numpy.full(tuple(shape), -1, dtype=numpy.int64)
        for k in range(len(self.boxes)):
            reader = numpy.where((reader < 0) & self._contains(k, centre), k, reader)
        aliasing = numpy.zeros(tuple(shape), dtype=bool)
        for k in range(len(self.boxes)):
            if not self.flags[k][own]:
                aliasing |= reader == k
        take_h, take_zero = [], []
        for corner in range(1 << len(axes)):
            coords = list(centre)
            for q, axis in enumerate(axes):
                coords[axis] = centre[axis] + (1 if corner >> q & 1 else -1)
            owned = numpy.zeros(tuple(shape), dtype=bool)
            for k in range(len(self.boxes)):
                owned |= (reader == k) & self._contains(k, coords)
            notowned = ~owned & (reader >= 0)
            take_h.append(aliasing & notowned)
            take_zero.append(notowned & ~aliasing if first_step
                             else numpy.zeros(tuple(shape), dtype=bool))
        return take_h, take_zero


def from_boxes(frames: Sequence[Tuple[int, int, int, bool, bool]], boxes, layer_flags
               ) -> ChunkDivision:
    """A lifted division: MEEP's chunk boxes ``(lo, hi]`` in doubled coordinates, with per
    box and axis the layer flag, over axes ``(origin, stored, owned, wraps, invariant)``.
    Only a product of per-axis pieces over unfolded axes (``owned == stored``) is taken."""
    axes: List[AxisCuts] = []
    for a, (origin, stored, owned, wraps, invariant) in enumerate(frames):
        if invariant:
            axes.append(AxisCuts.flat())
            continue
        if owned != stored:
            raise ValueError("a folded axis (owned rows differ from stored rows) is not built")
        if origin != -(stored - stored % 2):
            raise ValueError("an axis whose origin is not the centred grid's")
        spans = sorted({(int(lo[a]), int(hi[a])) for lo, hi in boxes})
        cuts = [(spans[0][0] - origin) // 2]
        for (lo, hi), nxt in zip(spans, spans[1:] + [None]):
            if nxt is not None and nxt[0] != hi:
                raise ValueError("the lifted chunk boxes do not tile the axis")
            cuts.append((hi - origin) // 2)
        if cuts[0] != 0 or cuts[-1] != stored:
            raise ValueError("the lifted chunk boxes do not span the axis")
        layer = []
        for lo, hi in spans:
            flags = {bool(f[a]) for (blo, bhi), f in zip(boxes, layer_flags)
                     if (int(blo[a]), int(bhi[a])) == (lo, hi)}
            if len(flags) != 1:
                raise ValueError("the lifted chunks disagree on a piece's layer flag")
            layer.append(flags.pop())
        axes.append(AxisCuts(int(stored), tuple(cuts), tuple(layer), bool(wraps)))
    division = ChunkDivision(tuple(axes))
    if division.count != len(boxes):
        raise ValueError("the lifted chunk boxes are not a product of per-axis pieces")
    return division
