"""Which cells step a component's conductive forms under MEEP's one-rank chunk division
(ORACLE / test infrastructure only).

MEEP keeps a structure chunk's conductivity array only when some sample the chunk loops over
is nonzero (structure.cpp:883-900); a chunk without one steps the lossless forms. Restated
from the reference's ChunkLayout: ``one_rank`` boxes (chunk_map.py:384-428), the owner of
each stored cell (``_resolved_rows``/``owner_volume`` :613-673, MEEP's ``owns`` lo < p <= hi),
the rows a chunk loops over (``loop_rows`` :675-718) and ``conductive_mask`` (:720-755).
Periodic and metallic Cartesian axes only.

Pinned by the ``conductive_chunks_pml_2d`` fixture digest (a block conductivity in one quadrant
of a metallic 2-D grid with a PML on both axes).
"""

from __future__ import annotations

import numpy


def _axis(n: int, invariant: bool, metallic: bool):
    """(origin, stored, owned, wraps, invariant) of one axis; a centred axis starts at
    -(n - n % 2) doubled (MEEP's icenter)."""
    if invariant:
        return 0, 1, 1, True, True
    return -(n - n % 2), n, n, not metallic, False


def _pieces(axis, low: float, high: float):
    origin, stored, owned, wraps, invariant = axis
    if invariant:
        return [(origin, origin)]
    lo_wall, hi_wall = origin, origin + 2 * owned
    a = lo_wall + 2 * min(owned, int(low + 1.5)) if low > 0 else lo_wall
    b = hi_wall - 2 * min(owned, int(high + 1.5)) if high > 0 else hi_wall
    cuts = [(lo_wall, a), (a, b), (b, hi_wall)] if a <= b else [(lo_wall, b), (b, a), (a, hi_wall)]
    return [c for c in cuts if c[0] < c[1]]


def _coordinates(axis, shift: int) -> numpy.ndarray:
    origin, stored, owned, wraps, invariant = axis
    if invariant:
        return numpy.zeros(stored, dtype=numpy.int64)
    lo, hi = origin, origin + 2 * owned
    p = origin + 2 * numpy.arange(stored, dtype=numpy.int64) + shift
    if wraps:
        p = numpy.where(p <= lo, p + 2 * owned, p)
        p = numpy.where(p > hi, p - 2 * owned, p)
    mine = (p > lo) & (p <= hi)
    good = numpy.flatnonzero(mine)
    for row in numpy.flatnonzero(~mine):  # a wall row takes the nearest owned row's place
        p[row] = p[good[numpy.argmin(numpy.abs(good - row))]]
    return p


def _loop(axis, shift: int, lo: int, hi: int) -> numpy.ndarray:
    origin, stored, owned, wraps, invariant = axis
    if invariant:
        return numpy.arange(stored)
    rows = (numpy.arange(lo + shift, hi + shift + 1, 2) - origin - shift) // 2
    return rows % owned if wraps else numpy.clip(rows, 0, stored - 1)


def conductive_cells(shape, invariant, metallic, faces, sigma: numpy.ndarray, iyee):
    """Bool volume (C order (nx, ny, nz)): the cells of one component lattice whose chunk keeps
    the conductivity ``sigma``."""
    axes = [_axis(shape[a], bool(invariant[a]), bool(metallic[a])) for a in range(3)]
    pieces = [_pieces(axes[a], *faces[a]) for a in range(3)]
    boxes = [(px, py, pz) for px in pieces[0] for py in pieces[1] for pz in pieces[2]]
    keeps = []
    for box in boxes:
        rows = [_loop(axes[a], int(iyee[a]), box[a][0], box[a][1]) for a in range(3)]
        keeps.append(bool(numpy.any(sigma[numpy.ix_(*rows)] != 0)))
    coords = numpy.meshgrid(*[_coordinates(axes[a], int(iyee[a])) for a in range(3)],
                            indexing="ij")
    out = numpy.zeros(tuple(shape), dtype=bool)
    found = numpy.zeros(tuple(shape), dtype=bool)
    for keep, box in zip(keeps, boxes):
        inside = ~found
        for a in range(3):
            if not axes[a][4]:
                inside &= (coords[a] > box[a][0]) & (coords[a] <= box[a][1])
        out[inside] = keep
        found |= inside
    assert found.all(), "a cell owned by no chunk"
    return out
