"""MEEP's ``sim.get_array(component=c)``, word for word: over the whole cell, or a box.

What a user of MEEP (and of the reference's ``GpuRunResult.get_array``) reads back is not the
stored Yee words but MEEP's centred lattice: at every cell centre the float32 sum of the
component's two or four nearest words, ``0.25 * (((a + b) + c) + d)`` in double, rounded to
float32 once (array_slice.cpp). Restated from the reference's readback.py for real fields on
Cartesian grids without mirror folding or a Bloch phase:

* ``chunk_words`` (:416-453): the stored words, with a metallic wall's own plane read as +0
  (MEEP allocates it and never steps it);
* ``_up`` (:455-488) and ``_four_words`` (:580-596): the next word up along each axis where the
  component's Yee shift is 0 -- the owned wrap on a periodic axis, +0 past a metallic wall (and
  the corner past one);
* the B reads under a PML (the rule the reference states as ``HAliasRule.masks`` :491-546
  and ``h_alias_rule`` :549-570, here from the division's per-axis pieces): in a chunk that
  holds H as B's alias (no layer on the component's own axis) a word the chunk does not own is
  the owner's H word, and a chunk with H of its own reads +0 there when the run has taken
  exactly one step (``first_step``);
* ``centred_sums`` (:604-635), ``CentredLattice`` (:751-819: the periodic duplicate plane below
  the first on a non-metallic axis), ``total_volume`` (:829-853) and ``get_array`` (:856-972:
  the lattice points inside the volume, MEEP's interpolation weights and float ``+=`` collapse
  on a one-cell axis, the ``n > 1`` axis filter).

At k = 0 every phase MEEP multiplies by is exactly ``(1, +0)`` (``pow(eikna, -1)`` of
``(1, +0)``), so the centred sum's double value is the word before rounding.
"""

from __future__ import annotations

import itertools
import math
from typing import Callable, List, Sequence

import numpy

from .chunks import IYEE, ChunkDivision

_F32 = numpy.float32
_F64 = numpy.float64
B_COMPONENTS = ("Bx", "By", "Bz")


def _slab(axis: int, index):
    out = [slice(None)] * 3
    out[axis] = index
    return tuple(out)


def _up(parts: numpy.ndarray, axis: int, metallic: bool) -> numpy.ndarray:
    """The next word up along ``axis``: +0 past a metallic wall, the wrap elsewhere."""
    q = numpy.empty_like(parts)
    q[_slab(axis, slice(None, -1))] = parts[_slab(axis, slice(1, None))]
    q[_slab(axis, -1)] = 0.0 if metallic else parts[_slab(axis, 0)]
    return q


def _four_words(parts: numpy.ndarray, axes: Sequence[int], metallic) -> List[numpy.ndarray]:
    if not axes:
        return [parts]
    if len(axes) == 1:
        return [parts, _up(parts, axes[0], metallic[axes[0]])]
    n1 = _up(parts, axes[0], metallic[axes[0]])
    n2 = _up(parts, axes[1], metallic[axes[1]])
    n12 = _up(n1, axes[1], metallic[axes[1]])
    if metallic[axes[0]]:  # the corner past a metal face is a zero_metal word
        n12[_slab(axes[0], -1)] = 0.0
    return [parts, n1, n2, n12]


def _chunk_words(stored: numpy.ndarray, component: str, invariant, metallic) -> numpy.ndarray:
    parts = numpy.array(stored, dtype=_F32, copy=True)
    for axis in range(3):
        if IYEE[component][axis] == 0 and not invariant[axis] and metallic[axis]:
            parts[_slab(axis, 0)] = 0.0
    return parts


def _alias_masks(division: ChunkDivision, component: str, axes: Sequence[int], shape,
                 first_step: bool):
    """Per word of the centred sum: (take the H word, take a zero).

    The chunk that reads a cell centre is the one owning the centre, row by row the piece
    holding the centre's row on each axis. Along an integer-lattice axis the word above the
    centre is always that chunk's own; the word below is another chunk's exactly at the first
    row of the reader's piece (its low face). A word the reader does not own is what MEEP's
    boundary exchange left in the reader's array there: where the reader holds H as B's alias
    (no layer on the component's own axis) that is the owner's H word, and where the reader
    keeps an H of its own it is +0 when the run has taken exactly one step (``first_step``).
    MEEP array_slice.cpp over the chunk arrays; the reference states the same rule as
    HAliasRule.masks (readback.py:491-546).
    """
    own = "xyz".index(component[1])

    def along(axis, rows):
        view = [1, 1, 1]
        view[axis] = shape[axis]
        return numpy.asarray(rows).reshape(view)

    aliasing = ~along(own, division.axes[own].layer_rows())
    below = [along(axis, division.axes[axis].first_rows()) for axis in axes]
    take_h, take_zero = [], []
    for corner in range(1 << len(axes)):  # bit q set: the word above the centre along axes[q]
        notowned = numpy.zeros((1, 1, 1), dtype=bool)
        for q in range(len(axes)):
            if not corner >> q & 1:
                notowned = notowned | below[q]
        take_h.append(numpy.broadcast_to(aliasing & notowned, shape))
        stale = notowned & ~aliasing if first_step else numpy.zeros((1, 1, 1), dtype=bool)
        take_zero.append(numpy.broadcast_to(stale, shape))
    return take_h, take_zero


def centred_sums(read: Callable[[str], numpy.ndarray], component: str, invariant, metallic,
                 division: ChunkDivision, pml_active: bool, first_step: bool) -> numpy.ndarray:
    """``0.25 * sum`` in float64 at every stored cell centre (doubled ``origin + 2 i + 1``)."""
    stored = read(component)
    parts = _chunk_words(stored, component, invariant, metallic)
    axes = [a for a in range(3) if not invariant[a] and IYEE[component][a] == 0]
    words = _four_words(parts, axes, metallic)
    if axes and component in B_COMPONENTS and pml_active:
        h_parts = _chunk_words(read("H" + component[1]), component, invariant, metallic)
        h_words = _four_words(h_parts, axes, metallic)
        take_h, take_zero = _alias_masks(division, component, axes, stored.shape, first_step)
        words = [numpy.where(zero, _F32(0.0), numpy.where(take, wh, wb)).astype(_F32)
                 for wb, wh, take, zero in zip(words, h_words, take_h, take_zero)]
    if len(words) == 1:
        f = words[0]
        sums = ((f + f) + f) + f
    elif len(words) == 2:
        f, a = words
        sums = ((f + a) + f) + a
    else:
        f, b, c, d = words
        sums = ((f + b) + c) + d
    return sums.astype(_F64) * 0.25


def _vec2diel_floor(coordinate: float, a: float) -> int:  # loop_in_chunks.cpp:229-236
    return 1 + 2 * int(math.floor(coordinate * a - .5))


def _vec2diel_ceil(coordinate: float, a: float) -> int:  # loop_in_chunks.cpp:238-245
    return 1 + 2 * int(math.ceil(coordinate * a - .5))


def meep_array(read: Callable[[str], numpy.ndarray], component: str, grid,
               wraps: Sequence[bool], division: ChunkDivision, pml_active: bool,
               step_count: int, center=None, size=None) -> numpy.ndarray:
    """MEEP's ``get_array(component)`` (float32, MEEP's rank and axis order): over the whole
    cell (``fields::total_volume()``), or over the box ``center -/+ size/2`` as ``mp.Volume``
    forms it. ``read(name)`` returns a stored volume, C order (nx, ny, nz)."""
    if (center is None) != (size is None):
        raise ValueError("pass both center and size, or neither (the whole cell)")
    invariant = tuple(grid.is_invariant(a) for a in range(3))
    metallic = tuple(not invariant[a] and not wraps[a] for a in range(3))
    q = centred_sums(read, component, invariant, metallic, division, pml_active,
                     first_step=step_count == 1)
    order = [a for a in range(3) if not invariant[a]]
    ivec = {}
    for axis in order:
        ivec[axis] = division.axes[axis].origin + 2 * numpy.arange(q.shape[axis]) + 1
        if not metallic[axis]:  # the periodic duplicate plane below the first
            q = numpy.concatenate((q[_slab(axis, slice(-1, None))], q), axis=axis)
            ivec[axis] = numpy.concatenate(([ivec[axis][0] - 2], ivec[axis]))
    a = float(grid.resolution)
    pick = [numpy.array([0])] * 3
    weights = {}
    collapse = []
    for axis in order:
        cut = division.axes[axis]
        if center is None:
            low = cut.origin * 0.5 * grid.dx  # total_volume: [little_corner, big_corner - 2]
            high = (cut.origin + 2 * cut.n - 2) * 0.5 * grid.dx
        else:
            low = float(center[axis]) - float(size[axis]) * 0.5
            high = float(center[axis]) + float(size[axis]) * 0.5
        lo, hi = min(low, high), max(low, high)
        i_s, i_e = _vec2diel_floor(lo, a), _vec2diel_ceil(hi, a)
        keep = numpy.nonzero((ivec[axis] >= i_s) & (ivec[axis] <= i_e))[0]
        if keep.size == 0:
            raise ValueError(f"the cell misses the {'xyz'[axis]} centred lattice of {component}")
        pick[axis] = keep
        if lo != hi:
            continue
        w0 = 1. - lo * a + 0.5 * i_s  # loop_in_chunks.cpp:260-261, left to right
        w1 = 1. + hi * a - 0.5 * i_e
        if keep.size == 2:
            weights[axis] = numpy.array([min(w0, w0), min(w1, w1)], dtype=_F64)
            collapse.append(axis)
        else:
            visited = int(ivec[axis][keep[0]])
            w = min(w0, w1) if i_e == i_s else (w0 if visited == i_s else w1)
            weights[axis] = numpy.array([w], dtype=_F64)
        if keep.size > 2 and lo == hi:
            raise RuntimeError("an empty dimension spanning more than two centred planes")
    values = q[numpy.ix_(*pick)]
    weight = None
    for axis in order:
        factor = _F64(1.0)
        if axis in weights:
            view = [1, 1, 1]
            view[axis] = weights[axis].size
            factor = weights[axis].reshape(view)
        weight = (_F64(1.0) * factor) if weight is None else (factor * weight)
    if weight is None:
        weight = _F64(1.0)
    store = (weight * values).astype(_F32)  # times the phase (1, +0): the same double
    if collapse:  # collapse_array: reduced[...] += full[...] in float32, row-major
        shape = list(store.shape)
        reduced = tuple(1 if axis in collapse else shape[axis] for axis in range(3))
        acc = numpy.zeros(reduced, dtype=_F32)
        ranges = [range(shape[axis]) if axis in collapse else (None,) for axis in range(3)]
        for combo in itertools.product(*ranges):
            sel = tuple(slice(combo[axis], combo[axis] + 1) if axis in collapse else slice(None)
                        for axis in range(3))
            acc = acc + store[sel]
        store = acc
    drop = tuple(axis for axis in range(3) if axis not in order or store.shape[axis] == 1)
    if drop:
        store = store.squeeze(axis=drop)
    return numpy.array(store, dtype=_F32, copy=True)
