"""Flux monitors: where a plane samples, what it transforms, and the spectrum read-out.

Mirrors reference ``meep_gpu.dft.FluxMonitor`` (dft.py:2814): each accumulating step the four
tangential components are averaged onto cell centres (dft.py:352, MEEP's yee2cent) and
transformed, E at ``t`` and H at ``t - dt/2`` with ``t = step_count*dt`` after the step
(dft.py:316-330, :3031-3063), scale ``dt*decimation/sqrt(2*pi)``. The transform runs on the
device in float64 (``fdtd_add_flux_plane``); the read-out (dft.py:3090-3130) is

    flux(f) = sum over the plane of Re(E1(f) * conj(H2(f)) - E2(f) * conj(H1(f))) * dA

with (E1, E2, H1, H2) the tangential components in cyclic order after the normal, so that
the normal Poynting component comes out with MEEP's sign for x, y and z normals alike.
A plane given as a whole-cell index box has unit weights; ``FluxMonitor.from_region`` places
a MEEP ``FluxRegion`` as MEEP does (its ladder weights, dft.py:604/:2753).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional, Sequence, Tuple

import numpy


def meep_frequency_span(fcen: float, df: float, nfreq: int) -> numpy.ndarray:
    if int(nfreq) < 1:
        raise ValueError("nfreq must be >= 1")
    if nfreq == 1:
        return numpy.array([float(fcen)])
    return numpy.linspace(fcen - 0.5 * df, fcen + 0.5 * df, int(nfreq))


def boundary_weight_ladder(where_min: float, where_max: float, a: float
                           ) -> Tuple[int, int, Tuple[float, float, float, float]]:
    """MEEP's integration ladder for one direction of a monitor volume on the centred grid.

    ``vec2diel_floor`` / ``vec2diel_ceil`` bracket the bounds on the odd doubled lattice (the
    cell centres sit at doubled coordinates 2k+1, i.e. at (k + 1/2)/a from the origin) and
    ``compute_boundary_weights`` picks the four end weights by the same case ladder
    (MEEP loop_in_chunks.cpp:233-297 as reference dft.py:604-690 states it, for parity 1).
    ``w0``/``w1`` are the overlaps of the first and last cell with the request, so an extended
    direction integrates to exactly ``(where_max - where_min) * a`` cells and a flat one is a
    linear interpolation (the weights sum to 1). Returns (is, ie, (s0, s1, e0, e1)).
    """
    is_d = 1 + 2 * math.floor(where_min * a - 0.5)
    ie_d = 1 + 2 * math.ceil(where_max * a - 0.5)
    w0 = 1.0 - where_min * a + 0.5 * is_d
    w1 = 1.0 + where_max * a - 0.5 * ie_d
    span = ie_d - is_d
    if span >= 6:
        s0, s1 = w0 * w0 / 2, 1 - (1 - w0) * (1 - w0) / 2
        e0, e1 = w1 * w1 / 2, 1 - (1 - w1) * (1 - w1) / 2
    elif span == 4:
        s0 = w0 * w0 / 2
        s1 = 1 - (1 - w0) * (1 - w0) / 2 - (1 - w1) * (1 - w1) / 2
        e0, e1 = w1 * w1 / 2, s1
    elif where_min == where_max:
        s0, s1, e0, e1 = w0, w1, w1, w0
    elif span == 2:
        s0 = w0 * w0 / 2 - (1 - w1) * (1 - w1) / 2
        e0 = w1 * w1 / 2 - (1 - w0) * (1 - w0) / 2
        s1, e1 = e0, s0
    else:
        s0 = s1 = e0 = e1 = 1.0
    return is_d, ie_d, (s0, s1, e0, e1)


def ladder_weight(index: int, count: int, s0: float, s1: float, e0: float, e1: float) -> float:
    """Weight of position ``index`` in a run of ``count`` (MEEP's IVEC_LOOP_WEIGHT1x)."""
    if 1 < index < count - 2:
        return 1.0
    if index == 0:
        return s0
    if index == 1:
        return s1
    if index == count - 1:
        return e0
    if index == count - 2:
        return e1
    return 1.0


def axis_cells(n: int, dx: float, where_min: float, where_max: float) -> Tuple[int, numpy.ndarray]:
    """(first cell index, per-cell weights) of one resolved direction of a monitor volume whose
    coordinates have their origin at the cell centre (the grid's cell centres at
    ``-n*dx/2 + (i + 1/2)*dx``; ``n`` even, so the centres are MEEP's odd doubled sites)."""
    if n % 2:
        raise ValueError("MEEP's monitor ladder needs an even cell count along a resolved axis")
    a = 1.0 / dx
    is_d, ie_d, (s0, s1, e0, e1) = boundary_weight_ladder(where_min, where_max, a)
    count = (ie_d - is_d) // 2 + 1
    first = (is_d - 1) // 2 + n // 2
    if first < 0 or first + count > n:
        raise ValueError(f"monitor extent [{where_min}, {where_max}] leaves the grid")
    w = numpy.array([ladder_weight(q, count, s0, s1, e0, e1) for q in range(count)])
    return first, w


@dataclass
class FluxMonitor:
    """A planar flux monitor: one or two sample planes of cell-centred fields with weights.

    The plain form (``normal, lo, cnt``) is one plane of whole cells at weight 1. A monitor
    built by ``from_region`` is placed where MEEP places it: the centre planes bracketing the
    requested coordinate along the normal, interpolated linearly, and the ladder weights
    across the plane (``boundary_weight_ladder``); MEEP weights the E transforms only
    (reference dft.py:2753-2800, :3107-3130), which is the same sum.
    """
    normal: int
    lo: Tuple[int, int, int]
    cnt: Tuple[int, int, int]
    frequencies: numpy.ndarray
    dx: float
    dt: float
    decimation: int = 1
    weight: float = 1.0
    plane_id: Optional[int] = None
    _sim: object = field(default=None, repr=False)
    # (normal index, weight) per sample plane; the transverse weights over the plane's cells
    normal_planes: Optional[list] = None
    cell_weights: Optional[numpy.ndarray] = None
    area_element: Optional[float] = None   # dx per resolved transverse axis (dx**2 in 3-D)
    plane_ids: Optional[list] = None

    def __post_init__(self):
        if self.cnt[self.normal] != 1:
            raise ValueError("a flux plane is one cell thick along its normal")
        self.frequencies = numpy.asarray(self.frequencies, dtype=numpy.float64)
        if self.normal_planes is None:
            self.normal_planes = [(int(self.lo[self.normal]), 1.0)]
        if self.area_element is None:
            self.area_element = self.dx * self.dx

    @classmethod
    def from_region(cls, grid, center, size, frequencies, decimation: int = 1) -> "FluxMonitor":
        """MEEP's ``sim.add_flux(..., mp.FluxRegion(center, size))`` on this grid: exactly one
        axis of ``size`` is zero (the normal); coordinates have their origin at the cell centre."""
        flat = [a for a in range(3) if not grid.is_invariant(a) and size[a] == 0]
        if len(flat) != 1:
            raise ValueError("a flux region is flat along exactly one resolved axis")
        normal = flat[0]
        lo, cnt, weights = [0, 0, 0], [1, 1, 1], []
        planes = None
        for a in range(3):
            n = grid.shape[a]
            if grid.is_invariant(a):
                weights.append(numpy.ones(1))
                continue
            lo_w = center[a] - 0.5 * size[a]
            hi_w = center[a] + 0.5 * size[a]
            first, w = axis_cells(n, grid.dx, lo_w, hi_w)
            if a == normal:  # a plane of weight zero contributes nothing: not registered
                planes = [(first + q, float(w[q])) for q in range(w.size) if w[q] != 0.0]
                lo[a], cnt[a] = first, 1
                weights.append(numpy.ones(1))
            else:
                lo[a], cnt[a] = first, w.size
                weights.append(w)
        cw = weights[0][:, None, None] * weights[1][None, :, None] * weights[2][None, None, :]
        resolved = sum(1 for a in range(3) if a != normal and not grid.is_invariant(a))
        return cls(normal, tuple(lo), tuple(cnt), frequencies, grid.dx, grid.dt, decimation,
                   normal_planes=planes, cell_weights=cw, area_element=grid.dx ** resolved)

    @property
    def omegas(self) -> numpy.ndarray:
        return 2.0 * math.pi * self.frequencies

    @property
    def scale(self) -> float:
        return self.dt * self.decimation / math.sqrt(2.0 * math.pi)

    @property
    def num_frequencies(self) -> int:
        return int(self.frequencies.size)

    def plane_boxes(self):
        """(lo, cnt, weight) of every sample plane."""
        for index, w in self.normal_planes:
            lo = list(self.lo)
            lo[self.normal] = index
            yield tuple(lo), tuple(self.cnt), w

    def attach(self, sim) -> int:
        self._sim = sim
        self.plane_ids = [sim.add_flux_plane(self.normal, lo, cnt, self.omegas, self.scale,
                                             self.decimation)
                          for lo, cnt, _ in self.plane_boxes()]
        self.plane_id = self.plane_ids[0]
        return self.plane_id

    def transforms(self) -> numpy.ndarray:
        """complex128 [4 (E1, E2, H1, H2)][nf][box cells, C order] of the first sample plane."""
        if self._sim is None:
            raise RuntimeError("the monitor is not attached to a running simulation")
        return self._sim.flux_read(self.plane_id)

    def get_flux_spectrum(self) -> numpy.ndarray:
        if self._sim is None:
            raise RuntimeError("the monitor is not attached to a running simulation")
        total = None
        for pid, (_, _, w) in zip(self.plane_ids, self.plane_boxes()):
            part = flux_spectrum(self._sim.flux_read(pid), self.area_element, self.weight * w,
                                 self.cell_weights)
            total = part if total is None else total + part
        return total


def flux_spectrum(transforms: numpy.ndarray, area_element: float, weight: float = 1.0,
                  cell_weights: Optional[numpy.ndarray] = None) -> numpy.ndarray:
    """sum over the plane of w * Re(E1 conj(H2) - E2 conj(H1)) * dA, per frequency."""
    e1, e2, h1, h2 = transforms
    axes = tuple(range(1, e1.ndim))
    s = numpy.real(e1 * numpy.conj(h2) - e2 * numpy.conj(h1))
    if cell_weights is not None:
        s = s * cell_weights.reshape((1,) + s.shape[1:])
    return numpy.sum(s, axis=axes) * area_element * weight
