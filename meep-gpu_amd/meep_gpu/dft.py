"""Flux monitors: where a plane samples, what it transforms, and the spectrum read-out.

Mirrors reference ``meep_gpu.dft.FluxMonitor`` (dft.py:2814): each accumulating step the four
tangential components are averaged onto cell centres (dft.py:352, MEEP's yee2cent) and
transformed, E at ``t`` and H at ``t - dt/2`` with ``t = step_count*dt`` after the step
(dft.py:316-330, :3031-3063), scale ``dt*decimation/sqrt(2*pi)``. The transform runs on the
device in float64 (``fdtd_add_flux_plane``); the read-out (dft.py:3090-3130) is

    flux(f) = sum over the plane of Re(E1(f) * conj(H2(f)) - E2(f) * conj(H1(f))) * dA

with (E1, E2, H1, H2) the tangential components in cyclic order after the normal, so that
the normal Poynting component comes out with MEEP's sign for x, y and z normals alike.
Planes are given as whole-cell index boxes with unit weights (the fractional-cell ladder of a
MEEP-lifted monitor, dft.py:604/:2753, belongs to the lift).
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


@dataclass
class FluxMonitor:
    """One planar flux monitor registered on a simulation (a box one cell thick)."""
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

    def __post_init__(self):
        if self.cnt[self.normal] != 1:
            raise ValueError("a flux plane is one cell thick along its normal")
        self.frequencies = numpy.asarray(self.frequencies, dtype=numpy.float64)

    @property
    def omegas(self) -> numpy.ndarray:
        return 2.0 * math.pi * self.frequencies

    @property
    def scale(self) -> float:
        return self.dt * self.decimation / math.sqrt(2.0 * math.pi)

    @property
    def num_frequencies(self) -> int:
        return int(self.frequencies.size)

    def attach(self, sim) -> int:
        self._sim = sim
        self.plane_id = sim.add_flux_plane(self.normal, self.lo, self.cnt, self.omegas, self.scale,
                                           self.decimation)
        return self.plane_id

    def transforms(self) -> numpy.ndarray:
        """complex128 [4 (E1, E2, H1, H2)][nf][box cells, C order]."""
        if self._sim is None:
            raise RuntimeError("the monitor is not attached to a running simulation")
        return self._sim.flux_read(self.plane_id)

    def get_flux_spectrum(self) -> numpy.ndarray:
        return flux_spectrum(self.transforms(), self.dx, self.weight)


def flux_spectrum(transforms: numpy.ndarray, dx: float, weight: float = 1.0) -> numpy.ndarray:
    e1, e2, h1, h2 = transforms
    axes = tuple(range(1, e1.ndim))
    s = numpy.real(e1 * numpy.conj(h2) - e2 * numpy.conj(h1))
    return numpy.sum(s, axis=axes) * dx * dx * weight
