"""Lorentz/Drude susceptibilities and the float32 recurrence words the ADE kernel steps with.

Mirrors reference ``meep_gpu.dispersion.Susceptibility`` (dispersion.py:184). The model, in
MEEP units (frequency f = omega/2pi, gamma likewise):

    d2P/dt2 + gamma dP/dt + omega0^2 P = sigma omega0^2 E      (Lorentz; Drude drops omega0^2 P)

discretised as MEEP's update_P (susceptibility.cpp:188-256):

    P' = gamma1inv * ((P*(2 - od_denom) - gamma1*P_prev) + od*(sigma*W))

The four words are formed exactly as MEEP forms them (``realnum_constants``,
dispersion.py:243-290): omega0, gamma and dt rounded to float32 first, 2*pi*omega0 in double
then rounded, gamma*2 in float then *pi in double then rounded, everything after that in
float32, left to right.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Tuple

import numpy

LORENTZIAN = "lorentzian"
DRUDE = "drude"

f32 = numpy.float32


@dataclass(frozen=True)
class Susceptibility:
    frequency: float
    gamma: float = 0.0
    kind: str = LORENTZIAN

    def __post_init__(self):
        if self.kind not in (LORENTZIAN, DRUDE):
            raise ValueError(f"kind must be {LORENTZIAN!r} or {DRUDE!r}, got {self.kind!r}")
        if not math.isfinite(self.frequency) or self.frequency <= 0:
            raise ValueError("susceptibility frequency must be finite and > 0")
        if not math.isfinite(self.gamma) or self.gamma < 0:
            raise ValueError("susceptibility gamma must be finite and >= 0 (no gain)")

    @property
    def is_drude(self) -> bool:
        return self.kind == DRUDE

    def realnum_constants(self, dt: float) -> Tuple[numpy.float32, ...]:
        """``(two_minus_od, gamma1, omega0dtsqr, gamma1inv)`` as MEEP's float32 words."""
        w0 = f32(self.frequency)
        g = f32(self.gamma)
        step = f32(dt)
        w2pi = f32((2.0 * math.pi) * float(w0))
        g2pi = f32(float(g * f32(2)) * math.pi)
        od = ((w2pi * w2pi) * step) * step
        half = (g2pi * step) / f32(2)
        g1inv = f32(1) / (f32(1) + half)
        g1 = f32(1) - half
        return (f32(2) - (f32(0) if self.is_drude else od), g1, od, g1inv)

    def chi1(self, frequency: float) -> complex:
        """Continuum susceptibility at sigma = 1 (MEEP susceptibility.cpp:297-307)."""
        f0, f = self.frequency, frequency
        if self.is_drude:
            return (f0 * f0) / complex(-f * f, -self.gamma * f)
        return (f0 * f0) / complex(f0 * f0 - f * f, -self.gamma * f)

    def stable(self, dt: float) -> bool:
        """The recurrence's characteristic roots lie in the unit disc (no growth)."""
        a, g1, od, g1inv = (float(v) for v in self.realnum_constants(dt))
        # z^2 - g1inv*a z + g1inv*g1 = 0 for the homogeneous update
        roots = numpy.roots([1.0, -g1inv * a, g1inv * g1])
        return bool(numpy.all(numpy.abs(roots) <= 1.0 + 1e-12))
