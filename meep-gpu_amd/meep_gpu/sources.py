"""Source time functions and deposit tables.

Time functions mirror MEEP's ``gaussian_src_time`` / ``continuous_src_time`` as the reference
states them (sources.py:577-660, :1765-1776, :177-197):

* Gaussian: ``width = 1/fwidth``, ``peak = start + cutoff*width``; the dipole is
  ``exp(-i*omega*(t - peak)) * exp(-(t-peak)*(t-peak)/(2*width*width)) * amp_factor`` with
  ``amp_factor = 1/(-2*pi*i*f)`` and a hard cut where ``float32(|t - peak|) > cutoff_time``;
  ``cutoff_time`` is ``cutoff*width`` shrunk by 0.9 until the envelope there is >= 1e-100,
  rounded to float32.
* The current MEEP deposits is the forward difference of the dipole,
  ``(dipole(t+dt) - dipole(t))/dt``, per part.

The engine never evaluates a time function on the device: ``current_table`` evaluates the
current in double on the host for every step of a run (electric sources at (n+1/2)*dt,
magnetic at n*dt, driver.py:3536/:3568) and the table is uploaded once, so a captured or
batched step sequence needs no host work.
"""

from __future__ import annotations

import cmath
import math
from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence, Tuple

import numpy


def amp_factor(frequency: float) -> complex:
    if not math.isfinite(frequency) or frequency == 0.0:
        raise ValueError(f"source frequency must be finite and nonzero, got {frequency!r}")
    return 1.0 / (-2j * math.pi * frequency)


def gaussian_cutoff_time(cutoff: float, width: float) -> float:
    t = cutoff * width
    while math.exp(-t * t / (2 * width * width)) < 1e-100:
        t *= 0.9
    return float(numpy.float32(t))


@dataclass
class GaussianSource:
    """MEEP's GaussianSource time dependence: ``frequency``, ``fwidth`` (MEEP units)."""
    frequency: float
    fwidth: float
    start_time: float = 0.0
    cutoff: float = 5.0
    width: float = field(init=False)
    peak_time: float = field(init=False)
    cutoff_time: float = field(init=False)

    def __post_init__(self):
        if self.fwidth <= 0:
            raise ValueError("fwidth must be positive")
        self.width = 1.0 / self.fwidth
        self.peak_time = self.start_time + self.cutoff * self.width
        self.cutoff_time = gaussian_cutoff_time(self.cutoff, self.width)
        self._omega = 2 * math.pi * self.frequency
        self._amp = amp_factor(self.frequency)

    def dipole(self, t: float) -> complex:
        rel = t - self.peak_time
        if numpy.float32(abs(rel)) > self.cutoff_time:
            return 0j
        envelope = math.exp(-rel * rel / (2 * self.width * self.width))
        return cmath.exp(-1j * self._omega * rel) * envelope * self._amp

    def current(self, t: float, dt: float) -> complex:
        late, early = self.dipole(t + dt), self.dipole(t)
        return complex((late.real - early.real) / dt, (late.imag - early.imag) / dt)

    def last_time(self) -> float:
        return self.peak_time + self.cutoff_time


@dataclass
class ContinuousSource:
    """MEEP's continuous_src_time without a ramp: ``exp(-i*omega*t) * amp_factor`` for t >= start."""
    frequency: float
    start_time: float = 0.0

    def dipole(self, t: float) -> complex:
        if t < self.start_time:
            return 0j
        return cmath.exp(-1j * 2 * math.pi * self.frequency * t) * amp_factor(self.frequency)

    def current(self, t: float, dt: float) -> complex:
        late, early = self.dipole(t + dt), self.dipole(t)
        return complex((late.real - early.real) / dt, (late.imag - early.imag) / dt)


def current_table(time_function, dt: float, steps: int, magnetic: bool) -> numpy.ndarray:
    """(steps, 2) float64 current words, one per step, at the source's slot time."""
    out = numpy.zeros((steps, 2), dtype=numpy.float64)
    offset = 0.0 if magnetic else 0.5 * dt
    for n in range(steps):
        value = complex(time_function.current(n * dt + offset, dt))
        out[n, 0] = value.real
        out[n, 1] = value.imag
    return out


@dataclass
class DepositBox:
    """Where a source deposits: a box of cells of one component with per-cell amplitudes."""
    component: str                       # 'Ex', 'Ey', 'Ez' (into D) or 'Hx'.. (into B)
    lo: Tuple[int, int, int]
    cnt: Tuple[int, int, int]
    amp_re: numpy.ndarray                # float64, shape cnt (C order)
    amp_im: numpy.ndarray

    @property
    def magnetic(self) -> bool:
        return self.component[0] in "HB"


def point_box(component: str, index: Sequence[int], amplitude: complex = 1.0) -> DepositBox:
    cnt = (1, 1, 1)
    return DepositBox(component, tuple(int(v) for v in index), cnt,
                      numpy.full(cnt, complex(amplitude).real),
                      numpy.full(cnt, complex(amplitude).imag))


def plane_box(component: str, axis: int, index: int, shape: Sequence[int],
              amplitude: complex = 1.0, margin: Sequence[int] = (0, 0, 0)) -> DepositBox:
    """A uniform sheet normal to ``axis`` at ``index``, covering the grid minus ``margin`` cells."""
    lo = [int(m) for m in margin]
    cnt = [int(n) - 2 * int(m) for n, m in zip(shape, margin)]
    lo[axis], cnt[axis] = int(index), 1
    return DepositBox(component, tuple(lo), tuple(cnt),
                      numpy.full(tuple(cnt), complex(amplitude).real),
                      numpy.full(tuple(cnt), complex(amplitude).imag))
