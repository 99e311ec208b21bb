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


#: Yee offsets in half cells per component (reference fields.py:218-223).
_IYEE = {"Ex": (1, 0, 0), "Ey": (0, 1, 0), "Ez": (0, 0, 1),
         "Hx": (0, 1, 1), "Hy": (1, 0, 1), "Hz": (1, 1, 0)}


def _zero_extent_weights(coord: float, n: int, dx: float, shift: int) -> Tuple[int, Tuple[float, ...]]:
    """The two samples a zero-size source direction lands between and their linear weights.

    MEEP's boundary weights for a zero-extent direction (``where_min == where_max``,
    loop_in_chunks.cpp:257-297, reference sources.py:657-696): the point at ``coord`` (origin
    at the cell centre) splits between the component's samples below and above it as
    ``(1 - f, f)``. A point exactly on a sample keeps both, the upper with weight 0.
    """
    pos = (coord + 0.5 * n * dx) / dx - 0.5 * shift  # in sample units of this component
    i0 = int(math.floor(pos + 1e-12))
    f = pos - i0
    if abs(f) < 1e-12:
        f = 0.0
    if not 0 <= i0 < n - 1:
        raise ValueError(f"a source at {coord} is not between two interior samples")
    return i0, (1.0 - f, f)


def meep_volume_source(grid, component: str, center: Sequence[float], size: Sequence[float],
                       amplitude: complex = 1.0) -> DepositBox:
    """A MEEP ``mp.Source`` whose size is zero or the whole cell along each axis.

    Zero-size directions interpolate linearly between the two nearest samples of the
    component and scale the amplitude by ``a`` (the resolution, MEEP sources.cpp:481-484,
    reference sources.py:860-866): a point source in 3-D deposits ``amp * a^3 * w`` -- the
    discrete delta of a current density. A direction the source spans whole takes every
    sample with weight 1 (a periodic axis has no boundary sample to halve).
    """
    iyee = _IYEE[component]
    a = grid.resolution
    lo, cnt, weights = [], [], []
    amp = complex(amplitude)
    for d in range(3):
        n = grid.shape[d]
        if grid.is_invariant(d):
            lo.append(0)
            cnt.append(1)
            weights.append(numpy.ones(1))
            continue
        if size[d] == 0.0:
            i0, w = _zero_extent_weights(float(center[d]), n, grid.dx, iyee[d])
            lo.append(i0)
            cnt.append(2)
            weights.append(numpy.array(w))
            amp = amp * a
        elif abs(size[d] - n * grid.dx) < 1e-9 * n * grid.dx:
            lo.append(0)
            cnt.append(n)
            weights.append(numpy.ones(n))
        else:
            raise ValueError("a source extent must be zero or the whole cell along each axis")
    w = weights[0][:, None, None] * weights[1][None, :, None] * weights[2][None, None, :]
    return DepositBox(component, tuple(lo), tuple(cnt), w * amp.real, w * amp.imag)
