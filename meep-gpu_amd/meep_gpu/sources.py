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
from typing import Callable, List, Optional, Sequence, Tuple

import numpy

from .dft import boundary_weight_ladder, ladder_weight


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


def current_table(time_function, dt: float, steps: int, magnetic: bool,
                  integrated: bool = False) -> numpy.ndarray:
    """(steps, 2) float64 words, one per step, at the source's slot time: the current, for a
    deposit (electric at (n+1/2)*dt, magnetic at n*dt), or for an integrated source the dipole
    the constitutive update subtracts (electric at (n+1)*dt, step.cpp:105-109; magnetic at
    (n+1/2)*dt)."""
    out = numpy.zeros((steps, 2), dtype=numpy.float64)
    if integrated:
        offset = 0.5 * dt if magnetic else dt
        value_at = lambda t: time_function.dipole(t)
    else:
        offset = 0.0 if magnetic else 0.5 * dt
        value_at = lambda t: time_function.current(t, dt)
    for n in range(steps):
        value = complex(value_at(n * dt + offset))
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
    #: MEEP's ``is_integrated`` (a polarization source): nothing is deposited into D; the
    #: constitutive update reads D less ``amp * dipole`` at these cells (update_eh.cpp:66-140)
    integrated: bool = False

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


def _cmul(p: Tuple[float, float], q: Tuple[float, float]) -> Tuple[float, float]:
    """C++ ``complex<double>`` multiply with every product rounded: (ac - bd, ad + bc)."""
    return (p[0] * q[0] - p[1] * q[1], p[0] * q[1] + p[1] * q[0])


def _end_weights(start: int, end: int, is_d: int, ie_d: int, ends):
    """The end weights of a run [start, end] of doubled coordinates clipped out of [is, ie]
    (reference sources.py:1097-1121): a run that lost its first (last) point takes the next
    end weight, and a run of one, two or three points keeps the smaller of the weights that
    meet on a point."""
    s0, s1, e0, e1 = ends
    ws0 = ws1 = we0 = we1 = 1.0
    if start == is_d:
        ws0, ws1 = s0, s1
    elif start == is_d + 2:
        ws0 = s1
    if end == ie_d:
        we0, we1 = e0, e1
    elif end == ie_d - 2:
        we0 = e1
    if end == start:
        ws0 = we0 = ws1 = we1 = min(ws0, we0)
    elif end == start + 2:
        ws0 = we1 = min(ws0, we1)
        ws1 = we0 = min(ws1, we0)
    elif end == start + 4:
        ws1 = we1 = min(ws1, we1)
    return ws0, ws1, we0, we1


def _axis_passes(n: int, a: float, iyee: int, where_min: float, where_max: float, wraps: bool):
    """One resolved direction of a source request: for each lattice shift along it, the
    stored rows it deposits into and their weights, as contiguous (first row, weights) runs.

    MEEP loop_in_chunks.cpp:349-418 on one chunk: the bounds and boundary weights on the
    component's lattice, the owned window ``little_corner + 2 - iyee .. big_corner - iyee``
    with the axis starting at doubled ``io = -(n - n % 2)`` (the centred origin of an even
    or odd count), and on a periodic axis every shift of one lattice vector that meets the
    window. A row that the window maps past the last stored row is the periodic image of row
    0, or on a metallic high wall a plane zeroed before anything reads it (no deposit).
    Returns {shift: [(first_row, weights), ...]}, shifts in increasing order.
    """
    is_d, ie_d, ends = boundary_weight_ladder(where_min, where_max, a, 1 - iyee)
    io = -(n - n % 2)
    low, high = io + 2 - iyee, io + 2 * n - iyee
    shifts = [0]
    if wraps:
        half = 0.5 / a
        lattice = 2 * n * half
        shifts = range(int(math.floor((where_min - (io + 2 * n) * half) / lattice)),
                       int(math.ceil((where_max - io * half) / lattice)) + 1)
    out = {}
    for k in shifts:
        sh = 2 * n * k
        start, end = max(is_d - sh, low), min(ie_d - sh, high)
        if start > end:
            continue
        ws = _end_weights(start + sh, end + sh, is_d, ie_d, ends)
        count = (end - start) // 2 + 1
        runs, rows, weights = [], [], []
        for q in range(count):
            row = (start + 2 * q - io) // 2
            if row >= n:
                if not wraps:
                    continue
                row -= n
            if rows and row != rows[-1] + 1:
                runs.append((rows[0], numpy.array(weights)))
                rows, weights = [], []
            rows.append(row)
            weights.append(ladder_weight(q, count, *ws))
        if rows:
            runs.append((rows[0], numpy.array(weights)))
        if runs:
            out[k] = runs
    return out


def _fold_parity(grid, component: str, axis: int) -> int:
    """The component's sign under the fold of ``axis`` (MEEP's phase_shift): a vector odd
    about the plane its own axis is normal to, a pseudovector even there and odd about the
    others, times the plane's declared phase."""
    own = "xyz".index(component[1]) == axis
    declared = grid.folds[axis][0]
    return -declared if own != (component[0] in "HB") else declared


def _folded_axis_passes(grid, component: str, d: int, a: float, iyee: int, where_min: float,
                        where_max: float):
    """A mirror-folded direction of a source request: MEEP's two symmetry passes over it
    (``loop_in_chunks`` with the images, sources.cpp:456-489). The request's samples are laid
    on the full cell (its origin at the mirror plane); pass 0 takes those at and above the
    component's first owned sample (doubled ``-iyee``) directly, pass 1 those below it through
    the reflection, each clipped run with MEEP's end weights (``_end_weights``), and each
    sample lands where ``Grid.fold_site`` serves it. Returns {pass: [(first_row, weights)]},
    each weight times the component's parity once per reflection."""
    parity = _fold_parity(grid, component, d)
    is_d, ie_d, ends = boundary_weight_ladder(where_min, where_max, a, 1 - iyee)
    out = {}
    for bit, (start, end) in ((0, (max(is_d, -iyee), ie_d)), (1, (is_d, min(ie_d, -iyee - 2)))):
        if start > end:
            continue
        ws = _end_weights(start, end, is_d, ie_d, ends)
        count = (end - start) // 2 + 1
        served = {}
        for q in range(count):
            site = grid.fold_site(d, start + 2 * q, iyee)
            if site is not None:
                served[site[0]] = ladder_weight(q, count, *ws) * (parity ** site[1])
        runs, rows, weights = [], [], []
        for row in sorted(served):
            if rows and row != rows[-1] + 1:
                runs.append((rows[0], numpy.array(weights)))
                rows, weights = [], []
            rows.append(row)
            weights.append(served[row])
        if rows:
            runs.append((rows[0], numpy.array(weights)))
        if runs:
            out[bit] = runs
    return out


def meep_volume_source(grid, component: str, center: Sequence[float], size: Sequence[float],
                       amplitude: complex = 1.0, metallic: Sequence[bool] = (False, False, False),
                       integrated: bool = False) -> List[DepositBox]:
    """A MEEP ``mp.Source`` with a constant amplitude, placed as MEEP places it on one chunk.

    MEEP's ``add_volume_source`` (sources.cpp:456-489, reference sources.py:787-1140): the
    amplitude is scaled by ``a`` (the resolution) once per zero-size resolved direction -- a
    point source in 3-D deposits ``amp * a^3 * w``, the discrete delta of a current density --
    and each resolved direction integrates the request over the component's samples with
    MEEP's boundary weights: a zero-size direction interpolates linearly between the two
    samples around the point (one sample with weight 1 when the point is on it), an extended
    one takes the ladder's end weights at its two ends. A direction the dimension does not
    carry is one sample of weight 1.

    The result is one box per lattice-shift pass (MEEP's ``src_vol`` of each shift, deposited
    in that order), split where the owned window wraps: a request that reaches a periodic
    face deposits the two images that meet on a row separately, rounding twice as MEEP does.
    The point weight is ``w_z * (w_y * w_x)`` (vec.hpp:381-383, exact in any grouping of the
    two-factor products) and the amplitude meets it, and the shift phase 1, as complex
    products. ``integrated`` marks the boxes as MEEP's polarization source (electric only).
    On a mirror-folded axis the request is laid on the full cell and each sample served as
    MEEP's symmetry passes serve it (``_folded_axis_passes``): a sample below the plane
    deposits at its mirror image with the component's parity.
    """
    if integrated and component[0] in "HB":
        raise ValueError("a magnetic integrated source is not supported: the engine keeps H "
                         "apart from B only on the layer rows")
    iyee = _IYEE[component]
    a = float(grid.resolution)
    inva = 1.0 / a
    amp = (float(complex(amplitude).real), float(complex(amplitude).imag))
    folds = tuple(getattr(grid, "folds", (None, None, None)))
    full = tuple(getattr(grid, "shape_full", grid.shape))
    lengths = tuple(full[d] * grid.dx for d in range(3))
    per_axis = []
    for d in range(3):
        n = grid.shape[d]
        if grid.is_invariant(d):
            per_axis.append({0: [(0, numpy.ones(1))]})
            continue
        lo_w = float(center[d]) - float(size[d]) / 2
        hi_w = float(center[d]) + float(size[d]) / 2
        span = hi_w - lo_w
        if span < 0.0:
            raise ValueError("a source size must be non-negative")
        if span > lengths[d] + inva:
            raise ValueError(f"source size {size[d]} exceeds the cell length {lengths[d]} by "
                             f"more than one cell (MEEP's add_volume_source aborts)")
        if span > lengths[d]:  # MEEP snaps a request wider by under a pixel to one cell
            lo_w = lo_w - (span - lengths[d]) * 0.5
            hi_w = lo_w + lengths[d]
        if hi_w - lo_w == 0.0:
            amp = (amp[0] * a, amp[1] * a)
        if folds[d] is not None:
            per_axis.append(_folded_axis_passes(grid, component, d, a, iyee[d], lo_w, hi_w))
        else:
            per_axis.append(_axis_passes(n, a, iyee[d], lo_w, hi_w, not metallic[d]))
    pass_amp = _cmul(amp, (1.0, -0.0))
    boxes: List[DepositBox] = []
    folded = [d for d in range(3) if folds[d] is not None]
    # the symmetry passes outermost (their images in chain order); within each, MEEP's shift
    # odometer runs its first direction fastest
    for sn in range(1 << len(folded)):
        bit = {d: sn >> q & 1 for q, d in enumerate(folded)}

        def keys(d):
            if d in bit:
                return [bit[d]] if bit[d] in per_axis[d] else []
            return sorted(per_axis[d])

        for kz in keys(2):
            for ky in keys(1):
                for kx in keys(0):
                    for x0, wx in per_axis[0][kx]:
                        for y0, wy in per_axis[1][ky]:
                            for z0, wz in per_axis[2][kz]:
                                w = (wx[:, None, None] * wy[None, :, None]) * wz[None, None, :]
                                re = w * pass_amp[0]
                                im = w * pass_amp[1]
                                # times the amplitude function's value, one
                                boxes.append(DepositBox(component, (x0, y0, z0), w.shape,
                                                        re * 1.0 - im * 0.0, re * 0.0 + im * 1.0,
                                                        integrated))
    return boxes


def polarization_boxes(grid, box: DepositBox, metallic: Sequence[bool]) -> List[DepositBox]:
    """An electric polarization (integrated) box as the engine's update_E subtracts it, box by
    box at each box's own cells.

    MEEP writes ``f_minus_p`` at the box's cells and then runs the D boundary passes again
    (reference sources.py:1697-1747, ``_boundary_passes``): a ghost row of a folded axis takes
    its image row's ``f_minus_p`` times the component's parity, and a wall row of a metallic
    axis is cleared. So the box is returned without its ghost and wall rows, followed by one
    image box per set of ghost planes its cells are images of: on those planes' ghost rows, the
    amplitudes times the product of the planes' parities. A parity is a sign, so at a ghost the
    engine's ``float(double(parity * D) - parity * r)`` is exactly ``parity`` times the image's
    word. Ghost rows (the engine's fold fill): row 0 for a component whose lattice puts a
    sample on the plane, the far row of a periodic fold for one whose lattice does not."""
    if not box.integrated or box.magnetic:
        return [box]
    comp = "xyz".index(box.component[1])
    n = grid.shape
    re = numpy.asarray(box.amp_re, dtype=numpy.float64).reshape(box.cnt)
    im = numpy.asarray(box.amp_im, dtype=numpy.float64).reshape(box.cnt)
    lo = list(box.lo)
    hi = [l + c for l, c in zip(box.lo, box.cnt)]
    ghosts = {}  # axis -> (ghost row, image row, parity)
    for a in range(3):
        fold = grid.folds[a]
        on_plane = a != comp  # an electric component's integer lattice on this axis
        cut = []
        if fold is not None:
            if on_plane:
                ghosts[a] = (0, 2, _fold_parity(grid, box.component, a))
            elif fold[1] == 2:
                ghosts[a] = (n[a] - 1, fold[2], _fold_parity(grid, box.component, a))
            if a in ghosts:
                cut.append(ghosts[a][0])
        elif metallic[a] and on_plane:
            cut.append(0)
        for row in cut:
            if lo[a] < hi[a] and row == lo[a]:
                lo[a] += 1
            elif lo[a] < hi[a] and row == hi[a] - 1:
                hi[a] -= 1

    def piece(rows, sign: float, at):
        sl = tuple(slice(r0 - box.lo[a], r1 - box.lo[a]) for a, (r0, r1) in enumerate(rows))
        return DepositBox(box.component, tuple(at), tuple(r1 - r0 for r0, r1 in rows),
                          re[sl] * sign, im[sl] * sign, True)

    out = []
    if all(l < h for l, h in zip(lo, hi)):
        out.append(piece(list(zip(lo, hi)), 1.0, lo))
    axes = sorted(ghosts)
    for mask in range(1, 1 << len(axes)):
        rows, at, sign = [], [], 1.0
        for a in range(3):
            if a in axes and mask >> axes.index(a) & 1:
                ghost, image, parity = ghosts[a]
                if not lo[a] <= image < hi[a]:
                    break
                rows.append((image, image + 1))
                at.append(ghost)
                sign *= parity
            else:
                if lo[a] >= hi[a]:
                    break
                rows.append((lo[a], hi[a]))
                at.append(lo[a])
        else:
            out.append(piece(rows, sign, at))
    return out
