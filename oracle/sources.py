"""MEEP's source placement on one chunk, restated for the oracle (test infrastructure only).

Where a ``mp.Source`` of a given centre and size deposits, and with which amplitude per
point, for a real-storage Cartesian run on one chunk and no Bloch phase. Restated from the
reference's ``meep_gpu/sources.py``:

* ``_axis_bounds``  <- ``_source_geometry`` :860-885 (``vec2diel_floor/ceil`` :647-655 on the
  component's lattice, ``compute_boundary_weights`` :657-695: MEEP loop_in_chunks.cpp:233-297)
* ``_owned``        <- ``_source_geometry`` :886-918 (the owned window ``little_corner + 2 -
  iyee .. big_corner - iyee``, vec.hpp:1102-1107; an axis starts at ``-(n - n % 2)``
  doubled, grid.py:1095-1120)
* ``_axis_weights`` <- ``_src_vol_pieces`` :1097-1121 (the clip to the chunk and the minima
  of a short run) and ``_ivec_loop_weight_1d`` :698-718 (MEEP vec.hpp:372-378)
* ``deposit_points`` <- ``_source_geometry`` :844-1025 and ``_src_vol_pieces`` :1122-1140:
  ``amp *= a`` per zero-size resolved direction (sources.cpp:481-484), the lattice-shift
  odometer over periodic axes (loop_in_chunks.cpp:391-418), the weight product
  ``w[2] * (w[1] * (1.0 * w[0]))`` over the ``yucky`` loop order (vec.hpp:381-383; 2-D
  loops (Z, X, Y), 3-D (X, Y, Z)) and the complex products with the shift phase (1 here)
  and the amplitude function (``one``).

The restatement works in MEEP's doubled integer coordinates: sample ``i`` of a component
whose Yee shift along an axis is ``s`` sits at ``io + 2 i + s`` with ``io = -(n - n % 2)``,
i.e. at ``(io/2 + i + s/2) * dx`` from the cell centre.
"""

from __future__ import annotations

import math
from typing import List, Sequence, Tuple

import numpy

#: Yee shifts in half cells per component (MEEP's iyee_shift)
IYEE = {"Ex": (1, 0, 0), "Ey": (0, 1, 0), "Ez": (0, 0, 1),
        "Hx": (0, 1, 1), "Hy": (1, 0, 1), "Hz": (1, 1, 0)}


def _cmul(p: Tuple[float, float], q: Tuple[float, float]) -> Tuple[float, float]:
    """C++ ``complex<double>`` multiply, every product rounded: (ac - bd, ad + bc)."""
    return (p[0] * q[0] - p[1] * q[1], p[0] * q[1] + p[1] * q[0])


def _axis_bounds(where_min: float, where_max: float, iyee: int, a: float, inva: float):
    """(is, ie, (s0, s1, e0, e1)) of one resolved direction on the component's lattice."""
    iyee_c = 1 - iyee  # the Centered shift (1) minus the component's
    yee_c = iyee_c * 0.5 * inva
    is_d = 1 + 2 * int(math.floor((where_min + yee_c) * a - 0.5)) - iyee_c
    ie_d = 1 + 2 * int(math.ceil((where_max + yee_c) * a - 0.5)) - iyee_c
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
    elif where_min == where_max:  # zero extent: linear interpolation (before span == 2)
        s0, s1, e0, e1 = w0, w1, w1, w0
    elif span == 2:
        s0 = w0 * w0 / 2 - (1 - w1) * (1 - w1) / 2
        e0 = w1 * w1 / 2 - (1 - w0) * (1 - w0) / 2
        s1, e1 = e0, s0
    else:
        s0 = s1 = e0 = e1 = 1.0
    return is_d, ie_d, (s0, s1, e0, e1)


def _owned(n: int, iyee: int) -> Tuple[int, int, int]:
    """(io, lowest owned doubled coordinate, highest) of one axis of a one-chunk grid."""
    io = -(n - n % 2)
    return io, io + 2 - iyee, io + 2 * n - iyee


def _ladder(index: int, count: int, s0: float, s1: float, e0: float, e1: float) -> float:
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


def _axis_weights(start: int, end: int, is_d: int, ie_d: int, ends):
    """The four end weights of a clipped run [start, end] (doubled, unshifted frame)."""
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
        w = min(ws0, we1)
        ws0 = we1 = w
        w = min(ws1, we0)
        ws1 = we0 = w
    elif end == start + 4:
        ws1 = we1 = min(ws1, we1)
    return ws0, ws1, we0, we1


def deposit_points(shape: Sequence[int], invariant: Sequence[bool], resolution: float,
                   component: str, center: Sequence[float], size: Sequence[float],
                   amplitude: complex = 1.0, wraps: Sequence[bool] = (True, True, True)
                   ) -> Tuple[numpy.ndarray, numpy.ndarray, numpy.ndarray]:
    """(C-order flat cell indices, amp_re, amp_im) of every point the source deposits into,
    in MEEP's placement order (shift odometer outermost, then x, y, z)."""
    a = float(resolution)
    inva = 1.0 / a
    directions = tuple(d for d in range(3) if not invariant[d])
    two_d = len(directions) == 2
    yucky = (None, 0, 1) if two_d else (0, 1, 2)
    iyee = IYEE[component]
    where_min = [float(center[d]) - float(size[d]) / 2 for d in range(3)]
    where_max = [float(center[d]) + float(size[d]) / 2 for d in range(3)]
    amp = (float(complex(amplitude).real), float(complex(amplitude).imag))
    for d in directions:
        span = where_max[d] - where_min[d]
        if span > shape[d] * inva:
            raise ValueError("the oracle places sources no wider than the cell")
        if span == 0.0:
            amp = (amp[0] * a, amp[1] * a)
    bounds = {}
    for d in directions:
        bounds[d] = _axis_bounds(where_min[d], where_max[d], iyee[d], a, inva)
    owned = {d: _owned(int(shape[d]), iyee[d]) for d in directions}
    shifts = {}
    for d in directions:
        if not wraps[d]:
            shifts[d] = (0,)
            continue
        io, _, _ = owned[d]
        big = io + 2 * int(shape[d])
        lattice = 2 * int(shape[d]) * 0.5 * inva
        lo = int(math.floor((where_min[d] - big * 0.5 * inva) / lattice))
        hi = int(math.ceil((where_max[d] - io * 0.5 * inva) / lattice))
        shifts[d] = tuple(range(lo, hi + 1))
    cells: List[int] = []
    re: List[float] = []
    im: List[float] = []

    def each_shift():
        index = [0] * len(directions)
        while True:
            yield {d: shifts[d][index[q]] for q, d in enumerate(directions)}
            for q, d in enumerate(directions):  # first direction fastest
                if index[q] + 1 < len(shifts[d]):
                    index[q] += 1
                    break
                index[q] = 0
            else:
                return

    for ishift in each_shift():
        pass_amp = _cmul(amp, (1.0, -0.0))  # amp * conj(phase), phase 1 without Bloch
        runs = {}
        for d in directions:
            is_d, ie_d, ends = bounds[d]
            shifti = 2 * int(shape[d]) * ishift[d]
            io, low, high = owned[d]
            start = max(is_d - shifti, low)
            end = min(ie_d - shifti, high)
            if start > end:
                runs = None
                break
            w = _axis_weights(start + shifti, end + shifti, is_d, ie_d, ends)
            runs[d] = (start, end, w, io)
        if runs is None:
            continue
        axes = [range(runs[d][0], runs[d][1] + 1, 2) if d in runs else (None,) for d in range(3)]
        for px in axes[0]:
            for py in axes[1]:
                for pz in axes[2]:
                    coords = (px, py, pz)
                    factor = {}
                    cell = [0, 0, 0]
                    inert = False
                    for d in directions:
                        start, end, w, io = runs[d]
                        count = (end - start) // 2 + 1
                        factor[d] = _ladder((coords[d] - start) // 2, count, *w)
                        row = (coords[d] - io) // 2
                        if wraps[d]:
                            row %= int(shape[d])
                        elif row >= int(shape[d]):
                            inert = True  # the plane on a metallic high wall: zeroed unread
                        cell[d] = row
                    if inert:
                        continue
                    f = [1.0 if d is None else factor[d] for d in yucky]
                    weight = f[2] * (f[1] * (1.0 * f[0]))
                    scaled = (weight * pass_amp[0], weight * pass_amp[1])
                    product = _cmul(scaled, (1.0, 0.0))  # the amplitude function: one
                    cells.append((cell[0] * int(shape[1]) + cell[1]) * int(shape[2]) + cell[2])
                    re.append(product[0])
                    im.append(product[1])
    return (numpy.asarray(cells, dtype=numpy.int64), numpy.asarray(re, dtype=numpy.float64),
            numpy.asarray(im, dtype=numpy.float64))
