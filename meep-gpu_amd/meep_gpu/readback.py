"""MEEP's ``sim.get_array(component=c)``, word for word: over the whole cell, or a box.

What a user of MEEP (and of the reference's ``GpuRunResult.get_array``) reads back is not the
stored Yee words but MEEP's centred lattice: at every cell centre the sum of the component's
two or four nearest words, ``0.25 * (((a + b) + c) + d)`` in double, times the point's phase,
rounded to float32 once (array_slice.cpp). Restated from the reference's readback.py for
Cartesian grids: real or complex storage, mirror folds and Bloch phases.

* The words MEEP's chunk holds (``chunk_words``, reference :416-453): the stored words; on a
  folded axis the low ghost of a Yee-shift-0 component is its mirror image, stored row 2 times
  the component's parity; a metallic wall's own plane is +0; on a Bloch-phased axis a
  shift-0 component's stored row 0 is the window top, and the word below the window is that
  top word connected by conj(phase) (one plane more, :375-402).
* The next word up along each axis where the component's Yee shift is 0 (``_up`` :455-488):
  the wrap on a periodic axis, +0 past a metallic wall or a folded metallic far face, the
  image of stored row ``N - stored + 2`` about a periodic fold's second mirror.
* The B reads under a PML: in a chunk that holds H as B's alias (no layer on the component's
  own axis) a word the reading chunk does not own is the owner's H word, and a chunk with H of
  its own reads +0 there after exactly one step (the rule the reference states as
  ``HAliasRule.masks`` :491-546; here each division answers it, ``reader_masks``).
* The whole-cell lattice (``CentredLattice`` :751-819): a folded axis unfolded about its
  plane (stored centre -1 unphased, the rest mirrored), then the periodic duplicate plane below
  the first on an unfolded non-metallic axis and on a periodic fold of odd count; each point's
  phase is MEEP's ``shift_phase * S.phase_shift`` (loop_in_chunks.cpp:202, :412-417,
  vec.cpp:1251-1369) in ``complex<double>`` with its signed zeros.
* The volume (``total_volume`` :829-853, ``get_array`` :856-972): the lattice points inside
  it, MEEP's interpolation weights and float ``+=`` collapse on a one-cell axis, the ``n > 1``
  axis filter.

The integer powers of a Bloch factor are the C++ runtime's ``pow(complex<double>, int)``, which
differs between libstdc++ (repeated squaring, and libgcc's ``__divdc3`` for a negative
exponent) and libc++ (``exp(n * log(x))``); ``runtime`` names the one MEEP ran with.
"""

from __future__ import annotations

import itertools
import math
import sys
from dataclasses import dataclass
from typing import Callable, List, Optional, Sequence, Tuple

import numpy

from .chunks import IYEE

_F32 = numpy.float32
_F64 = numpy.float64
B_COMPONENTS = ("Bx", "By", "Bz")
Cplx = Tuple[float, float]
#: a connection within this of +1 is a copy, of -1 a negation (boundaries.cpp:362-366)
CONNECT_TOLERANCE = 1e-13


def host_runtime() -> str:
    """The C++ runtime a MEEP built on this host would use: libc++ on Darwin, else libstdc++."""
    return "libc++" if sys.platform == "darwin" else "libstdc++"


def _slab(axis: int, index):
    out = [slice(None)] * 3
    out[axis] = index
    return tuple(out)


# --- complex<double> arithmetic as MEEP's compiled C++ forms it -----------------------------
def cmul(a: Cplx, b: Cplx) -> Cplx:
    """``(ac - bd, ad + bc)``, one rounding per operation (no contraction), finite operands."""
    return (a[0] * b[0] - a[1] * b[1], a[0] * b[1] + a[1] * b[0])


_TINY = 2.2250738585072014e-308      # DBL_MIN
_EPS = 2.220446049250313e-16         # DBL_EPSILON
_HUGE_HALF = 1.7976931348623157e308 / 2


def _divide(a: float, b: float, c: float, d: float) -> Cplx:
    """(a + ib) / (c + id) by libgcc's ``__divdc3`` (libgcc2.c, GCC >= 12): Smith's method on
    the larger denominator part, after scaling all four operands by a power of two when the
    divisor is huge or tiny (or a numerator part is subnormal), with the alternate order for a
    subnormal ratio. Finite operands."""
    swap = abs(c) < abs(d)
    big, small = (d, c) if swap else (c, d)
    if abs(big) >= _HUGE_HALF:
        a, b, c, d = a * 0.5, b * 0.5, c * 0.5, d * 0.5
    elif abs(big) < _EPS or ((abs(a) < _TINY or abs(b) < _TINY)
                             and max(abs(a), abs(b), abs(big)) < _HUGE_HALF * _EPS):
        s = 1.0 / _EPS
        a, b, c, d = a * s, b * s, c * s, d * s
    if swap:
        r = c / d
        den = (c * r) + d
        if abs(r) > _TINY:
            return (((a * r) + b) / den, ((b * r) - a) / den)
        return (((c * (a / d)) + b) / den, ((c * (b / d)) - a) / den)
    r = d / c
    den = (d * r) + c
    if abs(r) > _TINY:
        return (((b * r) + a) / den, (b - (a * r)) / den)
    return ((a + (d * (b / c))) / den, (b - (d * (a / c))) / den)


def _power_unsigned(x: Cplx, n: int) -> Cplx:
    """libstdc++'s repeated squaring (``__complex_pow_unsigned``); ``(1, 0)`` for n = 0."""
    acc = x if n & 1 else (1.0, 0.0)
    n >>= 1
    while n:
        x = cmul(x, x)
        if n & 1:
            acc = cmul(acc, x)
        n >>= 1
    return acc


def cpow_int(x: Cplx, n: int, runtime: str) -> Cplx:
    """``pow(complex<double> x, int n)`` of the named C++ runtime (see the module docstring)."""
    if runtime == "libstdc++":
        if n < 0:
            p = _power_unsigned(x, -n)
            return _divide(1.0, 0.0, p[0], p[1])
        return _power_unsigned(x, n)
    if runtime != "libc++":
        raise ValueError(f"runtime must be 'libstdc++' or 'libc++', got {runtime!r}")
    if x[0] == 1.0 and x[1] == 0.0 and math.copysign(1.0, x[1]) > 0:
        return (1.0, 0.0)  # log(1, +0) = (0, +0): every power is (1, +0)
    lg = (math.log(math.hypot(x[0], x[1])), math.atan2(x[1], x[0]))
    y = cmul((float(n), 0.0), lg)
    if y[1] == 0:
        return (math.exp(y[0]), math.copysign(0.0, y[1]))
    e = math.exp(y[0])
    return (e * math.cos(y[1]), e * math.sin(y[1]))


# --- the frame of a read ---------------------------------------------------------------------
@dataclass(frozen=True)
class ReadFrame:
    """What the read-out needs beyond the stored words: per axis invariance, an unfolded
    metallic wall, a fold ``(declared phase, kind 1 metallic / 2 periodic, far row)``, the
    full count, the Bloch factor (or None) and the doubled coordinate of stored row 0."""
    invariant: Tuple[bool, bool, bool]
    metallic: Tuple[bool, bool, bool]
    folds: Tuple
    shape_full: Tuple[int, int, int]
    phases: Tuple[Optional[complex], ...]
    origin: Tuple[int, int, int]
    runtime: str

    @classmethod
    def of(cls, grid, wraps: Sequence[bool], phases=None, runtime: Optional[str] = None):
        invariant = tuple(grid.is_invariant(a) for a in range(3))
        folds = tuple(getattr(grid, "folds", (None, None, None)))
        metallic = tuple(not invariant[a] and folds[a] is None and not wraps[a]
                         for a in range(3))
        origin = tuple(0 if invariant[a] else (-2 if folds[a] is not None
                                               else -(grid.shape[a] - grid.shape[a] % 2))
                       for a in range(3))
        phases = tuple(phases) if phases is not None else (None, None, None)
        return cls(invariant, metallic, folds,
                   tuple(getattr(grid, "shape_full", grid.shape)), phases, origin,
                   runtime or host_runtime())

    def parity(self, component: str, axis: int) -> int:
        """The component's sign under the fold of ``axis``: a vector (E, D) is odd about the
        plane its own axis is normal to, a pseudovector (H, B) even there and odd about the
        others, each times the plane's declared phase (MEEP's phase_shift)."""
        own = "xyz".index(component[1]) == axis
        flip = own != (component[0] in "HB")
        return -self.folds[axis][0] if flip else self.folds[axis][0]

    def eikna(self, axis: int) -> Cplx:
        z = self.phases[axis]
        return (1.0, 0.0) if z is None else (float(complex(z).real), float(complex(z).imag))

    def window_axes(self, component: str) -> Tuple[int, ...]:
        """Axes where stored row 0 of this component holds MEEP's owned top word: a Bloch
        phase and Yee shift 0 there."""
        return tuple(a for a in range(3) if self.phases[a] is not None
                     and not self.invariant[a] and IYEE[component][a] == 0)

    def metal_top(self, axis: int) -> bool:
        """Is the word one past the top stored row a zero_metal word?"""
        fold = self.folds[axis]
        return self.metallic[axis] or (fold is not None and fold[1] == 1)

    @property
    def directions(self) -> Tuple[int, ...]:
        return tuple(a for a in range(3) if not self.invariant[a])


def connect(parts: Sequence[numpy.ndarray], phase: Cplx) -> List[numpy.ndarray]:
    """One not-owned word from its owned source (step.cpp:171-221): within the tolerance of +1
    a copy, of -1 a negation of each part, otherwise the ``complex<float>`` product with the
    phase rounded to float, each product and sum rounded on its own."""
    re, im = phase
    if abs(im) < CONNECT_TOLERANCE and abs(re - 1.0) < CONNECT_TOLERANCE:
        return [numpy.array(p, dtype=_F32, copy=True) for p in parts]
    if abs(im) < CONNECT_TOLERANCE and abs(re + 1.0) < CONNECT_TOLERANCE:
        return [numpy.negative(numpy.asarray(p, dtype=_F32)) for p in parts]
    if len(parts) != 2:
        raise ValueError("a Bloch product needs complex storage")
    pr, pi = _F32(re), _F32(im)
    a, b = (numpy.asarray(p, dtype=_F32) for p in parts)
    return [(a * pr) - (b * pi), (a * pi) + (b * pr)]


def _ghost_phase(frame: ReadFrame, below: Sequence[int]) -> Cplx:
    """The phase of the word below the window on the axes ``below``: conj(eikna) per axis, in
    direction order, from 1 (boundaries.cpp:203-226)."""
    ph = (1.0, 0.0)
    for axis in frame.directions:
        if axis in below:
            e = frame.eikna(axis)
            ph = cmul(ph, (e[0], -e[1]))
    return ph


def chunk_words(parts: Sequence[numpy.ndarray], component: str,
                frame: ReadFrame) -> List[numpy.ndarray]:
    """The words MEEP's chunk holds at every stored cell (one plane more on a window axis)."""
    parts = [numpy.array(p, dtype=_F32, copy=True) for p in parts]
    shift = IYEE[component]
    for axis in range(3):  # mirror ghosts, x y z order (a corner takes both signs)
        if frame.folds[axis] is None or shift[axis] != 0:
            continue
        image = connect([p[_slab(axis, 2)] for p in parts],
                        (float(frame.parity(component, axis)), 0.0))
        for p, w in zip(parts, image):
            p[_slab(axis, 0)] = w
    for axis in range(3):  # a wall's own plane, after the ghosts
        if frame.metallic[axis] and shift[axis] == 0:
            for p in parts:
                p[_slab(axis, 0)] = 0.0
    window = frame.window_axes(component)
    if not window:
        return parts
    ext = parts
    for axis in window:  # the top word moves to index N, index 0 becomes the word below
        ext = [numpy.concatenate((e, e[_slab(axis, slice(0, 1))]), axis=axis) for e in ext]
    out = [numpy.array(e, copy=True) for e in ext]
    for size in range(1, len(window) + 1):
        for below in itertools.combinations(window, size):
            dst, src = [slice(None)] * 3, [slice(None)] * 3
            for axis in window:
                dst[axis], src[axis] = (0, -1) if axis in below else (slice(1, None),
                                                                      slice(1, None))
            words = connect([e[tuple(src)] for e in ext], _ghost_phase(frame, below))
            for o, w in zip(out, words):
                o[tuple(dst)] = w
    return out


def _up(parts: Sequence[numpy.ndarray], component: str, axis: int,
        frame: ReadFrame) -> List[numpy.ndarray]:
    """The next word up along ``axis`` at every cell, as MEEP's chunk holds it."""
    out = []
    for p in parts:
        q = numpy.empty_like(p)
        q[_slab(axis, slice(None, -1))] = p[_slab(axis, slice(1, None))]
        out.append(q)
    top = _slab(axis, -1)
    if axis in frame.window_axes(component):  # the extended plane is read; this one is cropped
        for q, p in zip(out, parts):
            q[top] = p[top]
        return out
    if frame.metal_top(axis):
        for q in out:
            q[top] = 0.0
        return out
    fold = frame.folds[axis]
    if fold is not None:  # the image about a periodic fold's second mirror
        image = connect([p[_slab(axis, fold[2])] for p in parts],
                        (float(frame.parity(component, axis)), 0.0))
    else:
        image = [numpy.array(p[_slab(axis, 0)], copy=True) for p in parts]
    for q, w in zip(out, image):
        q[top] = w
    return out


def _words(parts, component: str, axes: Sequence[int], frame: ReadFrame):
    if not axes:
        return [parts]
    if len(axes) == 1:
        return [parts, _up(parts, component, axes[0], frame)]
    n1 = _up(parts, component, axes[0], frame)
    n2 = _up(parts, component, axes[1], frame)
    n12 = _up(n1, component, axes[1], frame)
    if frame.metal_top(axes[0]):  # the corner past a metal face is a zero_metal word
        for q in n12:
            q[_slab(axes[0], -1)] = 0.0
    return [parts, n1, n2, n12]


def _as_parts(array: numpy.ndarray) -> List[numpy.ndarray]:
    a = numpy.asarray(array)
    if a.dtype.kind == "c":
        return [numpy.ascontiguousarray(a.real, dtype=_F32),
                numpy.ascontiguousarray(a.imag, dtype=_F32)]
    return [numpy.asarray(a, dtype=_F32)]


def centred_sums(read: Callable[[str], numpy.ndarray], component: str, invariant, metallic,
                 division, pml_active: bool, first_step: bool,
                 frame: Optional[ReadFrame] = None) -> List[numpy.ndarray]:
    """``0.25 * sum`` in float64 per stored part at every stored cell centre (doubled
    ``origin + 2 i + 1``). Without a ``frame``: unfolded axes, no Bloch phase."""
    if frame is None:
        frame = ReadFrame(tuple(invariant), tuple(metallic), (None, None, None), (0, 0, 0),
                          (None, None, None), (0, 0, 0), host_runtime())
    stored = _as_parts(read(component))
    shape = stored[0].shape
    axes = [a for a in range(3) if not frame.invariant[a] and IYEE[component][a] == 0]
    crop = tuple(slice(0, n) for n in shape)
    words = [[w[crop] for w in word]
             for word in _words(chunk_words(stored, component, frame), component, axes, frame)]
    if axes and component in B_COMPONENTS and pml_active:
        h = _as_parts(read("H" + component[1]))
        h_words = [[w[crop] for w in word]
                   for word in _words(chunk_words(h, component, frame), component, axes, frame)]
        take_h, take_zero = division.reader_masks(component, axes, shape, first_step)
        words = [[numpy.where(zero, _F32(0.0), numpy.where(take, wh, wb)).astype(_F32)
                  for wb, wh in zip(word_b, word_h)]
                 for word_b, word_h, take, zero in zip(words, h_words, take_h, take_zero)]
    if len(words) == 1:
        sums = [((f + f) + f) + f for f in words[0]]
    elif len(words) == 2:
        sums = [((f + a) + f) + a for f, a in zip(*words)]
    else:
        sums = [((f + b) + c) + d for f, b, c, d in zip(*words)]
    return [s.astype(_F64) * 0.25 for s in sums]


# --- MEEP's symmetry and lattice phases -------------------------------------------------------
def phase_shift(frame: ReadFrame, component: str, reflected: int) -> Cplx:
    """``symmetry::phase_shift`` of the image reflected on the folded axes in bit mask
    ``reflected`` (vec.cpp:1251-1369): the product of those planes' phases, each ``1 *
    declared``, nested from the last plane of the chain outwards; negated (both parts) when the
    component's own direction is flipped, and once more per flipped direction for a
    pseudovector."""
    planes = [a for a in range(3) if frame.folds[a] is not None]  # the chain, x y z order
    hit = [a for a in planes if reflected >> a & 1]
    if not hit:
        return (1.0, 0.0)
    ph = None
    for axis in reversed(hit):
        own = cmul((1.0, 0.0), (float(frame.folds[axis][0]), 0.0))
        ph = cmul((1.0, 0.0), own) if ph is None else cmul(ph, own)
    flip = "xyz".index(component[1]) in hit
    if component[0] in "HB":
        flip = flip != (len(hit) % 2 == 1)
    return (-ph[0], -ph[1]) if flip else ph


def shift_phase(frame: ReadFrame, ishift: Sequence[int]) -> Cplx:
    """``ph = 1; ph *= pow(eikna[d], ishift[d])`` over the directions (loop_in_chunks.cpp:412)."""
    ph = (1.0, 0.0)
    for axis in frame.directions:
        ph = cmul(ph, cpow_int(frame.eikna(axis), int(ishift[axis]), frame.runtime))
    return ph


def _unfold(a: numpy.ndarray, axis: int, n_full: int, centre: int) -> numpy.ndarray:
    """A folded axis's centred points over the full cell: stored centre i at doubled 2i - 1,
    so index ``centre - 1`` is stored 0 (the owned pixel at -1), ``centre + k`` stored
    ``1 + k``, and below ``centre - 1`` the mirror images of stored 2, 3, ..."""
    shape = list(a.shape)
    shape[axis] = n_full
    out = numpy.zeros(tuple(shape), dtype=a.dtype)
    out[_slab(axis, slice(centre, None))] = a[_slab(axis, slice(1, 1 + n_full - centre))]
    out[_slab(axis, centre - 1)] = a[_slab(axis, 0)]
    out[_slab(axis, slice(None, centre - 1))] = a[_slab(axis, slice(centre, 1, -1))]
    return out


def _vec2diel_floor(coordinate: float, a: float) -> int:  # loop_in_chunks.cpp:229-236
    return 1 + 2 * int(math.floor(coordinate * a - .5))


def _vec2diel_ceil(coordinate: float, a: float) -> int:  # loop_in_chunks.cpp:238-245
    return 1 + 2 * int(math.ceil(coordinate * a - .5))


def meep_array(read: Callable[[str], numpy.ndarray], component: str, grid,
               wraps: Sequence[bool], division, pml_active: bool, step_count: int,
               center=None, size=None, phases=None, runtime: Optional[str] = None
               ) -> numpy.ndarray:
    """MEEP's ``get_array(component)`` (float32, or complex64 from complex storage; MEEP's
    rank and axis order): over the whole cell (``fields::total_volume()``), or over the box
    ``center -/+ size/2`` as ``mp.Volume`` forms it. ``read(name)`` returns a stored volume, C
    order (nx, ny, nz), float32 or complex64; ``phases`` per axis the Bloch factor or None;
    ``runtime`` the C++ runtime of MEEP's integer powers (the host's by default)."""
    if (center is None) != (size is None):
        raise ValueError("pass both center and size, or neither (the whole cell)")
    frame = ReadFrame.of(grid, wraps, phases, runtime)
    q = centred_sums(read, component, frame.invariant, frame.metallic, division, pml_active,
                     step_count == 1, frame)
    complex_storage = len(q) == 2
    stored_shape = q[0].shape
    reflected = numpy.zeros(stored_shape, dtype=numpy.uint8)
    shifted = numpy.zeros(stored_shape, dtype=numpy.uint8)
    order = list(frame.directions)
    ivec = {axis: frame.origin[axis] + 2 * numpy.arange(stored_shape[axis]) + 1 for axis in order}
    for axis in order:  # unfold every folded axis about its plane
        if frame.folds[axis] is None:
            continue
        n_full = frame.shape_full[axis]
        centre = n_full // 2
        q = [_unfold(p, axis, n_full, centre) for p in q]
        reflected = _unfold(reflected, axis, n_full, centre)
        shifted = _unfold(shifted, axis, n_full, centre)
        reflected[_slab(axis, slice(None, centre - 1))] |= numpy.uint8(1 << axis)
        ivec[axis] = 2 * numpy.arange(n_full) - 2 * centre + 1
    for axis in order:  # the periodic duplicate plane one lattice vector below the first
        fold = frame.folds[axis]
        if fold is not None:
            duplicate = fold[1] == 2 and frame.shape_full[axis] % 2 == 1
        else:
            duplicate = not frame.metallic[axis]
        if not duplicate:
            continue
        top = _slab(axis, slice(-1, None))
        q = [numpy.concatenate((p[top], p), axis=axis) for p in q]
        reflected = numpy.concatenate((reflected[top], reflected), axis=axis)
        shifted = numpy.concatenate((shifted[top] | numpy.uint8(1 << axis), shifted), axis=axis)
        ivec[axis] = numpy.concatenate(([ivec[axis][0] - 2], ivec[axis]))
    a = float(grid.resolution)
    pick = [numpy.array([0])] * 3
    weights = {}
    collapse = []
    for axis in order:
        if center is None:  # total_volume: [origin, origin + 2 n - 2], a fold's mirrored
            low = frame.origin[axis]
            high = low + 2 * stored_shape[axis] - 2
            if frame.folds[axis] is not None:
                reach = max(abs(low), abs(high))
                low, high = -reach, reach
            low, high = low * 0.5 * grid.dx, high * 0.5 * grid.dx
        else:
            low = float(center[axis]) - float(size[axis]) * 0.5
            high = float(center[axis]) + float(size[axis]) * 0.5
        lo, hi = min(low, high), max(low, high)
        i_s, i_e = _vec2diel_floor(lo, a), _vec2diel_ceil(hi, a)
        keep = numpy.nonzero((ivec[axis] >= i_s) & (ivec[axis] <= i_e))[0]
        if keep.size == 0:
            raise ValueError(f"the volume misses the {'xyz'[axis]} centred lattice of {component}")
        pick[axis] = keep
        if lo != hi:
            continue
        w0 = 1. - lo * a + 0.5 * i_s  # loop_in_chunks.cpp:260-261, left to right
        w1 = 1. + hi * a - 0.5 * i_e
        if keep.size == 2:
            weights[axis] = numpy.array([min(w0, w0), min(w1, w1)], dtype=_F64)
            collapse.append(axis)
        elif keep.size == 1:
            visited = int(ivec[axis][keep[0]])
            w = min(w0, w1) if i_e == i_s else (w0 if visited == i_s else w1)
            weights[axis] = numpy.array([w], dtype=_F64)
        else:
            raise RuntimeError("an empty dimension spanning more than two centred planes")
    sel = numpy.ix_(*pick)
    values = [p[sel] for p in q]
    ref, sh = reflected[sel], shifted[sel]
    pr = numpy.empty(ref.shape, dtype=_F64)
    pi = numpy.empty(ref.shape, dtype=_F64)
    key = ref.astype(numpy.uint16) | (sh.astype(numpy.uint16) << 8)
    for value in numpy.unique(key):
        mask, bits = int(value) & 0xFF, int(value) >> 8
        ishift = [-1 if bits >> axis & 1 else 0 for axis in range(3)]
        ph = cmul(shift_phase(frame, ishift), phase_shift(frame, component, mask))
        pr[key == value], pi[key == value] = ph
    weight = None
    for axis in order:  # IVEC_LOOP_WEIGHT: w3 * (w2 * (dV * w1))
        factor = _F64(1.0)
        if axis in weights:
            view = [1, 1, 1]
            view[axis] = weights[axis].size
            factor = weights[axis].reshape(view)
        weight = (_F64(1.0) * factor) if weight is None else (factor * weight)
    if weight is None:
        weight = _F64(1.0)
    wr = weight * values[0]  # weight * complex<double>(f[0], f[1]) * ph, separate operations
    wi = weight * (values[1] if complex_storage else numpy.zeros_like(values[0]))
    store = [((wr * pr) - (wi * pi)).astype(_F32), ((wr * pi) + (wi * pr)).astype(_F32)]
    if collapse:  # collapse_array: reduced[...] += full[...] in float32, row-major
        shape = list(store[0].shape)
        reduced = tuple(1 if axis in collapse else shape[axis] for axis in range(3))
        acc = [numpy.zeros(reduced, dtype=_F32), numpy.zeros(reduced, dtype=_F32)]
        ranges = [range(shape[axis]) if axis in collapse else (None,) for axis in range(3)]
        for combo in itertools.product(*ranges):
            part = tuple(slice(combo[axis], combo[axis] + 1) if axis in collapse else slice(None)
                         for axis in range(3))
            acc = [acc[k] + store[k][part] for k in range(2)]
        store = acc
    drop = tuple(axis for axis in range(3) if axis not in order or store[0].shape[axis] == 1)
    if drop:
        store = [s.squeeze(axis=drop) for s in store]
    if not complex_storage:
        return numpy.array(store[0], dtype=_F32, copy=True)
    out = numpy.empty(store[0].shape, dtype=numpy.complex64)
    out.real, out.imag = store
    return out
