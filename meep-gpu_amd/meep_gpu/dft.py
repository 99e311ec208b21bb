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

``DftFieldsMonitor`` is MEEP's ``add_dft_fields`` (reference ``DFTMonitor``, dft.py:1960): the
transform of up to any number of components over a box of cells, cell-centred (yee2cent) or at
each component's own Yee site, on the same device machinery as the flux planes (at most four
components per engine region); ``get_dft_array`` reduces a zero-thickness axis with MEEP's
interpolation weights as ``sim.get_dft_array`` does (reference from_meep.py:886-955).
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


def boundary_weight_ladder(where_min: float, where_max: float, a: float, iyee_c: int = 0
                           ) -> Tuple[int, int, Tuple[float, float, float, float]]:
    """MEEP's integration ladder for one direction of a volume on the centred grid.

    ``vec2diel_floor`` / ``vec2diel_ceil`` bracket the bounds on the odd doubled lattice (the
    cell centres sit at doubled coordinates 2k+1, i.e. at (k + 1/2)/a from the origin) and
    ``compute_boundary_weights`` picks the four end weights by the same case ladder
    (MEEP loop_in_chunks.cpp:233-297 as reference dft.py:604-690 states it, for parity 1).
    ``iyee_c`` is the Centered shift minus a component's Yee shift along the direction (0
    for a monitor's cell centres, 1 for a component on the integer lattice,
    loop_in_chunks.cpp:352-359, reference sources.py:870-885): the bounds move by
    ``iyee_c/2`` cells onto that component's samples and the weights stay the overlaps of
    the original request. ``w0``/``w1`` are the overlaps of the first and last sample with
    the request, so an extended direction integrates to exactly ``(where_max - where_min) *
    a`` cells and a flat one is a linear interpolation (the weights sum to 1).
    Returns (is, ie, (s0, s1, e0, e1)).
    """
    yee_c = iyee_c * 0.5 * (1.0 / a)
    is_d = 1 + 2 * math.floor((where_min + yee_c) * a - 0.5) - iyee_c
    ie_d = 1 + 2 * math.ceil((where_max + yee_c) * a - 0.5) - iyee_c
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


def folded_axis_cells(grid, axis: int, where_min: float, where_max: float, sign: int = 1
                      ) -> Tuple[int, numpy.ndarray]:
    """(first stored row, per-row weights) of one direction of a monitor volume on a
    mirror-folded axis: the ladder of the request over the full cell (its origin at the
    mirror plane), each cell centre served where MEEP's symmetry images serve it
    (``Grid.fold_site``: the site at doubled -1 directly, those below it at their mirror
    images); a row reached through an odd number of reflections adds its weight times
    ``sign`` (the image's flux sign), and two sites meeting on a row add."""
    a = 1.0 / grid.dx
    is_d, ie_d, (s0, s1, e0, e1) = boundary_weight_ladder(where_min, where_max, a)
    count = (ie_d - is_d) // 2 + 1
    rows = {}
    for q in range(count):
        site = grid.fold_site(axis, is_d + 2 * q, 1)
        if site is None:
            raise ValueError(f"monitor extent [{where_min}, {where_max}] leaves the grid")
        w = ladder_weight(q, count, s0, s1, e0, e1) * (sign if site[1] % 2 else 1)
        rows[site[0]] = rows.get(site[0], 0.0) + w
    first, last = min(rows), max(rows)
    out = numpy.zeros(last - first + 1)
    for row, w in rows.items():
        out[row - first] = w
    return first, out


def _folded_centres(grid, axis: int, where_min: float, where_max: float):
    """The ladder's cell centres of a request over a folded axis's full cell: per centre, the
    stored row serving it, its reflections and its weight (``Grid.fold_site``)."""
    a = 1.0 / grid.dx
    is_d, ie_d, ends = boundary_weight_ladder(where_min, where_max, a)
    count = (ie_d - is_d) // 2 + 1
    rows, flips, weights = [], [], []
    for q in range(count):
        site = grid.fold_site(axis, is_d + 2 * q, 1)
        if site is None:
            raise ValueError(f"monitor extent [{where_min}, {where_max}] leaves the grid")
        rows.append(site[0])
        flips.append(site[1])
        weights.append(ladder_weight(q, count, *ends))
    return (numpy.asarray(rows, dtype=numpy.int64), numpy.asarray(flips, dtype=numpy.int64),
            numpy.asarray(weights))


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


_AXES = "xyz"


def tangential_names(normal: int) -> Tuple[Tuple[str, str], Tuple[str, str]]:
    """(E names, H names) of the two tangential components, in cyclic order after the normal."""
    t1, t2 = (normal + 1) % 3, (normal + 2) % 3
    return ("E" + _AXES[t1], "E" + _AXES[t2]), ("H" + _AXES[t1], "H" + _AXES[t2])


@dataclass(frozen=True)
class FluxPlaneDft:
    """One sample plane's accumulated transforms, detached from the run (host copies), as
    reference ``FluxPlaneDft`` (dft.py:2618-2635): a named ``(nf, nx, ny, nz)`` complex block per
    tangential component on each side, and the frequencies they were accumulated at."""
    E: dict
    H: dict
    frequencies: tuple


@dataclass(frozen=True)
class FluxDftData:
    """What ``FluxMonitor.get_dft_data`` returns (MEEP's ``sim.get_flux_data(flux)``), as
    reference ``FluxDftData`` (dft.py:2638-2658). Opaque: the only supported use is handing it
    back to ``load_dft_data`` / ``load_minus_dft_data`` of a monitor built the same way in a
    later run. One record per sample plane of the monitor, in the monitor's own order."""
    planes: tuple


def _load_flux_planes(data, monitor: "FluxMonitor", description: str) -> None:
    """Replace each sample plane's transforms from ``data`` (reference ``_load_flux_planes``,
    dft.py:661-733, after MEEP's ``load_flux_data``: every element is assigned, nothing is
    added). Checked plane by plane -- plane count, frequencies, component sets, block shapes --
    before anything is written, so a refusal leaves the monitor untouched."""
    if not isinstance(data, FluxDftData):
        raise TypeError(f"{description} takes the object a flux monitor's get_dft_data() "
                        f"returned (FluxDftData), got {type(data).__name__}")
    ids = monitor._attached_ids(description)
    if len(data.planes) != len(ids):
        raise ValueError(f"{description} was given data for {len(data.planes)} flux plane(s) "
                         f"but this monitor has {len(ids)}")
    e_names, h_names = tangential_names(monitor.normal)
    blocks = []
    for index, (record, (lo, cnt, _)) in enumerate(zip(data.planes, monitor.plane_boxes())):
        where = f"{description}, plane {index}"
        if len(record.frequencies) != monitor.num_frequencies:
            raise ValueError(f"{where}: the data holds {len(record.frequencies)} frequencies "
                             f"and this monitor accumulates {monitor.num_frequencies}")
        for saved, mine in zip(record.frequencies, monitor.frequencies):
            if not math.isclose(float(saved), float(mine), rel_tol=1e-12, abs_tol=0.0):
                raise ValueError(f"{where}: the data was accumulated at frequencies "
                                 f"{record.frequencies}, this monitor at "
                                 f"{tuple(monitor.frequencies)}")
        shape = (monitor.num_frequencies,) + tuple(cnt)
        stack = []
        for side, names, block in (("E", e_names, record.E), ("H", h_names, record.H)):
            if set(block) != set(names):
                raise ValueError(f"{where}: the data's {side} components are {sorted(block)}, "
                                 f"expected {sorted(names)}")
            for name in names:
                arr = numpy.asarray(block[name])
                if arr.shape != shape:
                    raise ValueError(f"{where}: {name} in the data has shape {arr.shape} and "
                                     f"this monitor's plane is {shape}")
                stack.append(arr)
        blocks.append(numpy.stack(stack))
    for pid, block in zip(ids, blocks):
        monitor._sim.flux_write(pid, block)


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
        axis of ``size`` is zero (the normal); coordinates have their origin at the cell centre.
        On a mirror-folded axis the region is laid on the full cell and folded onto the stored
        rows (``folded_axis_cells``): a mirror image counts with the sign its flux takes, -1
        through a plane normal to the monitor's normal (the image points the other way), +1
        through one along it (E_t1 H_t2 and E_t2 H_t1 both keep their sign there)."""
        flat = [a for a in range(3) if not grid.is_invariant(a) and size[a] == 0]
        if len(flat) != 1:
            raise ValueError("a flux region is flat along exactly one resolved axis")
        normal = flat[0]
        lo, cnt, weights = [0, 0, 0], [1, 1, 1], []
        planes = None
        folds = getattr(grid, "folds", (None, None, None))
        for a in range(3):
            n = grid.shape[a]
            if grid.is_invariant(a):
                weights.append(numpy.ones(1))
                continue
            lo_w = center[a] - 0.5 * size[a]
            hi_w = center[a] + 0.5 * size[a]
            if folds[a] is not None:
                first, w = folded_axis_cells(grid, a, lo_w, hi_w, -1 if a == normal else 1)
            else:
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

    def _attached_ids(self, what: str) -> list:
        if self._sim is None or self.plane_ids is None:
            raise RuntimeError(f"{what}: the monitor is not attached to a running simulation "
                               f"(freeze the driver first)")
        return list(self.plane_ids)

    def get_dft_data(self) -> FluxDftData:
        """Every sample plane's accumulated transforms as host copies, MEEP's
        ``sim.get_flux_data(flux)`` (reference ``FluxMonitor.get_dft_data``, dft.py:3261)."""
        e_names, h_names = tangential_names(self.normal)
        planes = []
        for pid in self._attached_ids("get_dft_data"):
            t = self._sim.flux_read(pid)
            planes.append(FluxPlaneDft(E={e_names[0]: t[0].copy(), e_names[1]: t[1].copy()},
                                       H={h_names[0]: t[2].copy(), h_names[1]: t[3].copy()},
                                       frequencies=tuple(float(f) for f in self.frequencies)))
        return FluxDftData(planes=tuple(planes))

    def load_dft_data(self, data: FluxDftData) -> None:
        """Replace the transforms with ``data``: MEEP's ``load_flux_data`` (reference
        dft.py:3288-3295). The accumulation that follows adds on top."""
        _load_flux_planes(data, self, "FluxMonitor.load_dft_data")

    def scale_dfts(self, scale: complex) -> None:
        """Multiply every transform, E and H sides alike, by ``scale`` (reference
        dft.py:3274-3286, MEEP's dft_flux::scale_dfts)."""
        for pid in self._attached_ids("scale_dfts"):
            self._sim.flux_scale(pid, complex(scale))

    def load_minus_dft_data(self, data: FluxDftData) -> None:
        """``load_dft_data`` then negate: MEEP's ``load_minus_flux_data`` (reference
        dft.py:3297-3316). The run then ends holding E2 - E1 and H2 - H1, and the flux it
        reports is Re[(E2 - E1) x conj(H2 - H1)]: the fields are subtracted, not the powers."""
        self.load_dft_data(data)
        self.scale_dfts(-1.0)

    def get_flux_spectrum(self) -> numpy.ndarray:
        """The flux per frequency (reference ``get_flux_spectrum``, dft.py:3090): each sample
        plane's sum reduced on the device, scaled by the area element and its weight."""
        if self._sim is None:
            raise RuntimeError("the monitor is not attached to a running simulation")
        total = None
        for pid, (_, _, w) in zip(self.plane_ids, self.plane_boxes()):
            part = self._sim.flux_spectrum(pid, self.cell_weights) * self.area_element * (
                self.weight * w)
            total = part if total is None else total + part
        return total

    def get_flux_spectrum_host(self) -> numpy.ndarray:
        """The same spectrum summed on the host over the downloaded transforms."""
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


_FIELD_NAMES = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


@dataclass
class DftFieldsMonitor:
    """The DFT of field components over the cells [lo, lo+cnt) (reference ``DFTMonitor``,
    dft.py:1960-2300): every accumulating step each component is sampled -- at the cell centre
    (MEEP's yee2cent, dft.py:352, the default) or at its own Yee site (``yee_grid``) -- E at t
    and H at t - dt/2, times exp(i omega t) dt*decimation/sqrt(2 pi), summed in float64.

    ``empty_axes`` maps an axis the requested volume is flat along to the interpolation weights
    of the one or two cells registered there (the ladder's, ``boundary_weight_ladder``);
    ``get_dft_array`` sums that axis away with them, and drops invariant axes, as MEEP's
    ``get_dft_array`` does (process_dft_component and collapse_array, reference
    from_meep.py:886-955).

    ``folded`` maps a mirror-folded axis to its requested cell centres over the full cell:
    ``(stored row offsets from lo, reflections, declared phase)``; ``get_dft_array`` reads each
    centre from the row serving it, times the component's parity per reflection (MEEP's
    symmetry images, ``Grid.fold_site``), and a flat folded axis's weights are per centre.
    """
    components: Tuple[str, ...]
    lo: Tuple[int, int, int]
    cnt: Tuple[int, int, int]
    frequencies: numpy.ndarray
    dt: float
    decimation: int = 1
    yee_grid: bool = False
    empty_axes: dict = field(default_factory=dict)
    invariant: Tuple[bool, bool, bool] = (False, False, False)
    folded: dict = field(default_factory=dict)
    _sim: object = field(default=None, repr=False)
    _slots: dict = field(default_factory=dict, repr=False)

    def __post_init__(self):
        self.components = tuple(str(c) for c in self.components)
        if not self.components:
            raise ValueError("a DFT field monitor needs at least one component")
        unknown = [c for c in self.components if c not in _FIELD_NAMES]
        if unknown or len(set(self.components)) != len(self.components):
            raise ValueError(f"DFT field components must be distinct names of {_FIELD_NAMES}, "
                             f"got {self.components}")
        self.frequencies = numpy.atleast_1d(numpy.asarray(self.frequencies, dtype=numpy.float64))
        if self.frequencies.size < 1:
            raise ValueError("a DFT field monitor needs at least one frequency")
        if int(self.decimation) < 1:
            raise ValueError("decimation must be a positive integer")
        self.lo = tuple(int(v) for v in self.lo)
        self.cnt = tuple(int(v) for v in self.cnt)
        for axis, w in self.empty_axes.items():
            sites = len(self.folded[axis][0]) if axis in self.folded else self.cnt[axis]
            if len(w) != sites:
                raise ValueError(f"axis {axis}: {len(w)} weights for {sites} cells")

    @classmethod
    def from_region(cls, grid, components, center, size, frequencies, decimation: int = 1,
                    yee_grid: bool = False) -> "DftFieldsMonitor":
        """The cells MEEP's axis ladder registers for the volume (center, size) -- coordinates
        with their origin at the cell centre -- as for a flux region: an extended axis takes
        the ladder's cells, a flat one the one or two cell-centre planes bracketing it, with
        their interpolation weights (reference ``set_region_from_volume`` /
        ``_register_volume``, dft.py:604-706)."""
        lo, cnt, empty, folded = [0, 0, 0], [1, 1, 1], {}, {}
        folds = getattr(grid, "folds", (None, None, None))
        for a in range(3):
            if grid.is_invariant(a):
                if size[a] != 0:
                    raise ValueError(f"axis {a} is invariant: the region must be flat along it")
                continue
            if folds[a] is not None:
                if yee_grid:
                    raise ValueError("a yee_grid region on a mirror-folded axis is not built: "
                                     "each component's own lattice folds differently")
                rows, flips, w = _folded_centres(grid, a, center[a] - 0.5 * size[a],
                                                 center[a] + 0.5 * size[a])
                if size[a] == 0:
                    keep = [q for q in range(w.size) if w[q] != 0.0] or [0]
                    rows, flips, w = rows[keep], flips[keep], w[keep]
                    empty[a] = tuple(float(x) for x in w)
                lo[a], cnt[a] = int(rows.min()), int(rows.max() - rows.min() + 1)
                folded[a] = (rows - rows.min(), flips, int(folds[a][0]))
                continue
            first, w = axis_cells(grid.shape[a], grid.dx, center[a] - 0.5 * size[a],
                                  center[a] + 0.5 * size[a])
            if size[a] == 0:
                if yee_grid:
                    raise ValueError("a yee_grid region flat along an axis interpolates on the "
                                     "component's own lattice, which this build does not place; "
                                     "give it one cell of thickness instead")
                keep = [q for q in range(w.size) if w[q] != 0.0] or [0]
                first, w = first + keep[0], w[keep[0]:keep[-1] + 1]
                empty[a] = tuple(float(x) for x in w)
            lo[a], cnt[a] = first, w.size
        return cls(tuple(components), tuple(lo), tuple(cnt), frequencies, grid.dt, decimation,
                   yee_grid, empty, tuple(grid.is_invariant(a) for a in range(3)), folded)

    @property
    def omegas(self) -> numpy.ndarray:
        return 2.0 * math.pi * self.frequencies

    @property
    def scale(self) -> float:
        return self.dt * self.decimation / math.sqrt(2.0 * math.pi)

    @property
    def num_frequencies(self) -> int:
        return int(self.frequencies.size)

    def attach(self, sim) -> None:
        """Register the engine regions: the components four at a time."""
        self._sim, self._slots = sim, {}
        for g in range(0, len(self.components), 4):
            group = self.components[g:g + 4]
            rid = sim.add_dft_region(self.lo, self.cnt, group, self.omegas, self.scale,
                                     self.decimation, self.yee_grid)
            for slot, name in enumerate(group):
                self._slots[name] = (rid, slot)

    def get_dft(self, component: str, freq_index: int = 0) -> numpy.ndarray:
        """The component's transform over the box, complex128 (cnt_x, cnt_y, cnt_z), at one
        frequency (reference ``DFTMonitor.get_dft``, dft.py:2259)."""
        if self._sim is None:
            raise RuntimeError("the monitor is not attached to a running simulation")
        if component not in self._slots:
            raise ValueError(f"component {component!r} is not monitored (monitored: "
                             f"{self.components})")
        if not 0 <= int(freq_index) < self.num_frequencies:
            raise IndexError(f"frequency index {freq_index} of {self.num_frequencies}")
        rid, slot = self._slots[component]
        return self._sim.flux_read(rid)[slot, int(freq_index)].copy()

    def get_dft_array(self, component: str, freq_index: int = 0) -> numpy.ndarray:
        """``sim.get_dft_array(dft_fields, component, freq_index)``: the block with every flat
        axis summed away under its interpolation weights and every invariant axis dropped."""
        block = self.get_dft(component, freq_index)
        for axis, (rows, flips, declared) in self.folded.items():  # the full cell's centres
            own = "xyz".index(component[1]) == axis
            parity = -declared if own != (component[0] == "H") else declared
            sign = numpy.where(flips % 2 == 1, float(parity), 1.0)
            view = [1, 1, 1]
            view[axis] = sign.size
            block = block.take(rows, axis=axis) * sign.reshape(view)
        for axis in (2, 1, 0):
            if axis in self.empty_axes:
                w = numpy.asarray(self.empty_axes[axis], dtype=numpy.float64)
                block = numpy.tensordot(block, w, axes=([axis], [0]))
            elif self.invariant[axis]:
                block = block.take(0, axis=axis)
        return block
